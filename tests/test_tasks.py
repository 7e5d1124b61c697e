"""Per-midpoint tasks (the SPITS module's job manager / worker / committer roles,
include/hss_host.h hss_host_job_*) on the CPU: the job hands out one task per CDP of the
line in ascending midpoint order, result records are committed in any order, and the
committed job writes the same files, byte for byte, as the reference's own Dumper wrote
for the same values (the dump fixtures, written by the reference-built checker).  The
worker side needs a GPU (tests/test_gpu_tasks.py)."""
import json
import os
import re
import struct

import numpy as np
import pytest

from highperformanceseismicstacking_amd import tasks

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PROGRAM = {"greedy": tasks.LINEAR_SEARCH, "de": tasks.DIFFERENTIAL_EVOLUTION}


def _command_line(a, out):
    args = ["--aph", a["aph"], "--apm", a["apm"], "--tau", a["tau"], "--azimuth", a["azimuth"],
            "--input", a["gather"], "--output", out, "--traveltime", a["model"],
            "--lower-bounds", *a["lower"], "--upper-bounds", *a["upper"], "--verbose", 0]
    if a["model"] == "zocrs":
        args += ["--v0", a["v0"], "--ain", a["ain"], "--bpctg", a["bpctg"]]
    if a["model"] == "oct":
        args += ["--h0", a["h0"]]
    if a["method"] == "greedy":
        args += ["--granularity", *([4] * len(a["lower"]))]
    else:
        args += ["--generations", 2, "--population-size", 4]
    return args


@pytest.mark.parametrize("case", ["dump_basic_cmp", "dump_basic_zocrs", "dump_basic_oct",
                                  "dump_filters_cmp"])
def test_committed_job_writes_the_reference_dumper_files(case, tmp_path, monkeypatch):
    fix = os.path.join(GOLD, case)
    a = json.load(open(os.path.join(fix, "args.json")))
    monkeypatch.chdir(GOLD)  # gather.txt names the input as given
    job = tasks.TaskJob(PROGRAM[a["method"]], _command_line(a, str(tmp_path)))
    values = np.fromfile(os.path.join(fix, "values.f32"), np.float32)
    stats = np.fromfile(os.path.join(fix, "stats.f32"), np.float32).reshape(-1, 5)
    ncdp = len(stats)
    assert job.tasks == ncdp and job.task_bytes == 4
    per = values.size // ncdp
    assert job.result_bytes == 4 * (1 + per + 5)
    m0s = []
    while (t := job.next_task()) is not None:
        m0s.append(struct.unpack("<f", t)[0])
    assert job.next_task() is None
    assert len(m0s) == ncdp and m0s == sorted(m0s)
    # records in reverse order: the committer keys them by midpoint
    for k in reversed(range(ncdp)):
        rec = (np.float32(m0s[k]).tobytes() + values[k * per:(k + 1) * per].tobytes() +
               stats[k].tobytes())
        job.commit_task(rec)
    out = job.commit()
    job.close()
    name = os.path.basename(out.rstrip("/"))
    assert name.startswith(a["directory"]) and re.fullmatch(r"\d{8}_\d{4}", name[len(a["directory"]):])
    assert sorted(os.listdir(out)) == a["files"]
    for f in a["files"]:
        assert open(os.path.join(out, f), "rb").read() == open(os.path.join(fix, f), "rb").read(), f


def test_job_rejects_bad_records_and_programs(tmp_path, monkeypatch):
    a = json.load(open(os.path.join(GOLD, "dump_basic_cmp", "args.json")))
    monkeypatch.chdir(GOLD)
    with pytest.raises(tasks.TaskError) as e:
        tasks.TaskJob(7, _command_line(a, str(tmp_path)))
    assert e.value.rc == -1
    with pytest.raises(tasks.TaskError) as e:  # a usage error: a required option missing
        tasks.TaskJob(tasks.LINEAR_SEARCH, ["--input", a["gather"]])
    assert e.value.rc == -1
    job = tasks.TaskJob(tasks.LINEAR_SEARCH, _command_line(a, str(tmp_path)))
    m0 = struct.unpack("<f", job.next_task())[0]
    with pytest.raises(tasks.TaskError) as e:  # short record
        job.commit_task(np.float32(m0).tobytes())
    assert e.value.rc == -1
    with pytest.raises(tasks.TaskError) as e:  # a midpoint that is no CDP of the line
        job.commit_task(np.float32(m0 + 0.25).tobytes() + bytes(job.result_bytes - 4))
    assert e.value.rc == -1
    job.commit_task(np.float32(m0).tobytes() + bytes(job.result_bytes - 4))
    job.close()
