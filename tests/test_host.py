"""The C++17 host layer on the CPU (libhss_host.so through include/hss_host.h): the SU
reader against what the reference's own Gather::readGatherFromFile kept for the same
files (fixtures written by the reference-built checker), the traveltime bounds against
the reference's TraveltimeBuilder, the output files byte for byte against the reference's
own Dumper, and the command-line rules."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from highperformanceseismicstacking_amd import PKG_DIR

GOLD = os.path.join(os.path.dirname(__file__), "golden")
HOST_LIB = os.path.join(PKG_DIR, "libhss_host.so")


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOST_LIB):
        pytest.fail("libhss_host.so is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(HOST_LIB)
    L.hss_host_last_error.restype = ctypes.c_char_p
    vp, f = ctypes.c_void_p, ctypes.c_float
    L.hss_host_read_gather.argtypes = [ctypes.c_char_p, f, f, f, f, vp, vp, vp, vp, vp, vp, vp, vp]
    L.hss_host_gather_arrays.argtypes = [vp, vp, vp, vp, vp]
    L.hss_host_traveltime_bounds.argtypes = [ctypes.c_int, vp, vp, ctypes.c_int32, f, f, f, vp, vp]
    L.hss_host_parse_command_line.argtypes = [ctypes.c_int, ctypes.c_int, vp]
    L.hss_host_xcd_balanced_order.argtypes = [vp, ctypes.c_int32, vp]
    L.hss_host_dump_results.argtypes = [ctypes.c_char_p] * 4 + [ctypes.c_int, vp, vp,
                                        ctypes.c_int32, f, f, f, f, vp, vp, ctypes.c_char_p,
                                        ctypes.c_int32]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def read_gather(host, path, aph, apm, tau, az, lowest=None, highest=None):
    sc = [np.zeros(1, np.int32), np.zeros(1, np.float32), np.zeros(1, np.int32),
          np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1, np.int32)]
    lo = np.array([lowest], np.float32) if lowest is not None else None
    hi = np.array([highest], np.float32) if highest is not None else None
    rc = host.hss_host_read_gather(path.encode(), aph, apm, tau, az,
                                   _p(lo) if lo is not None else None,
                                   _p(hi) if hi is not None else None, *[_p(x) for x in sc])
    assert rc == 0, host.hss_host_last_error()
    ns, dt, tau_i, w, ntr, ncdp = (x[0] for x in sc)
    mid = np.zeros(ntr, np.float32)
    hoff = np.zeros(ntr, np.float32)
    samples = np.zeros((ntr, ns), np.float32)
    cdps = np.zeros(ncdp, np.float32)
    counts = np.zeros(ncdp, np.int32)
    assert host.hss_host_gather_arrays(_p(mid), _p(hoff), _p(samples), _p(cdps), _p(counts)) == 0
    return dict(ns=ns, dt=dt, tau_index=tau_i, w=w, mid=mid, hoff=hoff, samples=samples,
                cdps=cdps, counts=counts)


@pytest.mark.parametrize("name", ["basic", "filters"])
def test_su_reader_matches_reference_reader(host, name):
    d = np.load(os.path.join(GOLD, f"gather_{name}.npz"))
    aph, apm, tau, az = (float(x) for x in d["params"])
    lowest = None if str(d["lowest"]) == "-" else float(str(d["lowest"]))
    highest = None if str(d["highest"]) == "-" else float(str(d["highest"]))
    got = read_gather(host, os.path.join(GOLD, f"gather_{name}.su"), aph, apm, tau, az,
                      lowest, highest)
    assert got["ns"] == int(d["ns"]) and got["dt"] == d["dt"]
    assert got["tau_index"] == int(d["tau_index"]) and got["w"] == int(d["w"])
    # the CDP map, including the entry at m = 0 the reader's extra read at end of file adds
    assert np.array_equal(got["cdps"], d["cdps"])
    assert got["cdps"][0] == 0.0
    assert np.array_equal(got["mid"].view(np.uint32), d["mid"].view(np.uint32))
    assert np.array_equal(got["hoff"].view(np.uint32), d["hoff"].view(np.uint32))
    assert np.array_equal(got["samples"].view(np.uint32), d["samples"].view(np.uint32))
    assert got["counts"].sum() == len(d["mid"])


def test_su_reader_errors(host, tmp_path):
    ns_ = np.zeros(1, np.int32)
    args = [np.zeros(1, t) for t in (np.int32, np.float32, np.int32, np.int32, np.int32, np.int32)]
    rc = host.hss_host_read_gather(str(tmp_path / "missing.su").encode(), 1.0, 0.0, 0.02, 90.0,
                                   None, None, *[_p(a) for a in args])
    assert rc == -1 and b"cannot open input file" in host.hss_host_last_error()
    # a record cut inside its samples is a runtime error (Trace.cpp:58-60)
    raw = open(os.path.join(GOLD, "gather_basic.su"), "rb").read()
    cut = tmp_path / "cut.su"
    cut.write_bytes(raw[: 240 + 4 * 32 + 240 + 40])
    rc = host.hss_host_read_gather(str(cut).encode(), 1e4, 0.0, 0.02, 90.0, None, None,
                                   *[_p(a) for a in args])
    assert rc == -3 and b"ends inside the samples" in host.hss_host_last_error()


def test_traveltime_bounds_match_reference_builder(host):
    d = np.load(os.path.join(GOLD, "traveltime.npz"))
    for row in d["zocrs"]:
        vmin, vmax, v0, ain, bp = row[:5]
        lo_ref = row[[5, 7, 9]].astype(np.float32)
        hi_ref = row[[6, 8, 10]].astype(np.float32)
        lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
        lo_in = np.array([vmin], np.float32)
        hi_in = np.array([vmax], np.float32)
        assert host.hss_host_traveltime_bounds(1, _p(lo_in), _p(hi_in), 1, v0, ain, bp, _p(lo), _p(hi)) == 0
        assert np.array_equal(lo, lo_ref) and np.array_equal(hi, hi_ref)
    lo_in = np.array([1.0], np.float32)
    rc = host.hss_host_traveltime_bounds(2, _p(lo_in), _p(lo_in), 1, 0, 0, 0, _p(lo), _p(hi))
    assert rc == -2  # OCT needs two bounds each: logic_error


def parse(host, program, *args):
    argv = [b"prog"] + [a.encode() for a in args]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    return host.hss_host_parse_command_line(program, len(argv), ctypes.cast(arr, ctypes.c_void_p))


BASE = ["--aph", "1000", "--apm", "150", "--azimuth", "0", "--input", "x.su", "--output", "out",
        "--tau", "0.02"]


def test_command_lines_of_the_reference_test_scripts(host):
    # the flag sets of the reference's run scripts (opencl/test/test_greedy.sh, test_de.sh)
    assert parse(host, 0, *BASE, "--traveltime", "cmp", "--lower-bounds", "450",
                 "--upper-bounds", "6000", "--granularity", "1024") == 0
    assert parse(host, 0, *BASE, "--traveltime", "zocrs", "--lower-bounds", "450",
                 "--upper-bounds", "6000", "--granularity", "100", "10", "10", "--ain", "60",
                 "--v0", "2000", "--bpctg", "0.1") == 0
    assert parse(host, 0, *BASE, "--traveltime", "oct", "--lower-bounds", "450", "-0.001",
                 "--upper-bounds", "6000", "0.001", "--granularity", "100", "100") == 0
    assert parse(host, 1, *BASE, "--traveltime", "cmp", "--lower-bounds", "450",
                 "--upper-bounds", "6000", "--generations", "32", "--population-size", "32",
                 "--thread-count", "128", "--verbose", "1") == 0
    assert parse(host, 2, *BASE, "--traveltime", "cmp", "--lower-bounds", "450",
                 "--upper-bounds", "6000", "--stack-datafiles", "velocity.su") == 0
    assert parse(host, 0, *BASE, "--traveltime=cmp", "--lower-bounds", "450",
                 "--upper-bounds", "6000", "--granularity", "64", "--kernel-path", "k/",
                 "--variant", "exact", "--storage", "f16", "--batch-size", "8") == 0


def test_command_line_errors(host):
    ok = ["--traveltime", "cmp", "--lower-bounds", "450", "--upper-bounds", "6000"]
    # missing required option -> invalid_argument (main prints it and the help, returns 0)
    assert parse(host, 0, *BASE, *ok) == -1
    assert b"Unable to parse input arguments" in host.hss_host_last_error()
    assert parse(host, 0, *BASE, *ok, "--granularity", "10", "--bogus", "1") == -1
    assert parse(host, 0, *BASE, "--traveltime", "cmx", "--lower-bounds", "450",
                 "--upper-bounds", "6000", "--granularity", "10") == -1
    assert parse(host, 0, *BASE, "--traveltime", "zocrs", "--lower-bounds", "450",
                 "--upper-bounds", "6000", "--granularity", "1", "1", "1") == -1
    assert b"Not all arguments" in host.hss_host_last_error()
    assert parse(host, 0, *BASE, "--traveltime", "oct", "--lower-bounds", "450",
                 "--upper-bounds", "6000", "--granularity", "10") == -2
    assert parse(host, 0, *BASE, *ok, "--granularity", "10", "--aph", "5") == -1
    assert parse(host, 0, *BASE, *ok, "--granularity", "ten") == -1


def test_xcd_balanced_order_is_a_balanced_permutation():
    """The launch order of a batch: a permutation whose eight contiguous chunks -- the
    split the kernels' workgroup order gives the XCDs (sizes n//8 + 1 for the first n%8,
    n//8 for the rest) -- carry total work within one midpoint of each other, each chunk
    largest first."""
    from highperformanceseismicstacking_amd import capi
    rng = np.random.default_rng(0)
    for n in (1, 7, 8, 9, 17, 500, 2001):
        work = rng.integers(24, 121, size=n)
        order = capi.xcd_balanced_order(work)
        assert sorted(order.tolist()) == list(range(n))
        q, r = divmod(n, 8)
        chunks, start = [], 0
        for x in range(8):
            sz = q + (1 if x < r else 0)
            chunks.append(work[order[start:start + sz]])
            start += sz
        sums = [int(c.sum()) for c in chunks if len(c)]
        if n >= 16:
            assert max(sums) - min(sums) <= int(work.max())
        for c in chunks:
            assert np.all(np.diff(c) <= 0)


@pytest.mark.parametrize("n", [0, 1, 5, 8, 13, 64, 300, 2001])
def test_xcd_balanced_order(host, n):
    """The launch order the host layer gives a batch (hss_midpoints.order): a permutation
    of the midpoints, the same as the Python binding's, largest work first in each of the
    8 bins, and bins whose total work differs by at most two midpoints' work."""
    from highperformanceseismicstacking_amd import capi
    rng = np.random.default_rng(n)
    work = rng.integers(0, 2500, n).astype(np.int32)
    work[: n // 4] = work[0] if n else 0  # ties keep the stable order
    order = np.full(max(n, 1), -7, np.int32)
    assert host.hss_host_xcd_balanced_order(_p(work), n, _p(order)) == 0
    order = order[:n]
    assert sorted(order.tolist()) == list(range(n))
    assert np.array_equal(order, capi.xcd_balanced_order(work))
    full, extra = divmod(n, 8)
    sizes = [full + (b < extra) for b in range(8)]
    bins = np.split(order, np.cumsum(sizes)[:-1])
    for b in bins:
        assert np.all(np.diff(work[b]) <= 0)
    sums = [int(work[b].sum()) for b in bins]
    assert max(sums) - min(sums) <= 2 * int(work.max(initial=0))
    assert host.hss_host_xcd_balanced_order(_p(work), -1, _p(order)) == -1


DUMP_CASES = sorted(d for d in os.listdir(GOLD) if d.startswith("dump_"))


@pytest.mark.parametrize("case", DUMP_CASES)
def test_output_files_equal_the_reference_dumper_byte_for_byte(host, case, tmp_path,
                                                               monkeypatch):
    """The output layer writes the same bytes as the reference's Dumper (fixtures written
    by the reference-built checker from common/src/output/Dumper.cpp, ResultSet.cpp and
    the result classes) for the same gather, traveltime and result set: every SU file
    (CDP headers, the end-of-file CDP's included, and samples), every CSV (default stream
    formatting of m0 and the statistic), gather.txt and traveltime.txt; only the
    directory's timestamp differs."""
    fix = os.path.join(GOLD, case)
    a = json.load(open(os.path.join(fix, "args.json")))
    monkeypatch.chdir(GOLD)                      # gather.txt names the input as given
    read_gather(host, a["gather"], a["aph"], a["apm"], a["tau"], a["azimuth"])
    values = np.fromfile(os.path.join(fix, "values.f32"), np.float32)
    stats = np.fromfile(os.path.join(fix, "stats.f32"), np.float32)
    lo = np.array(a["lower"], np.float32)
    hi = np.array(a["upper"], np.float32)
    model = {"cmp": 0, "zocrs": 1, "oct": 2}[a["model"]]
    buf = ctypes.create_string_buffer(4096)
    rc = host.hss_host_dump_results(str(tmp_path).encode(), a["gather"].encode(),
                                    a["data_name"].encode(), a["method"].encode(), model,
                                    _p(lo), _p(hi), len(lo), a["v0"], a["ain"], a["bpctg"],
                                    a["h0"], _p(values), _p(stats), buf, 4096)
    assert rc == 0, host.hss_host_last_error()
    out = buf.value.decode()
    name = os.path.basename(out.rstrip("/"))
    assert name.startswith(a["directory"]) and re.fullmatch(r"\d{8}_\d{4}", name[len(a["directory"]):])
    assert sorted(os.listdir(out)) == a["files"]
    for f in a["files"]:
        got = open(os.path.join(out, f), "rb").read()
        want = open(os.path.join(fix, f), "rb").read()
        assert got == want, f
