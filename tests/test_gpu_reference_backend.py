"""The drop-in proven with the reference's own drivers: oracle/_ref/ref_hip_backend runs
LinearSearchAlgorithm, DifferentialEvolutionAlgorithm and StretchFreeAlgorithm from the
reference's common/src (compiled unmodified) over the reference-side HIP backend of
highperformanceseismicstacking_amd/reference_backend (HipComputeAlgorithmBuilder,
HipDataContainer, HipDeviceContext) and libhss.so, for every CDP of a golden SU file.
With HSS_VARIANT=exact its results equal the oracle's bit for bit: grid search for the
three models (OCT over the reference's 1024 random selection samples), differential
evolution with a fixed Philox key (midpoint k of the run uses stream k), and stretch-free
re-stacking from a velocity file.  The driver is built in the build container from the
reference's sources (oracle/ref/Makefile); without it the tests are skipped."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from highperformanceseismicstacking_amd import synth
from helpers import grid_steps, zocrs_bounds

pytestmark = pytest.mark.gpu
f32 = np.float32
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_hip_backend")
GOLD = os.path.join(ROOT, "tests", "golden")
SU = os.path.join(GOLD, "gather_basic.su")
APH, APM, TAU, AZ = 500.0, 60.0, 0.02, 90.0


def _need_driver():
    if not os.path.exists(DRIVER):
        pytest.skip("reference-built backend driver absent (built only where the reference "
                    "sources are)")


def _run(method, model, lo, hi, extra, tmp, v0=0.0, ain=0.0, bpctg=0.0, h0=0.0, env=None):
    out = str(tmp / f"{method}_{model}.bin")
    args = [DRIVER, method, SU, APH, APM, TAU, AZ, model, len(lo), *lo, *hi, v0, ain, bpctg,
            h0, *extra, out]
    e = dict(os.environ, HSS_VARIANT="exact", **(env or {}))
    print("reference backend:", method, model, flush=True)
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120, env=e)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.fromfile(out, np.float32)


def _gather():
    """The reader's view of gather_basic.su with these read parameters (the CPU host
    layer's reader, itself checked against the reference reader)."""
    from highperformanceseismicstacking_amd import HOST_LIB_PATH
    import ctypes
    H = ctypes.CDLL(HOST_LIB_PATH)
    f = ctypes.c_float
    sc = [ctypes.c_int32(), ctypes.c_float(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(),
          ctypes.c_int32()]
    assert H.hss_host_read_gather(SU.encode(), f(APH), f(APM), f(TAU), f(AZ), None, None,
                                  *[ctypes.byref(x) for x in sc]) == 0
    ns, dt, tau, w, ntr, ncdp = (x.value for x in sc)
    mid, hoff = np.zeros(ntr, np.float32), np.zeros(ntr, np.float32)
    samples = np.zeros((ntr, ns), np.float32)
    cdps, counts = np.zeros(ncdp, np.float32), np.zeros(ncdp, np.int32)
    vp = ctypes.c_void_p
    assert H.hss_host_gather_arrays(vp(mid.ctypes.data), vp(hoff.ctypes.data),
                                    vp(samples.ctypes.data), vp(cdps.ctypes.data),
                                    vp(counts.ctypes.data)) == 0
    return dict(ns=ns, dt=f32(dt), tau=tau, mid=mid, hoff=hoff, samples=samples, cdps=cdps)


def _oracle_gather(G, idx, m0, h0=0.0, apm=APM):
    return O.Gather(G["samples"][idx], G["mid"][idx], G["hoff"][idx], G["dt"], G["tau"],
                    m0=float(m0), h0=h0, apm=apm)


@pytest.mark.parametrize("model", ["cmp", "zocrs", "oct"])
def test_reference_grid_driver_on_hip_backend_equals_oracle(model, tmp_path):
    _need_driver()
    G = _gather()
    if model == "cmp":
        lo, hi, gran, extra = [1500.0], [4500.0], [24], {}
        mcode = O.CMP
    elif model == "zocrs":
        lo, hi = zocrs_bounds(1600.0, 4800.0, 60.0, 2000.0, 0.5)
        gran, extra = [6, 4, 4], dict(v0=2000.0, ain=60.0, bpctg=0.5)
        mcode = O.ZOCRS
    else:
        lo, hi, gran, extra = [1500.0, -5e-4], [4500.0, 5e-4], [8, 4], dict(h0=100.0)
        mcode = O.OCT
    cli_lo = lo if model != "zocrs" else [1600.0]
    cli_hi = hi if model != "zocrs" else [4800.0]
    got = _run("greedy", model, cli_lo, cli_hi, gran, tmp_path, **extra)
    np_ = O.NPARAMS[mcode]
    ns = G["ns"]
    got = got.reshape(len(G["cdps"]), 2 + np_, ns)
    params = O.build_parameter_array(mcode, lo, grid_steps(lo, hi, gran), gran)
    h0 = extra.get("h0", 0.0)
    pairs = O.oct_selection_samples(lo[0], hi[0], lo[1], hi[1]) if mcode == O.OCT else None
    checked = 0
    for k, m0 in enumerate(G["cdps"]):
        mask = O.select_traces(mcode, G["mid"], G["hoff"], m0, apm=APM, h0=h0, ns=ns,
                               dt=G["dt"], pairs=pairs, npairs=1024 if pairs is not None else 0)
        idx = np.flatnonzero(mask)
        if idx.size == 0:
            assert not got[k].any(), k        # no selected trace: the results stay zero
            continue
        rres, _, _, _, _ = O.grid_f32(mcode, _oracle_gather(G, idx, m0, h0), params, full=False)
        assert np.array_equal(got[k].view(np.uint32), rres.view(np.uint32)), (k, m0)
        checked += 1
    assert checked >= 3


@pytest.mark.parametrize("model", ["cmp", "zocrs"])
def test_reference_de_driver_on_hip_backend_equals_oracle(model, tmp_path):
    _need_driver()
    G = _gather()
    seed, NP, gens = 987654321, 8, 5
    if model == "cmp":
        lo, hi, extra, mcode = [1500.0], [4500.0], {}, O.CMP
        cli_lo, cli_hi = lo, hi
    else:
        lo, hi = zocrs_bounds(1600.0, 4800.0, 60.0, 2000.0, 0.5)
        extra, mcode = dict(v0=2000.0, ain=60.0, bpctg=0.5), O.ZOCRS
        cli_lo, cli_hi = [1600.0], [4800.0]
    got = _run("de", model, cli_lo, cli_hi, [gens, NP], tmp_path,
               env={"HSS_DE_SEED": str(seed)}, **extra)
    np_ = O.NPARAMS[mcode]
    got = got.reshape(len(G["cdps"]), 2 + np_, G["ns"])
    checked = 0
    for k, m0 in enumerate(G["cdps"]):
        idx = np.flatnonzero(O.select_traces(mcode, G["mid"], G["hoff"], m0, apm=APM))
        if idx.size == 0:
            assert not got[k].any(), k
            continue
        ref = O.de_f32(mcode, _oracle_gather(G, idx, m0), lo, hi, NP, gens, seed, stream_id=k)
        assert np.array_equal(got[k].view(np.uint32), ref.view(np.uint32)), (k, m0)
        checked += 1
    assert checked >= 3


def test_reference_stretch_free_driver_on_hip_backend_equals_oracle(tmp_path):
    _need_driver()
    G = _gather()
    rng = np.random.default_rng(3)
    ns = G["ns"]
    vel = rng.uniform(1600.0, 4400.0, (len(G["cdps"]), ns)).astype(np.float32)
    # one velocity trace per CDP, in map order (what the reference reads per CDP)
    line = synth.Line(vel, np.zeros(len(vel), np.int32), np.zeros(len(vel), np.int32), 1, ns,
                      int(round(float(G["dt"]) * 1e6)), np.arange(len(vel), dtype=np.int32))
    vfile = str(tmp_path / "velocity.su")
    synth.write_su(vfile, line)
    got = _run("stretch_free", "cmp", [1500.0], [4500.0], [vfile], tmp_path)
    got = got.reshape(len(G["cdps"]), 3, ns)
    nvals = np.arange(-10, 11, dtype=np.float32)
    checked = 0
    for k, m0 in enumerate(G["cdps"]):
        idx = np.flatnonzero(O.select_traces(O.CMP, G["mid"], G["hoff"], m0))
        if idx.size == 0:
            assert not got[k].any(), k
            continue
        ref, _, _ = O.stretch_free_f32(O.CMP, _oracle_gather(G, idx, m0), vel[k], nvals)
        assert np.array_equal(got[k].view(np.uint32), ref.view(np.uint32)), (k, m0)
        checked += 1
    assert checked >= 3
