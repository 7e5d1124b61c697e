"""ctypes loader for the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, and only as
the checker.  The product path (libhss.so and the C++ host) never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

CMP, ZOCRS, OCT = 0, 1, 2
NPARAMS = {CMP: 1, ZOCRS: 3, OCT: 2}

_lib = None


class orc_gather(ctypes.Structure):
    _fields_ = [
        ("samples", ctypes.c_void_p), ("mid", ctypes.c_void_p), ("hoff", ctypes.c_void_p),
        ("hoff2", ctypes.c_void_p), ("ntr", ctypes.c_int), ("ns", ctypes.c_int),
        ("dt", ctypes.c_float), ("tau", ctypes.c_int), ("w", ctypes.c_int),
        ("m0", ctypes.c_float), ("h0", ctypes.c_float), ("apm", ctypes.c_float),
    ]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        fl, i = ctypes.c_float, ctypes.c_int
        L.orc_oct_displaced_midpoint.argtypes = [fl, fl, fl, fl, fl, fl, ctypes.POINTER(fl)]
        L.orc_oct_time.argtypes = [fl, fl, fl, fl, fl, fl, fl, fl, ctypes.POINTER(fl)]
        L.orc_philox4x32_10.argtypes = [vp, vp, vp]
        L.orc_eval_f32.argtypes = [i, vp, i, i, vp, vp, vp, vp, vp]
        L.orc_eval_f64.argtypes = [i, vp, i, i, vp, vp, vp, vp, vp]
        L.orc_build_parameter_array.argtypes = [i, vp, vp, vp, vp]
        L.orc_grid_f32.argtypes = [i, vp, vp, i, vp, vp, vp, vp, vp, i]
        L.orc_grid_f64.argtypes = [i, vp, vp, i, vp, vp, vp, vp, i]
        L.orc_stretch_free_f32.argtypes = [i, vp, vp, vp, i, vp, vp, vp, i]
        L.orc_stretch_free_f64.argtypes = [i, vp, vp, vp, i, vp, i]
        L.orc_de_f32.argtypes = [i, vp, vp, vp, i, i, ctypes.c_uint64, ctypes.c_uint32,
                                 vp, vp, vp, vp, i]
        L.orc_de_start_individual.argtypes = [i, vp, vp, vp, vp]
        L.orc_de_trial.argtypes = [i, i, vp, i, vp, vp, vp, vp, vp]
        L.orc_select_traces.argtypes = [i, vp, vp, i, fl, fl, fl, i, fl, vp, i, i, vp]
        L.orc_oct_selection_samples.argtypes = [fl, fl, fl, fl, i, vp]
        L.orc_eval_grad_f64.argtypes = [i, vp, i, vp, i, vp, vp, vp]
        L.orc_traveltime_f32.argtypes = [i, vp, i, fl, vp, ctypes.POINTER(fl)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Gather:
    """Host arrays of the traces one midpoint uses, and the orc_gather view of them."""

    def __init__(self, samples, mid, hoff, dt, tau, m0=0.0, h0=0.0, apm=0.0, hoff2=None):
        self.samples = np.ascontiguousarray(samples, dtype=np.float32)
        self.mid = np.ascontiguousarray(mid, dtype=np.float32)
        self.hoff = np.ascontiguousarray(hoff, dtype=np.float32)
        self.hoff2 = (np.ascontiguousarray(hoff2, dtype=np.float32) if hoff2 is not None
                      else (self.hoff * self.hoff).astype(np.float32))
        ntr, ns = self.samples.shape
        self.ns = ns
        self.w = 2 * int(tau) + 1
        self.s = orc_gather(_p(self.samples), _p(self.mid), _p(self.hoff), _p(self.hoff2),
                            ntr, ns, float(dt), int(tau), self.w, float(m0), float(h0), float(apm))

    @property
    def ref(self):
        return ctypes.byref(self.s)


def build_parameter_array(model, lower, step, count):
    n = NPARAMS[model]
    lo = np.array(list(lower[:n]) + [0] * (3 - n), dtype=np.float32)
    st = np.array(list(step[:n]) + [0] * (3 - n), dtype=np.float32)
    ct = np.array(list(count[:n]) + [1] * (3 - n), dtype=np.int32)
    P = int(np.prod(ct[:n]))
    out = np.empty(P * n, dtype=np.float32)
    lib().orc_build_parameter_array(model, _p(lo), _p(st), _p(ct), _p(out))
    return out


def grid_f32(model, g, params, nthreads=8, full=True):
    n = NPARAMS[model]
    params = np.ascontiguousarray(params, dtype=np.float32)
    P = params.size // n
    sembl = np.empty((g.ns, P), dtype=np.float32) if full else None
    stack = np.empty((g.ns, P), dtype=np.float32) if full else None
    res = np.empty((2 + n, g.ns), dtype=np.float32)
    best = np.empty(g.ns, dtype=np.int32)
    nu = np.zeros(1, dtype=np.float64)
    lib().orc_grid_f32(model, g.ref, _p(params), P, _p(sembl), _p(stack), _p(res), _p(best),
                       _p(nu), nthreads)
    return res, sembl, stack, best, float(nu[0])


def grid_f64(model, g, params, nthreads=8, full=True):
    n = NPARAMS[model]
    params = np.ascontiguousarray(params, dtype=np.float32)
    P = params.size // n
    sembl = np.empty((g.ns, P), dtype=np.float64) if full else None
    stack = np.empty((g.ns, P), dtype=np.float64) if full else None
    res = np.empty((2 + n, g.ns), dtype=np.float64)
    best = np.empty(g.ns, dtype=np.int32)
    lib().orc_grid_f64(model, g.ref, _p(params), P, _p(sembl), _p(stack), _p(res), _p(best),
                       nthreads)
    return res, sembl, stack, best


def eval_f32(model, g, s, prm, shift_n=0):
    prm = np.ascontiguousarray(prm, dtype=np.float32)
    S = np.zeros(1, np.float32)
    st = np.zeros(1, np.float32)
    used = np.zeros(1, np.uint32)
    nu = np.zeros(1, np.uint32)
    lib().orc_eval_f32(model, g.ref, s, shift_n, _p(prm), _p(S), _p(st), _p(used), _p(nu))
    return float(S[0]), float(st[0]), int(used[0]), int(nu[0])


def eval_f64(model, g, s, prm, shift_n=0):
    prm = np.ascontiguousarray(prm, dtype=np.float32)
    S = np.zeros(1, np.float64)
    st = np.zeros(1, np.float64)
    used = np.zeros(1, np.uint32)
    nu = np.zeros(1, np.uint32)
    lib().orc_eval_f64(model, g.ref, s, shift_n, _p(prm), _p(S), _p(st), _p(used), _p(nu))
    return float(S[0]), float(st[0]), int(used[0]), int(nu[0])


def stretch_free_f32(model, g, vparams, nvals, nthreads=8, full=False):
    vparams = np.ascontiguousarray(vparams, dtype=np.float32)
    nvals = np.ascontiguousarray(nvals, dtype=np.float32)
    nN = nvals.size
    sembl = np.empty((g.ns, nN), np.float32) if full else None
    stack = np.empty((g.ns, nN), np.float32) if full else None
    res = np.empty((3, g.ns), np.float32)
    lib().orc_stretch_free_f32(model, g.ref, _p(vparams), _p(nvals), nN, _p(sembl), _p(stack),
                               _p(res), nthreads)
    return res, sembl, stack


def stretch_free_f64(model, g, vparams, nvals, nthreads=8):
    vparams = np.ascontiguousarray(vparams, dtype=np.float32)
    nvals = np.ascontiguousarray(nvals, dtype=np.float32)
    res = np.empty((3, g.ns), np.float64)
    lib().orc_stretch_free_f64(model, g.ref, _p(vparams), _p(nvals), nvals.size, _p(res),
                               nthreads)
    return res


def de_f32(model, g, lower, upper, NP, G, seed, stream_id=0, nthreads=8, population=False,
           not_used=False):
    """result [(2+np)][ns]; with population=True also the final population x, fx; with
    not_used=True also the traces the model rejected over the whole run (last item)."""
    n = NPARAMS[model]
    lo = np.ascontiguousarray(lower[:n], dtype=np.float32)
    hi = np.ascontiguousarray(upper[:n], dtype=np.float32)
    res = np.empty((2 + n, g.ns), np.float32)
    x = np.empty((g.ns, NP, n), np.float32) if population else None
    fx = np.empty((g.ns, NP, 2), np.float32) if population else None
    nu = np.zeros(1, np.uint64) if not_used else None
    lib().orc_de_f32(model, g.ref, _p(lo), _p(hi), NP, G, seed, stream_id, _p(res), _p(x), _p(fx),
                     _p(nu), nthreads)
    out = (res, x, fx) if population else (res,)
    if not_used:
        out = out + (int(nu[0]),)
    return out if len(out) > 1 else out[0]


def de_start_individual(lower, upper, u):
    """x_d = lower_d + u_d * (upper_d - lower_d) for explicit uniforms u [D]."""
    lo = np.ascontiguousarray(lower, dtype=np.float32)
    hi = np.ascontiguousarray(upper, dtype=np.float32)
    u = np.ascontiguousarray(u, dtype=np.float32)
    out = np.empty(lo.size, np.float32)
    lib().orc_de_start_individual(lo.size, _p(lo), _p(hi), _p(u), _p(out))
    return out


def de_trial(x, i, lower, upper, u):
    """(mutant, trial) of individual i of the population x [NP][D] for explicit uniforms
    u [4D+1]: u[3d..3d+2] donor draws of parameter d, u[3D] the forced crossover index
    draw, u[3D+1+d] the crossover coin of parameter d."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    NP, D = x.shape
    lo = np.ascontiguousarray(lower, dtype=np.float32)
    hi = np.ascontiguousarray(upper, dtype=np.float32)
    u = np.ascontiguousarray(u, dtype=np.float32)
    assert u.size == 4 * D + 1 and lo.size == D and hi.size == D
    mut = np.empty(D, np.float32)
    tri = np.empty(D, np.float32)
    lib().orc_de_trial(D, NP, _p(x), int(i), _p(lo), _p(hi), _p(u), _p(mut), _p(tri))
    return mut, tri


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def select_traces(model, mid, hoff, m0, apm=0.0, h0=0.0, ns=0, dt=0.0, pairs=None, npairs=0,
                  per_sample=False):
    mid = np.ascontiguousarray(mid, dtype=np.float32)
    hoff = np.ascontiguousarray(hoff, dtype=np.float32)
    mask = np.zeros(mid.size, np.uint8)
    pr = np.ascontiguousarray(pairs, dtype=np.float32) if pairs is not None else None
    lib().orc_select_traces(model, _p(mid), _p(hoff), mid.size, float(m0), float(apm), float(h0),
                            int(ns), float(dt), _p(pr), int(npairs), int(per_sample), _p(mask))
    return mask


def oct_selection_samples(vmin, vmax, smin, smax, count=1024):
    out = np.empty(count * 2, np.float32)
    lib().orc_oct_selection_samples(vmin, vmax, smin, smax, count, _p(out))
    return out


def oct_displaced_midpoint(h, h0, t0, m0, c, slope):
    mh = ctypes.c_float(0)
    err = lib().orc_oct_displaced_midpoint(h, h0, t0, m0, c, slope, ctypes.byref(mh))
    return err, mh.value


def oct_time(h, h0, t0, m0, m, mh, c, slope):
    t = ctypes.c_float(0)
    err = lib().orc_oct_time(h, h0, t0, m0, m, mh, c, slope, ctypes.byref(t))
    return err, t.value


def traveltime_f32(model, g, i, t0, prm):
    """(kept, t) of trace i at t0 for the storage-form candidate prm: kept is 0 when the
    model rejects the trace (ZO-CRS negative radicand, OCT error or aperture)."""
    prm = np.ascontiguousarray(prm, dtype=np.float32)
    t = ctypes.c_float(0)
    kept = lib().orc_traveltime_f32(model, g.ref, int(i), float(t0), _p(prm), ctypes.byref(t))
    return kept, t.value


def eval_grad_f64(model, g, s, pv, f32_index=False):
    """FP64 (S, stack, dS/dparams) at sample s for the parameter point pv (v space)
    (oracle_grad.c).  f32_index=False: every step in double, including t/dt and the
    sample index; True: traveltime, index, weight and trace decisions in FP32 as the
    FP32 evaluation takes them, derivatives and sums in double."""
    n = NPARAMS[model]
    pv = np.ascontiguousarray(pv, dtype=np.float64)
    S = np.zeros(1)
    T = np.zeros(1)
    grad = np.zeros(n)
    lib().orc_eval_grad_f64(model, g.ref, int(s), _p(pv), int(bool(f32_index)), _p(S), _p(T),
                            _p(grad))
    return float(S[0]), float(T[0]), grad
