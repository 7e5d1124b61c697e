"""ctypes binding of libhss.so (include/hss.h).

Device buffers are torch tensors on ``cuda:0`` (torch is only the allocator and the
stream here); every call goes through the C ABI.  Loading fails loudly when the HIP
library has not been built: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import LIB_PATH, MODEL_NPARAMS, F32, F16, BF16, EXACT, FAST

_lib = None


class HssError(RuntimeError):
    pass


class hss_traces(ctypes.Structure):
    _fields_ = [
        ("samples", ctypes.c_void_p),
        ("scale", ctypes.c_void_p),
        ("mid", ctypes.c_void_p),
        ("hoff", ctypes.c_void_p),
        ("hoff2", ctypes.c_void_p),
        ("ntr", ctypes.c_int32),
        ("ns", ctypes.c_int32),
        ("dt", ctypes.c_float),
        ("tau", ctypes.c_int32),
        ("w", ctypes.c_int32),
        ("storage", ctypes.c_int32),
    ]


class hss_midpoints(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("begin", ctypes.c_void_p),
        ("count", ctypes.c_void_p),
        ("m0", ctypes.c_void_p),
        ("index", ctypes.c_void_p),
        ("h0", ctypes.c_float),
        ("apm", ctypes.c_float),
        ("order", ctypes.c_void_p),
    ]


class hss_oct_candidates(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32),
        ("d_pairs", ctypes.c_void_p),
        ("npairs", ctypes.c_int32),
        ("d_vparams", ctypes.c_void_p),
        ("NP", ctypes.c_int32),
        ("lower", ctypes.c_float * 2),
        ("upper", ctypes.c_float * 2),
        ("seed", ctypes.c_uint64),
        ("stream_base", ctypes.c_uint32),
    ]


OCT_PAIRS, OCT_PER_SAMPLE, OCT_DE_START = 0, 1, 2


def lib():
    """The loaded libhss.so (raises HssError when it is not built)."""
    global _lib
    if _lib is None:
        path = os.environ.get("HSS_LIBRARY", LIB_PATH)  # another build of the same ABI
        if not os.path.exists(path):
            raise HssError(f"{path} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        vp, i32, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_float
        ci, u64, u32 = ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32
        tr, mp = ctypes.POINTER(hss_traces), ctypes.POINTER(hss_midpoints)
        L.hss_abi_version.restype = ci
        L.hss_status_string.restype = ctypes.c_char_p
        L.hss_status_string.argtypes = [ci]
        L.hss_device_count.argtypes = [ctypes.POINTER(ci)]
        L.hss_verify_index_ops.argtypes = [f32, ctypes.POINTER(u64), vp]
        L.hss_index_ops_state.argtypes = [f32, ctypes.POINTER(ci)]
        L.hss_build_parameter_array.argtypes = [ci, vp, vp, vp, vp, vp]
        L.hss_grid_semblance.argtypes = [tr, mp, ci, vp, i32, ci, vp, vp, vp, vp, vp]
        L.hss_grid_semblance_plain.argtypes = [tr, ci, vp, i32, i32, vp, vp, vp, f32, f32, vp,
                                               vp, vp, vp]
        L.hss_de_workspace_size.argtypes = [i32, i32, ci, i32, ctypes.POINTER(u64)]
        L.hss_de_semblance.argtypes = [tr, mp, ci, vp, vp, i32, i32, i32, i32, u64, u32, ci, vp,
                                       u64, vp, vp, vp, vp, vp]
        L.hss_de_semblance_noalloc.argtypes = [tr, mp, ci, vp, vp, i32, i32, i32, i32, u64, u32,
                                               ci, vp, vp, vp, vp, vp]
        L.hss_stretch_free.argtypes = [tr, mp, ci, vp, vp, i32, ci, vp, vp, vp, vp, vp]
        L.hss_semblance_gradient.argtypes = [tr, mp, ci, vp, vp, vp]
        L.hss_grid_set_config.argtypes = [ci, ci, ci, ci, i32]
        L.hss_semblance_vjp.argtypes = [i32, i32, i32, vp, vp, vp, vp]
        L.hss_grid_tuned_config.argtypes = [ci, ci, ci, ci, ctypes.POINTER(i32)]
        L.hss_grid_tuned_shape.argtypes = [ci, ci, ci, ci] + [ctypes.POINTER(i32)] * 3
        L.hss_grid_set_shape.argtypes = [ci, ci, ci, ci, i32, i32, i32]
        L.hss_polish.argtypes = [tr, mp, ci, vp, vp, i32, vp, vp, vp]
        L.hss_select_traces.argtypes = [tr, ci, i32, vp, f32, f32,
                                        ctypes.POINTER(hss_oct_candidates), vp, vp]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise HssError(f"hss error {rc}: {lib().hss_status_string(rc).decode()}")


def _stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _on(stream):
    """Context in which a call's temporaries are allocated (and zeroed) on the stream the
    call launches on, so the caching allocator orders their reuse after the kernel."""
    import contextlib

    import torch
    return torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def pow2_scale(samples):
    """Per-trace power-of-two scale so that |samples/scale| <= 1 (exact to undo)."""
    amax = np.abs(samples).max(axis=1).astype(np.float64)
    amax[amax == 0] = 1.0
    e = np.ceil(np.log2(amax))
    return np.exp2(e).astype(np.float32)


def encode_samples(samples, storage):
    """Returns (stored array as numpy, per-trace scale or None, decoded float32 copy).

    F16 stores samples/scale with a power-of-two per-trace scale (FP16's range); the
    decoded copy is float(stored) * scale computed in float32 -- the values the kernels
    read.  BF16 has FP32's exponent range, where a power-of-two scale changes no rounded
    bit, so it stores the samples themselves with no scale (and the kernels sum its windows
    as FP32 windows)."""
    samples = np.ascontiguousarray(samples, dtype=np.float32)
    if storage == F32:
        return samples, None, samples
    if storage == F16:
        scale = pow2_scale(samples)
        norm = (samples / scale[:, None]).astype(np.float32)
        stored = norm.astype(np.float16)
        dec = stored.astype(np.float32) * scale[:, None]
        return stored, scale, dec.astype(np.float32)
    if storage == BF16:
        u = samples.view(np.uint32)
        # round to nearest even on the upper 16 bits
        rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
        dec = (rounded.astype(np.uint32) << 16).view(np.float32)
        return rounded, None, dec.astype(np.float32)
    raise ValueError(storage)


class DeviceLine:
    """All traces of a line resident on the device (the hss_traces of the C ABI)."""

    def __init__(self, samples, mid, hoff, dt, tau, storage=F32, device="cuda"):
        import torch
        samples = np.asarray(samples, dtype=np.float32)
        self.ntr, self.ns = samples.shape
        self.dt = np.float32(dt)
        self.tau = int(tau)
        self.w = 2 * self.tau + 1
        self.storage = storage
        stored, scale, dec = encode_samples(samples, storage)
        self.decoded = dec
        mid = np.ascontiguousarray(mid, dtype=np.float32)
        hoff = np.ascontiguousarray(hoff, dtype=np.float32)
        self.mid_h, self.hoff_h = mid, hoff
        self.hoff2_h = (hoff * hoff).astype(np.float32)
        if storage == F32:
            self.samples = torch.from_numpy(stored).to(device)
        elif storage == F16:
            self.samples = torch.from_numpy(stored).to(device)
        else:
            self.samples = torch.from_numpy(stored.view(np.int16)).to(device)
        self.scale = torch.from_numpy(scale).to(device) if scale is not None else None
        self.mid = torch.from_numpy(mid).to(device)
        self.hoff = torch.from_numpy(hoff).to(device)
        self.hoff2 = torch.from_numpy(self.hoff2_h).to(device)
        self.struct = hss_traces(
            self.samples.data_ptr(), self.scale.data_ptr() if self.scale is not None else 0,
            self.mid.data_ptr(), self.hoff.data_ptr(), self.hoff2.data_ptr(),
            self.ntr, self.ns, float(self.dt), self.tau, self.w, storage)

    def sample_bytes(self):
        return {F32: 4, F16: 2, BF16: 2}[self.storage]


class DeviceMidpoints:
    """Midpoint batch: trace ranges (and optional index list) into a DeviceLine.
    `order` is the launch order (a permutation of range(n); "balanced" for
    xcd_balanced_order(count)); it changes the schedule only, never a result."""

    def __init__(self, begin, count, m0, index=None, h0=0.0, apm=0.0, device="cuda", order=None):
        import torch
        self.begin_h = np.ascontiguousarray(begin, dtype=np.int32)
        self.count_h = np.ascontiguousarray(count, dtype=np.int32)
        self.m0_h = np.ascontiguousarray(m0, dtype=np.float32)
        self.index_h = np.ascontiguousarray(index, dtype=np.int32) if index is not None else None
        self.n = len(self.begin_h)
        if (self.count_h < 0).any() or (self.begin_h < 0).any():
            raise ValueError("negative trace range")
        self.begin = torch.from_numpy(self.begin_h).to(device)
        self.count = torch.from_numpy(self.count_h).to(device)
        self.m0 = torch.from_numpy(self.m0_h).to(device)
        self.index = (torch.from_numpy(np.ascontiguousarray(index, dtype=np.int32)).to(device)
                      if index is not None else None)
        if isinstance(order, str):
            if order != "balanced":
                raise ValueError(f"unknown order {order!r}")
            order = xcd_balanced_order(self.count_h)
        self.order_h = np.ascontiguousarray(order, dtype=np.int32) if order is not None else None
        if self.order_h is not None and self.order_h.shape != (self.n,):
            raise ValueError("order must hold one entry per midpoint")
        self.order = (torch.from_numpy(self.order_h).to(device)
                      if self.order_h is not None and self.n > 0 else None)
        self.struct = hss_midpoints(
            self.n, self.begin.data_ptr(), self.count.data_ptr(), self.m0.data_ptr(),
            self.index.data_ptr() if self.index is not None else 0, float(h0), float(apm),
            self.order.data_ptr() if self.order is not None else 0)


def verify_index_ops(dt, stream=None):
    """Mismatch count of the exhaustive check of the kernels' short sqrt and t/dt
    sequences for this dt on the current device (0 means the kernels use them)."""
    n = ctypes.c_uint64(0)
    check(lib().hss_verify_index_ops(ctypes.c_float(dt), ctypes.byref(n), _stream_handle(stream)))
    return int(n.value)


def index_ops_state(dt):
    """1 when the kernels use the verified short sqrt/division sequences for dt on the
    current device, 0 when the check failed, -1 when it has not completed yet."""
    st = ctypes.c_int(0)
    check(lib().hss_index_ops_state(ctypes.c_float(dt), ctypes.byref(st)))
    return int(st.value)


def build_parameter_array(model, lower, step, count, device="cuda", stream=None):
    import torch
    np_ = MODEL_NPARAMS[model]
    lo = (ctypes.c_float * 3)(*[float(x) for x in lower[:np_]] + [0.0] * (3 - np_))
    st = (ctypes.c_float * 3)(*[float(x) for x in step[:np_]] + [0.0] * (3 - np_))
    ct = (ctypes.c_int32 * 3)(*[int(x) for x in count[:np_]] + [1] * (3 - np_))
    P = int(np.prod(count[:np_]))
    with _on(stream):
        out = torch.empty(P * np_, dtype=torch.float32, device=device)
    check(lib().hss_build_parameter_array(model, ctypes.cast(lo, ctypes.c_void_p),
                                          ctypes.cast(st, ctypes.c_void_p),
                                          ctypes.cast(ct, ctypes.c_void_p),
                                          _ptr(out), _stream_handle(stream)))
    return out


def check_ranges(line, mids):
    """Every trace a midpoint names must exist in the line (checked on the host before
    any launch)."""
    end = mids.begin_h.astype(np.int64) + mids.count_h
    if mids.index_h is None:
        if mids.n and end.max(initial=0) > line.ntr:
            raise ValueError("midpoint trace range beyond the line")
    else:
        if mids.n and end.max(initial=0) > mids.index_h.size:
            raise ValueError("midpoint range beyond the index list")
        if mids.index_h.size and (mids.index_h.min() < 0 or mids.index_h.max() >= line.ntr):
            raise ValueError("trace index out of range")


def grid_semblance(line, mids, model, params, P, variant=FAST, full=False,
                   not_used=False, result=None, stream=None):
    """Returns (result [n,(2+np),ns], sembl [n,ns,P] or None, stack or None, not_used or None)."""
    import torch
    check_ranges(line, mids)
    if params.numel() < P * MODEL_NPARAMS[model]:
        raise ValueError("parameter array shorter than P candidates")
    dev = line.samples.device
    np_ = MODEL_NPARAMS[model]
    sembl = stack = nu = None
    with _on(stream):
        if result is None:
            result = torch.empty((mids.n, 2 + np_, line.ns), dtype=torch.float32, device=dev)
        if full:
            sembl = torch.empty((mids.n, line.ns, P), dtype=torch.float32, device=dev)
            stack = torch.empty((mids.n, line.ns, P), dtype=torch.float32, device=dev)
        if not_used:
            nu = torch.zeros(mids.n, dtype=torch.int64, device=dev)
    check(lib().hss_grid_semblance(ctypes.byref(line.struct), ctypes.byref(mids.struct), model,
                                   _ptr(params), int(P), int(variant), _ptr(result),
                                   _ptr(sembl), _ptr(stack), _ptr(nu), _stream_handle(stream)))
    return result, sembl, stack, nu


def grid_semblance_plain(line, model, params, P, begins, counts, m0s, h0=0.0, apm=0.0,
                         result=None, scratch=None, stream=None):
    """The plain per-gather baseline (one launch pair per midpoint, HSS_EXACT arithmetic).
    begins/counts/m0s are host arrays; returns result [n][(2+np)][ns]."""
    import numpy as np
    import torch
    np_ = {0: 1, 1: 3, 2: 2}[model]
    b = np.ascontiguousarray(begins, dtype=np.int32)
    c = np.ascontiguousarray(counts, dtype=np.int32)
    m = np.ascontiguousarray(m0s, dtype=np.float32)
    n = b.size
    with _on(stream):
        if result is None:
            result = torch.empty((n, 2 + np_, line.ns), dtype=torch.float32, device="cuda")
        if scratch is None:
            scratch = (torch.empty((line.ns, P), dtype=torch.float32, device="cuda"),
                       torch.empty((line.ns, P), dtype=torch.float32, device="cuda"))
    check(lib().hss_grid_semblance_plain(
        ctypes.byref(line.struct), model, _ptr(params), P, n,
        b.ctypes.data_as(ctypes.c_void_p), c.ctypes.data_as(ctypes.c_void_p),
        m.ctypes.data_as(ctypes.c_void_p), ctypes.c_float(h0), ctypes.c_float(apm),
        _ptr(scratch[0]), _ptr(scratch[1]), _ptr(result), _stream_handle(stream)))
    return result


def de_workspace_bytes(n_mid, ns, model, NP):
    out = ctypes.c_uint64(0)
    check(lib().hss_de_workspace_size(int(n_mid), int(ns), int(model), int(NP), ctypes.byref(out)))
    return int(out.value)


def de_semblance(line, mids, model, lower, upper, NP, G, seed, stream_base=0, variant=FAST,
                 population=False, result=None, workspace=None, stream=None, gen_range=None,
                 not_used=None):
    """Returns (result [n,(2+np),ns], pop_x [n,ns,NP,np] or None, pop_fx [n,ns,NP,2] or None).
    gen_range=(g0, g1) runs generations [g0, g1) only, resuming from `workspace`."""
    import torch
    check_ranges(line, mids)
    dev = line.samples.device
    np_ = MODEL_NPARAMS[model]
    lo = (ctypes.c_float * 3)(*[float(x) for x in list(lower)[:np_]] + [0.0] * (3 - np_))
    hi = (ctypes.c_float * 3)(*[float(x) for x in list(upper)[:np_]] + [0.0] * (3 - np_))
    nbytes = de_workspace_bytes(mids.n, line.ns, model, NP)
    g0, g1 = gen_range if gen_range is not None else (0, G)
    px = pfx = None
    with _on(stream):
        if workspace is None:
            workspace = torch.empty(max(nbytes, 4) // 4, dtype=torch.float32, device=dev)
        if result is None:
            result = torch.empty((mids.n, 2 + np_, line.ns), dtype=torch.float32, device=dev)
        if population:
            px = torch.empty((mids.n, line.ns, NP, np_), dtype=torch.float32, device=dev)
            pfx = torch.empty((mids.n, line.ns, NP, 2), dtype=torch.float32, device=dev)
    check(lib().hss_de_semblance(ctypes.byref(line.struct), ctypes.byref(mids.struct), model,
                                 ctypes.cast(lo, ctypes.c_void_p), ctypes.cast(hi, ctypes.c_void_p),
                                 int(NP), int(G), int(g0), int(g1), ctypes.c_uint64(seed),
                                 ctypes.c_uint32(stream_base), int(variant), _ptr(workspace),
                                 ctypes.c_uint64(workspace.numel() * 4), _ptr(result), _ptr(px),
                                 _ptr(pfx), _ptr(not_used), _stream_handle(stream)))
    return result, px, pfx


DE_SCRATCH_BYTES = 64 << 20   # HSS_DE_SCRATCH_BYTES


def de_semblance_noalloc(line, mids, model, lower, upper, NP, G, seed, stream_base=0,
                         variant=FAST, result=None, stream=None, gen_range=None, not_used=None):
    """hss_de_semblance_noalloc: DE with the populations in the library's static device
    scratch (no workspace argument).  `result` must be given by a caller that cannot
    allocate; here it is allocated when None.  Returns result [n,(2+np),ns]."""
    import torch
    check_ranges(line, mids)
    np_ = MODEL_NPARAMS[model]
    lo = (ctypes.c_float * 3)(*[float(x) for x in list(lower)[:np_]] + [0.0] * (3 - np_))
    hi = (ctypes.c_float * 3)(*[float(x) for x in list(upper)[:np_]] + [0.0] * (3 - np_))
    g0, g1 = gen_range if gen_range is not None else (0, G)
    if result is None:
        with _on(stream):
            result = torch.empty((mids.n, 2 + np_, line.ns), dtype=torch.float32,
                                 device=line.samples.device)
    check(lib().hss_de_semblance_noalloc(
        ctypes.byref(line.struct), ctypes.byref(mids.struct), model,
        ctypes.cast(lo, ctypes.c_void_p), ctypes.cast(hi, ctypes.c_void_p), int(NP), int(G),
        int(g0), int(g1), ctypes.c_uint64(seed), ctypes.c_uint32(stream_base), int(variant),
        _ptr(result), None, None, _ptr(not_used), _stream_handle(stream)))
    return result


def stretch_free(line, mids, model, vparams, nvals, variant=FAST, full=False, not_used=False,
                 stream=None):
    """vparams: device tensor [n, np, ns] (velocity, then A,B or slope).
    Returns (result [n,3,ns], sembl [n,ns,nN] or None, stack or None, not_used or None)."""
    import torch
    check_ranges(line, mids)
    dev = line.samples.device
    np_ = MODEL_NPARAMS[model]
    if vparams.numel() < mids.n * np_ * line.ns:
        raise ValueError("vparams shorter than [n, np, ns]")
    nv = np.ascontiguousarray(nvals, dtype=np.float32)
    nN = nv.size
    sembl = stack = nu = None
    with _on(stream):
        result = torch.empty((mids.n, 3, line.ns), dtype=torch.float32, device=dev)
        if full:
            sembl = torch.empty((mids.n, line.ns, nN), dtype=torch.float32, device=dev)
            stack = torch.empty((mids.n, line.ns, nN), dtype=torch.float32, device=dev)
        if not_used:
            nu = torch.zeros(mids.n, dtype=torch.int64, device=dev)
    check(lib().hss_stretch_free(ctypes.byref(line.struct), ctypes.byref(mids.struct), model,
                                 _ptr(vparams), nv.ctypes.data_as(ctypes.c_void_p), int(nN),
                                 int(variant), _ptr(result), _ptr(sembl), _ptr(stack), _ptr(nu),
                                 _stream_handle(stream)))
    return result, sembl, stack, nu


def select_traces(line, model, m0, apm=0.0, h0=0.0, oct_mode=None, pairs=None, vparams=None,
                  NP=0, lower=(0.0, 0.0), upper=(0.0, 0.0), seed=0, stream_base=0, stream=None):
    """Trace masks [n, ntr] (uint8 tensor on the device) over the whole line."""
    import torch
    dev = line.samples.device
    m0 = np.ascontiguousarray(m0, dtype=np.float32)
    with _on(stream):
        dm0 = torch.from_numpy(m0).to(dev)
        mask = torch.zeros((m0.size, line.ntr), dtype=torch.uint8, device=dev)
    octc = None
    if model == 2:
        octc = hss_oct_candidates()
        octc.mode = oct_mode
        if oct_mode == OCT_PAIRS:
            octc.d_pairs = pairs.data_ptr()
            octc.npairs = pairs.numel() // 2
        elif oct_mode == OCT_PER_SAMPLE:
            if vparams.numel() < m0.size * 2 * line.ns:
                raise ValueError("vparams shorter than [n, 2, ns]")
            octc.d_vparams = vparams.data_ptr()
        else:
            octc.NP = int(NP)
            octc.lower[0], octc.lower[1] = float(lower[0]), float(lower[1])
            octc.upper[0], octc.upper[1] = float(upper[0]), float(upper[1])
            octc.seed = int(seed)
            octc.stream_base = int(stream_base)
    check(lib().hss_select_traces(ctypes.byref(line.struct), int(model), int(m0.size), _ptr(dm0),
                                  float(apm), float(h0),
                                  ctypes.byref(octc) if octc is not None else None,
                                  _ptr(mask), _stream_handle(stream)))
    return mask


def semblance_gradient(line, mids, model, points, out=None, stream=None):
    """S, stack and dS/d(parameters) at the points in rows 2.. of `points`
    [n][(2+np)][ns] (v space).  Returns out [n][(2+np)][ns]: S, stack, gradient rows."""
    import torch
    if out is None:
        with _on(stream):
            out = torch.empty_like(points)
    check(lib().hss_semblance_gradient(ctypes.byref(line.struct), ctypes.byref(mids.struct),
                                       int(model), _ptr(points), _ptr(out),
                                       _stream_handle(stream)))
    return out


def polish(line, mids, model, lower, upper, steps, start, result=None, stream=None):
    """Gauss-Newton polish of the points in rows 2.. of `start` [n][(2+np)][ns] within
    [lower, upper] (v space).  Returns result [n][(2+np)][ns] (S, stack, parameters)."""
    import numpy as np
    import torch
    lo = np.ascontiguousarray(lower, dtype=np.float32)
    hi = np.ascontiguousarray(upper, dtype=np.float32)
    if result is None:
        with _on(stream):
            result = torch.empty_like(start)
    check(lib().hss_polish(ctypes.byref(line.struct), ctypes.byref(mids.struct), int(model),
                           lo.ctypes.data_as(ctypes.c_void_p), hi.ctypes.data_as(ctypes.c_void_p),
                           int(steps), _ptr(start), _ptr(result), _stream_handle(stream)))
    return result


def grid_set_config(model, w, variant, storage, cpl):
    """Pins the grid launch shape (candidates per lane) for a key; 0 forgets it."""
    check(lib().hss_grid_set_config(int(model), int(w), int(variant), int(storage), int(cpl)))


def grid_set_shape(model, w, variant, storage, cpl, t0_tiles, staged=0):
    """Pins the grid launch shape (candidates per lane, t0 tiles per workgroup, LDS-staged
    traces); cpl 0 forgets the key."""
    check(lib().hss_grid_set_shape(int(model), int(w), int(variant), int(storage), int(cpl),
                                   int(t0_tiles), int(staged)))


def grid_tuned_shape(model, w, variant, storage):
    """(candidates per lane, t0 tiles per workgroup, staged) chosen for a key ((0, 0, 0):
    not tuned)."""
    c, t, st = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
    check(lib().hss_grid_tuned_shape(int(model), int(w), int(variant), int(storage),
                                     ctypes.byref(c), ctypes.byref(t), ctypes.byref(st)))
    return int(c.value), int(t.value), int(st.value)


def grid_tuned_config(model, w, variant, storage):
    """Candidates per lane chosen for a key (0: not tuned yet)."""
    c = ctypes.c_int32(0)
    check(lib().hss_grid_tuned_config(int(model), int(w), int(variant), int(storage),
                                      ctypes.byref(c)))
    return int(c.value)


def semblance_function():
    """torch.autograd.Function over the adjoint kernel: S = f(points), where points is a
    [n, np, ns] tensor of parameters (v space) for every (midpoint, sample) of a batch.
    forward returns S [n, ns]; backward multiplies the incoming gradient by dS/dpoints
    from hss_semblance_gradient (computed in the same launch as S).  The line and the
    midpoint batch ride along as non-differentiable arguments."""
    import torch

    class SemblanceFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, points, line, mids, model):
            n, np_, ns = points.shape
            rows = torch.zeros((n, 2 + np_, ns), dtype=torch.float32, device=points.device)
            rows[:, 2:, :] = points.detach()
            out = semblance_gradient(line, mids, model, rows)
            ctx.save_for_backward(out)
            ctx.shape = (n, np_, ns)
            return out[:, 0, :].clone()

        @staticmethod
        def backward(ctx, grad_s):
            (rows,) = ctx.saved_tensors
            n, np_, ns = ctx.shape
            g = grad_s.contiguous().to(torch.float32)
            out = torch.empty((n, np_, ns), dtype=torch.float32, device=rows.device)
            check(lib().hss_semblance_vjp(n, np_, ns, _ptr(g), _ptr(rows), _ptr(out),
                                          _stream_handle(None)))
            return out, None, None, None

    return SemblanceFunction


def xcd_balanced_order(work, bins=8):
    """Launch order of a batch's midpoints (the C++ host layer's xcd_balanced_order): sort
    by work, deal to `bins` bins in snake order, concatenate -- every XCD then receives
    the same number of midpoints with about the same total work, largest first."""
    import numpy as np
    work = np.asarray(work)
    idx = np.argsort(-work, kind="stable")
    out = [[] for _ in range(bins)]
    full = len(idx) // bins  # the last partial round fills bins 0.. (larger bins first)
    for k, i in enumerate(idx):
        rnd, pos = divmod(k, bins)
        out[bins - 1 - pos if (rnd & 1) and rnd < full else pos].append(int(i))
    return np.array([i for b in out for i in b], dtype=np.int64)
