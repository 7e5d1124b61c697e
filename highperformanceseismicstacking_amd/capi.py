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
    ]


def lib():
    """The loaded libhss.so (raises HssError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HssError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_float
        L.hss_abi_version.restype = ctypes.c_int
        L.hss_status_string.restype = ctypes.c_char_p
        L.hss_status_string.argtypes = [ctypes.c_int]
        L.hss_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.hss_build_parameter_array.argtypes = [ctypes.c_int, vp, vp, vp, vp, vp]
        L.hss_grid_semblance.argtypes = [
            ctypes.POINTER(hss_traces), ctypes.POINTER(hss_midpoints), ctypes.c_int,
            vp, i32, ctypes.c_int, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise HssError(f"hss error {rc}: {lib().hss_status_string(rc).decode()}")


def _stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


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

    F16/BF16 store samples/scale with a power-of-two per-trace scale; the decoded copy
    is float(stored) * scale computed in float32 -- the values the kernels read."""
    samples = np.ascontiguousarray(samples, dtype=np.float32)
    if storage == F32:
        return samples, None, samples
    scale = pow2_scale(samples)
    norm = (samples / scale[:, None]).astype(np.float32)
    if storage == F16:
        stored = norm.astype(np.float16)
        dec = stored.astype(np.float32) * scale[:, None]
        return stored, scale, dec.astype(np.float32)
    if storage == BF16:
        u = norm.view(np.uint32)
        # round to nearest even on the upper 16 bits
        rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
        dec = (rounded.astype(np.uint32) << 16).view(np.float32) * scale[:, None]
        return rounded, scale, dec.astype(np.float32)
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
    """Midpoint batch: trace ranges (and optional index list) into a DeviceLine."""

    def __init__(self, begin, count, m0, index=None, h0=0.0, apm=0.0, device="cuda"):
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
        self.struct = hss_midpoints(
            self.n, self.begin.data_ptr(), self.count.data_ptr(), self.m0.data_ptr(),
            self.index.data_ptr() if self.index is not None else 0, float(h0), float(apm))


def build_parameter_array(model, lower, step, count, device="cuda", stream=None):
    import torch
    np_ = MODEL_NPARAMS[model]
    lo = (ctypes.c_float * 3)(*[float(x) for x in lower[:np_]] + [0.0] * (3 - np_))
    st = (ctypes.c_float * 3)(*[float(x) for x in step[:np_]] + [0.0] * (3 - np_))
    ct = (ctypes.c_int32 * 3)(*[int(x) for x in count[:np_]] + [1] * (3 - np_))
    P = int(np.prod(count[:np_]))
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
    if result is None:
        result = torch.empty((mids.n, 2 + np_, line.ns), dtype=torch.float32, device=dev)
    sembl = stack = nu = None
    if full:
        sembl = torch.empty((mids.n, line.ns, P), dtype=torch.float32, device=dev)
        stack = torch.empty((mids.n, line.ns, P), dtype=torch.float32, device=dev)
    if not_used:
        nu = torch.zeros(mids.n, dtype=torch.int64, device=dev)
    check(lib().hss_grid_semblance(ctypes.byref(line.struct), ctypes.byref(mids.struct), model,
                                   _ptr(params), int(P), int(variant), _ptr(result),
                                   _ptr(sembl), _ptr(stack), _ptr(nu), _stream_handle(stream)))
    return result, sembl, stack, nu
