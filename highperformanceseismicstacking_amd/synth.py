"""Seeded synthetic prestack data for the semblance workloads, and SU file I/O.

The reference ships no datasets (its sample files are absent), so tests and the
benchmark generate lines here: Ricker wavelets on CMP hyperbolae, ZO-CRS surfaces or
dipping OCT events, plus Gaussian noise, with integer source/receiver coordinates so
that midpoints and half offsets come out exactly as the reference's SU reader computes
them (common/src/model/Utils.cpp:8-52: f = scalco >= 0 ? scalco : -1/scalco,
m = f*(gx+sx)*0.5, h = f*(gx-sx)*0.5, projected on the azimuth).
"""
from dataclasses import dataclass, field

import numpy as np

f32 = np.float32

# 240-byte SU trace header (trace_info_t, common/include/model/Trace.hpp:9-91): the
# fields this code reads or writes, at their byte offsets; the rest stays zero.
SU_HEADER_DTYPE = np.dtype({
    "names": ["tracl", "tracr", "cdp", "offset", "scalco", "sx", "sy", "gx", "gy", "ns", "dt"],
    "formats": ["<i4", "<i4", "<i4", "<i4", "<i2", "<i4", "<i4", "<i4", "<i4", "<u2", "<u2"],
    "offsets": [0, 4, 20, 36, 70, 72, 76, 80, 84, 114, 116],
    "itemsize": 240,
})


def dt_seconds(dt_us):
    """dtInSeconds = (float)dt * 1e-6f (common/src/model/Gather.cpp:73)."""
    return f32(f32(dt_us) * f32(1e-6))


def tau_index(tau_s, dt_us):
    """tauIndexDisplacement = (int)(tau / dtInSeconds) in float (Gather.cpp:75)."""
    return int(np.trunc(f32(f32(tau_s) / dt_seconds(dt_us))))


def azimuth_rad(az_deg):
    """degreesToRad in float (common/src/model/Utils.cpp:54-56)."""
    return f32(f32(f32(az_deg) * f32(np.pi)) / f32(180.0))


def geometry_from_header(sx, gx, scalco, az_deg=90.0, sy=None, gy=None):
    """Midpoint and signed half offset in float32, op for op as Utils.cpp:8-52."""
    sx = np.asarray(sx, dtype=np.int64)
    gx = np.asarray(gx, dtype=np.int64)
    sy = np.zeros_like(sx) if sy is None else np.asarray(sy, dtype=np.int64)
    gy = np.zeros_like(gx) if gy is None else np.asarray(gy, dtype=np.int64)
    sc = np.asarray(scalco, dtype=np.int64) * np.ones_like(sx)
    scf = sc.astype(np.float32)
    with np.errstate(divide="ignore"):
        fsc = np.where(sc >= 0, scf, f32(-1.0) / scf).astype(np.float32)
    half = f32(0.5)
    mx = (fsc * (gx + sx).astype(np.float32)) * half
    my = (fsc * (gy + sy).astype(np.float32)) * half
    hx = (fsc * (gx - sx).astype(np.float32)) * half
    hy = (fsc * (gy - sy).astype(np.float32)) * half
    az = azimuth_rad(az_deg)
    sa, ca = f32(np.sin(az, dtype=np.float32)), f32(np.cos(az, dtype=np.float32))
    m = (mx * sa + my * ca).astype(np.float32)
    h = (hx * sa + hy * ca).astype(np.float32)
    return m, h


def ricker(t, f):
    a = (np.pi * f * t) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


@dataclass
class Line:
    samples: np.ndarray          # [ntr, ns] float32
    sx: np.ndarray               # [ntr] int32 (decimetres with scalco=-10)
    gx: np.ndarray
    scalco: int
    ns: int
    dt_us: int
    cdp: np.ndarray              # [ntr] CMP ordinal
    mid: np.ndarray = field(default=None)
    hoff: np.ndarray = field(default=None)

    def __post_init__(self):
        if self.mid is None:
            self.mid, self.hoff = geometry_from_header(self.sx, self.gx, self.scalco)

    @property
    def dt(self):
        return dt_seconds(self.dt_us)

    @property
    def ntr(self):
        return self.samples.shape[0]

    def cmp_ranges(self):
        """(begin, count, m0) of each distinct midpoint; traces are sorted by (m, h)."""
        m = self.mid
        starts = np.flatnonzero(np.r_[True, m[1:] != m[:-1]])
        counts = np.diff(np.r_[starts, len(m)])
        return starts.astype(np.int32), counts.astype(np.int32), m[starts].astype(np.float32)

    def subset(self, begin, count):
        sl = slice(begin, begin + count)
        return Line(self.samples[sl].copy(), self.sx[sl].copy(), self.gx[sl].copy(), self.scalco,
                    self.ns, self.dt_us, self.cdp[sl].copy(), self.mid[sl].copy(),
                    self.hoff[sl].copy())


def _add_events(samples, t_event, amp, f_hz, dt):
    """Adds a Ricker of peak frequency f_hz at time t_event[i] to trace i."""
    ntr, ns = samples.shape
    half = int(np.ceil(1.5 / (f_hz * dt))) + 1
    finite = np.isfinite(t_event)
    k0 = np.round(np.where(finite, t_event, -1e9) / dt).astype(np.int64)
    offs = np.arange(-half, half + 1)
    idx = k0[:, None] + offs[None, :]
    valid = (idx >= 0) & (idx < ns) & finite[:, None]
    tt = idx * dt - t_event[:, None]
    vals = (amp[:, None] * ricker(tt, f_hz)).astype(np.float32)
    rows = np.broadcast_to(np.arange(ntr)[:, None], idx.shape)
    samples[rows[valid], idx[valid]] += vals[valid]


def make_line(n_cmp, fold, offset_max, ns, dt_us, seed, *, offset_min=0.0, offsets=None,
              n_events=5, f_hz=(25.0, 25.0), v_range=(1500.0, 3500.0), noise=0.05,
              cmp_spacing=12.5, m_first=1000.0, kind="cmp", zocrs=None, oct_slope=None):
    """A line of CMP gathers with seeded Ricker events and Gaussian noise.

    fold: int, or (lo, hi) for a per-CMP fold drawn uniformly from [lo, hi].
    kind: "cmp" hyperbolae t^2 = t0^2 + 4h^2/v^2 (t0 drifting slowly along the line);
          "zocrs" curved dipping events t^2 = (t0 + A dm)^2 + B dm^2 + C h^2 around the
          line centre, with (A, B, C) drawn from the `zocrs` bounds;
          "oct" dipping hyperbolae t^2 = (t0 + p dm)^2 + 4h^2/v^2, p from `oct_slope`."""
    rng = np.random.default_rng(seed)
    dt = float(dt_seconds(dt_us))
    T = (ns - 1) * dt
    if np.isscalar(fold):
        folds = np.full(n_cmp, int(fold))
    else:
        folds = rng.integers(fold[0], fold[1] + 1, size=n_cmp)
    scalco = -10
    M = np.round((m_first + cmp_spacing * np.arange(n_cmp)) * 10).astype(np.int64)  # decimetres
    sx, gx, cdp = [], [], []
    for g in range(n_cmp):
        if offsets is not None:
            off = np.asarray(offsets, dtype=np.float64)
        else:
            off = np.linspace(offset_min, offset_max, folds[g])
        Hd = np.round(off * 10 / 2).astype(np.int64)  # half offset in decimetres
        sx.append(M[g] - Hd)
        gx.append(M[g] + Hd)
        cdp.append(np.full(len(Hd), g))
    sx = np.concatenate(sx).astype(np.int32)
    gx = np.concatenate(gx).astype(np.int32)
    cdp = np.concatenate(cdp).astype(np.int32)
    mid, hoff = geometry_from_header(sx, gx, scalco)
    ntr = len(sx)
    samples = np.zeros((ntr, ns), dtype=np.float32)
    m64 = mid.astype(np.float64)
    h64 = hoff.astype(np.float64)
    mc = 0.5 * (m64.min() + m64.max())
    dm = m64 - mc
    for e in range(n_events):
        t0 = rng.uniform(0.1 * T, 0.9 * T)
        v = rng.uniform(*v_range)
        amp = rng.uniform(0.5, 1.0) * rng.choice([-1.0, 1.0])
        fq = rng.uniform(*f_hz)
        if kind == "cmp":
            drift = rng.uniform(-5e-5, 5e-5)
            t0e = t0 + drift * dm
            arg = t0e ** 2 + 4.0 * h64 ** 2 / v ** 2
        elif kind == "zocrs":
            (alo, ahi), (blo, bhi), (clo, chi) = zocrs
            A = rng.uniform(alo, ahi)
            B = rng.uniform(blo, bhi)
            C = rng.uniform(clo, chi)
            arg = (t0 + A * dm) ** 2 + B * dm ** 2 + C * h64 ** 2
        elif kind == "oct":
            p = rng.uniform(*oct_slope)
            arg = (t0 + p * dm) ** 2 + 4.0 * h64 ** 2 / v ** 2
        else:
            raise ValueError(kind)
        with np.errstate(invalid="ignore"):
            te = np.where(arg >= 0, np.sqrt(np.maximum(arg, 0.0)), np.nan)
        _add_events(samples, te, np.full(ntr, amp), fq, dt)
    if noise > 0:
        samples += rng.normal(0.0, noise, size=samples.shape).astype(np.float32)
    return Line(samples, sx, gx, scalco, ns, int(dt_us), cdp, mid, hoff)


# ----------------------------------------------------------------------------------
# SU I/O: raw native-endian traces, 240-byte header + ns float32 (Trace.cpp:45-61).

def write_su(path, line, extra_headers=None):
    hdr = np.zeros(line.ntr, dtype=SU_HEADER_DTYPE)
    hdr["tracl"] = np.arange(1, line.ntr + 1)
    hdr["tracr"] = hdr["tracl"]
    hdr["cdp"] = line.cdp
    hdr["offset"] = (line.gx.astype(np.int64) - line.sx.astype(np.int64))
    hdr["scalco"] = line.scalco
    hdr["sx"] = line.sx
    hdr["gx"] = line.gx
    hdr["ns"] = line.ns
    hdr["dt"] = line.dt_us
    rec = np.zeros(line.ntr, dtype=np.dtype([("h", "V240"), ("d", "<f4", (line.ns,))]))
    rec["h"] = hdr.view("V240")
    rec["d"] = line.samples
    rec.tofile(path)


def read_su(path, ns=None):
    """Returns (headers structured array, samples [ntr, ns] float32)."""
    raw = np.fromfile(path, dtype=np.uint8)
    if ns is None:
        ns = int(raw[114:116].view("<u2")[0])
    rec_dtype = np.dtype([("h", "V240"), ("d", "<f4", (ns,))])
    n = raw.size // rec_dtype.itemsize
    rec = raw[: n * rec_dtype.itemsize].view(rec_dtype)
    hdr = rec["h"].copy().view(SU_HEADER_DTYPE)
    return hdr, np.ascontiguousarray(rec["d"])


# ----------------------------------------------------------------------------------
# The five workloads (sizes of the benchmark configurations).

def workload(name, seed=20240601):
    """Returns (Line, dict of search settings) for 'c1'..'c5'."""
    if name == "c1":
        line = make_line(1, 60, 2950.0, 1001, 4000, seed, offsets=np.arange(0, 3000, 50.0),
                         m_first=1000.0)
        return line, dict(model="cmp", tau=0.022, lower=[1500.0], upper=[4500.0],
                          granularity=[256], aph=1500.0, apm=0.0)
    if name == "c2":
        line = make_line(500, (48, 96), 4000.0, 1501, 2000, seed, offset_min=50.0,
                         v_range=(1500.0, 3500.0))
        return line, dict(model="cmp", tau=0.011, lower=[1400.0], upper=[5000.0],
                          granularity=[1024], aph=2000.0, apm=0.0)
    if name == "c5":
        line = make_line(2000, (24, 120), 4000.0, 2001, 2000, seed, offset_min=50.0,
                         f_hz=(15.0, 40.0), noise=0.1)
        return line, dict(model="cmp", tau=0.014, lower=[1400.0], upper=[5000.0],
                          granularity=[2048], aph=2000.0, apm=0.0)
    if name == "c3":
        line = make_line(200, 24, 2000.0, 1001, 4000, seed, offset_min=50.0, kind="zocrs",
                         zocrs=((-8e-4, 8e-4), (-8e-7, 8e-7), (2e-7, 1.5e-6)))
        return line, dict(model="zocrs", tau=0.022, lower=[1581.139], upper=[5000.0],
                          ain=90.0, v0=2000.0, bpctg=0.625, aph=1000.0, apm=250.0,
                          population=32, generations=50)
    if name == "c4":
        line = make_line(300, 16, 3000.0, 1251, 4000, seed, offset_min=200.0, kind="oct",
                         oct_slope=(-4e-4, 4e-4))
        return line, dict(model="oct", tau=0.022, lower=[1500.0, -5e-4], upper=[4500.0, 5e-4],
                          h0=0.0, aph=1500.0, apm=200.0, population=48, generations=40)
    raise ValueError(name)
