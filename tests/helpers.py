"""Shared set-up for the parity tests: grids, midpoint selections and comparisons."""
import numpy as np

import oracle as O
from highperformanceseismicstacking_amd import synth

f32 = np.float32

# Tolerances written where they are used (FP32 semblance/stack agreement for the
# contracted variant; ties for the best-candidate index).
S_RTOL = 1e-5
TIE_TOL = 1e-6


def grid_steps(lower, upper, granularity):
    """discretizationStep = (hi - lo) / G in float (LinearSearchAlgorithm.cpp:139)."""
    return [float(f32(f32(f32(hi) - f32(lo)) / f32(g))) for lo, hi, g in
            zip(lower, upper, granularity)]


def zocrs_bounds(lower_v, upper_v, ain_deg, v0, bpctg):
    """ZO-CRS parameter bounds (TraveltimeBuilder.cpp:45-76): v from the CLI, then
    A in +-2 sin(ain)/v0 and B in +-bpctg*4/vmin^2, all in float."""
    ain = f32(f32(f32(ain_deg) * f32(np.pi)) / f32(180.0))
    ua = f32(f32(f32(2.0) * f32(np.sin(ain, dtype=np.float32))) / f32(v0))
    ub = f32(f32(f32(bpctg) * f32(4.0)) / f32(f32(lower_v) * f32(lower_v)))
    return [float(lower_v), float(-ua), float(-ub)], [float(upper_v), float(ua), float(ub)]


def zocrs_window(line, m0, apm):
    """Traces with |m0 - m| <= apm: a contiguous range of the (m, h)-sorted line
    (zocrs/common.cu:14)."""
    sel = np.flatnonzero(np.abs(f32(m0) - line.mid) <= f32(apm))
    if sel.size == 0:
        return 0, 0
    assert np.all(np.diff(sel) == 1)
    return int(sel[0]), int(sel.size)


def oracle_gather(samples, line, begin, count, tau, m0, h0=0.0, apm=0.0):
    sl = slice(begin, begin + count)
    return O.Gather(samples[sl], line.mid[sl], line.hoff[sl], line.dt, tau, m0=m0, h0=h0,
                    apm=apm)


def compare_close(gpu_S, gpu_stack, ref_S, ref_stack):
    """Semblance within 1e-5 relative (1e-6 absolute floor; S is in [0, 1]) and stack
    within 1e-5 relative to the largest stack magnitude of the array (a stack near zero
    is a cancellation of O(1) terms)."""
    okS = np.abs(gpu_S - ref_S) <= S_RTOL * np.abs(ref_S) + 1e-6
    floor = S_RTOL * max(float(np.nanmax(np.abs(ref_stack))), 1e-30)
    okT = np.abs(gpu_stack - ref_stack) <= S_RTOL * np.abs(ref_stack) + floor
    return okS, okT


def argmax_mismatches(gpu_idx, ref_idx, ref_sembl_rows):
    """Samples where the picked candidate differs and is not a tie within 1e-6 in the
    oracle's own semblance values."""
    bad = []
    for s in np.flatnonzero(gpu_idx != ref_idx):
        row = ref_sembl_rows[s]
        gi, ri = int(gpu_idx[s]), int(ref_idx[s])
        if gi < 0 or ri < 0 or abs(float(row[gi]) - float(row[ri])) > TIE_TOL:
            bad.append(int(s))
    return bad


def best_index_from_result(result_v, params_c):
    """Candidate index whose 2/sqrt(c) equals the reported velocity (grid output)."""
    v_of = (f32(2.0) / np.sqrt(params_c.astype(np.float32))).astype(np.float32)
    return v_of, result_v
