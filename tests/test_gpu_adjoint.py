"""The semblance adjoint and the Gauss-Newton polish on the GPU (libhss.so, C ABI).

The gradient is compared with the FP64 oracle gradient (oracle_grad.c), which is itself
pinned against central finite differences of the all-FP64 semblance (test_oracle.py).
Here the oracle takes the traveltime, sample index, weight and trace decisions in FP32
exactly as the kernel does (f32_index) and the derivatives and sums in FP64, so the
comparison measures the kernel's FP32 derivative and summation error: within 1e-3
relative plus a floor of 1e-3 of the median gradient magnitude (the gradient is a
difference of two sums that cancel near a maximum), S within 1e-5 relative.
"""
import numpy as np
import pytest

import oracle as O
from highperformanceseismicstacking_amd import CMP, ZOCRS, OCT, F32, F16, FAST
from highperformanceseismicstacking_amd import capi, synth
from helpers import grid_steps, oracle_gather, zocrs_bounds, zocrs_window

pytestmark = pytest.mark.gpu

OMODEL = {CMP: O.CMP, ZOCRS: O.ZOCRS, OCT: O.OCT}
NP = {CMP: 1, ZOCRS: 3, OCT: 2}


def problem(model):
    """(line, tau, begins, counts, m0s, h0, apm, lower, upper) for one model."""
    if model == CMP:
        line, cfg = synth.workload("c1")
        b, c, m0 = line.cmp_ranges()
        lo, hi = cfg["lower"], cfg["upper"]
        h0 = apm = 0.0
    else:
        line, cfg = synth.workload("c3" if model == ZOCRS else "c4")
        _, _, m0all = line.cmp_ranges()
        m0 = m0all[[len(m0all) // 3, len(m0all) // 2]]
        ranges = [zocrs_window(line, m, cfg["apm"]) for m in m0]
        b = np.array([r[0] for r in ranges])
        c = np.array([r[1] for r in ranges])
        apm = cfg["apm"]
        h0 = cfg.get("h0", 0.0) if model == OCT else 0.0
        if model == ZOCRS:
            lo, hi = zocrs_bounds(cfg["lower"][0], cfg["upper"][0], cfg["ain"], cfg["v0"],
                                  cfg["bpctg"])
        else:
            lo, hi = cfg["lower"], cfg["upper"]
    tau = synth.tau_index(cfg["tau"], line.dt_us)
    return line, tau, b, c, m0, h0, apm, lo, hi


def random_points(model, n, ns, lo, hi, seed):
    rng = np.random.default_rng(seed)
    np_ = NP[model]
    pts = np.zeros((n, 2 + np_, ns), dtype=np.float32)
    for d in range(np_):
        a, b = lo[d], hi[d]
        if d == 0:
            a, b = a + 0.05 * (b - a), b - 0.05 * (b - a)
        pts[:, 2 + d, :] = rng.uniform(a, b, size=(n, ns)).astype(np.float32)
    return pts


@pytest.mark.parametrize("model", [CMP, ZOCRS, OCT])
@pytest.mark.parametrize("storage", [F32, F16])
def test_gradient_matches_fp64_oracle(model, storage):
    import torch
    line, tau, b, c, m0, h0, apm, lo, hi = problem(model)
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=storage)
    mp = capi.DeviceMidpoints(b, c, m0, h0=h0, apm=apm)
    pts = random_points(model, len(b), line.ns, lo, hi, seed=3 + model)
    out = capi.semblance_gradient(dl, mp, model, torch.from_numpy(pts).cuda())
    torch.cuda.synchronize()
    out = out.cpu().numpy()
    rng = np.random.default_rng(9)
    np_ = NP[model]
    bad = total = 0
    fails = []
    for g in range(len(b)):
        og = oracle_gather(dl.decoded, line, int(b[g]), int(c[g]), tau, float(m0[g]), h0=h0,
                           apm=apm)
        samples = rng.choice(np.arange(20, line.ns - 20), size=60, replace=False)
        ref = [O.eval_grad_f64(OMODEL[model], og, int(s), pts[g][2:, s].astype(np.float64),
                               f32_index=True) for s in samples]
        S64 = np.array([r[0] for r in ref])
        G64 = np.array([r[2] for r in ref])
        S32 = out[g][0, samples]
        G32 = out[g][2:, samples].T  # (samples, np)
        scale = np.median(np.abs(G64), axis=0) + 1e-30
        for i in range(len(samples)):
            total += 1
            okS = abs(S32[i] - S64[i]) <= 1e-5 * abs(S64[i]) + 1e-6
            okG = np.all(np.abs(G32[i] - G64[i]) <= 1e-3 * np.abs(G64[i]) + 1e-3 * scale)
            if not (okS and okG):
                bad += 1
                fails.append((g, int(samples[i]), pts[g][2:, samples[i]].tolist(), float(S32[i]),
                              S64[i], G32[i].tolist(), G64[i].tolist()))
    assert bad == 0, (bad, total, fails[:4])


@pytest.mark.parametrize("storage", [F32, F16])
def test_oct_gradient_over_offset_grouped_lists(storage):
    """OCT lists grouped by half offset (how the host layer passes them to the non-exact
    variants): the adjoint keeps each point's dual traveltime terms across a run of one
    offset (the same bits as computing them per trace, so the gradient over a list is the
    same whichever way the terms come).  Against the FP64 oracle over the same list, to
    the bounds above, at the points of the line-order test (seed 3 + OCT), so the two
    tests compare both trace loops at the same parameters.  (Other random points can fall
    where S is ~1e-3 and the gradient cancels to a few percent in FP32, in either loop.)"""
    import torch
    line, tau, b, c, m0, h0, apm, lo, hi = problem(OCT)
    lists = []
    for g in range(len(b)):
        idx = np.arange(b[g], b[g] + c[g])
        idx = idx[np.argsort(line.hoff[idx], kind="stable")]
        h = line.hoff[idx].view(np.uint32)
        runs = 1 + np.count_nonzero(h[1:] != h[:-1])
        assert 4 * runs <= len(idx), g  # long runs: the kernel keeps the terms
        lists.append(idx)
    counts = np.array([len(x) for x in lists])
    begins = np.r_[0, np.cumsum(counts)[:-1]]
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=storage)
    mp = capi.DeviceMidpoints(begins, counts, m0, index=np.concatenate(lists), h0=h0, apm=apm)
    pts = random_points(OCT, len(b), line.ns, lo, hi, seed=3 + OCT)
    out = capi.semblance_gradient(dl, mp, OCT, torch.from_numpy(pts).cuda())
    torch.cuda.synchronize()
    out = out.cpu().numpy()
    rng = np.random.default_rng(5)
    bad = total = 0
    fails = []
    for g, idx in enumerate(lists):
        og = O.Gather(dl.decoded[idx], line.mid[idx], line.hoff[idx], line.dt, tau,
                      m0=float(m0[g]), h0=h0, apm=apm)
        samples = rng.choice(np.arange(20, line.ns - 20), size=60, replace=False)
        ref = [O.eval_grad_f64(O.OCT, og, int(s), pts[g][2:, s].astype(np.float64),
                               f32_index=True) for s in samples]
        S64 = np.array([r[0] for r in ref])
        G64 = np.array([r[2] for r in ref])
        S32 = out[g][0, samples]
        G32 = out[g][2:, samples].T
        scale = np.median(np.abs(G64), axis=0) + 1e-30
        for i in range(len(samples)):
            total += 1
            okS = abs(S32[i] - S64[i]) <= 1e-5 * abs(S64[i]) + 1e-6
            okG = np.all(np.abs(G32[i] - G64[i]) <= 1e-3 * np.abs(G64[i]) + 1e-3 * scale)
            if not (okS and okG):
                bad += 1
                fails.append((g, int(samples[i]), float(S32[i]), S64[i], G32[i].tolist(),
                              G64[i].tolist()))
    assert total > 0 and bad == 0, (bad, total, fails[:4])


@pytest.mark.parametrize("model", [CMP, ZOCRS, OCT])
def test_polish_never_decreases_and_is_consistent(model):
    """From a coarse grid result: every polished S >= the S at the start point (same
    arithmetic), the polished S equals the gradient kernel's S at the polished point bit
    for bit, the points stay in the box, and the FP64 oracle confirms the gain."""
    import torch
    line, tau, b, c, m0, h0, apm, lo, hi = problem(model)
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau)
    mp = capi.DeviceMidpoints(b, c, m0, h0=h0, apm=apm)
    G = {CMP: [24], ZOCRS: [5, 3, 3], OCT: [8, 3]}[model]
    params = O.build_parameter_array(OMODEL[model], lo, grid_steps(lo, hi, G), G)
    P = params.size // NP[model]
    res, _, _, _ = capi.grid_semblance(dl, mp, model,
                                       torch.from_numpy(np.ascontiguousarray(params)).cuda(), P,
                                       variant=FAST)
    start = capi.semblance_gradient(dl, mp, model, res)
    pol = capi.polish(dl, mp, model, lo, hi, 10, res)
    again = capi.semblance_gradient(dl, mp, model, pol)
    torch.cuda.synchronize()
    res, start, pol, again = (x.cpu().numpy() for x in (res, start, pol, again))
    valid = res[:, 2, :] > 0
    assert np.all(pol[:, 0, :][valid] >= start[:, 0, :][valid])
    assert np.array_equal(pol[:, 0, :].view(np.uint32), again[:, 0, :].view(np.uint32))
    for d in range(NP[model]):
        v = pol[:, 2 + d, :][valid]
        assert np.all(v >= np.float32(lo[d])) and np.all(v <= np.float32(hi[d]))
    gain = pol[:, 0, :][valid] - start[:, 0, :][valid]
    assert np.mean(gain) > 0
    # the FP64 oracle agrees that the polished points are better where the kernel moved
    og = oracle_gather(line.samples, line, int(b[0]), int(c[0]), tau, float(m0[0]), h0=h0, apm=apm)
    moved = np.flatnonzero((np.abs(pol[0][2:] - res[0][2:]).sum(axis=0) > 0) & valid[0])
    rng = np.random.default_rng(1)
    better = 0
    pick = rng.choice(moved, size=min(40, moved.size), replace=False)
    for s in pick:
        S1 = O.eval_grad_f64(OMODEL[model], og, int(s), res[0][2:, s].astype(np.float64))[0]
        S2 = O.eval_grad_f64(OMODEL[model], og, int(s), pol[0][2:, s].astype(np.float64))[0]
        better += S2 > S1
    assert better >= 0.9 * len(pick), (better, len(pick))


def test_autograd_function_drives_a_torch_optimizer():
    """The autograd op: backward equals the gradient kernel's dS/dv, and a few steps of a
    torch optimizer on -sum(S) raise the mean semblance of grid winners' neighbours."""
    import torch
    line, tau, b, c, m0, h0, apm, lo, hi = problem(CMP)
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau)
    mp = capi.DeviceMidpoints(b, c, m0)
    Fn = capi.semblance_function()
    rng = np.random.default_rng(4)
    v0 = rng.uniform(2000.0, 3000.0, size=(1, 1, line.ns)).astype(np.float32)
    p = torch.tensor(v0, device="cuda", requires_grad=True)
    S = Fn.apply(p, dl, mp, CMP)
    S.sum().backward()
    rows = torch.zeros((1, 3, line.ns), device="cuda")
    rows[:, 2:, :] = torch.from_numpy(v0).cuda()
    ref = capi.semblance_gradient(dl, mp, CMP, rows)
    torch.cuda.synchronize()
    assert torch.equal(p.grad, ref[:, 2:, :])
    assert torch.equal(S.detach(), ref[:, 0, :])
    opt = torch.optim.Adam([p], lr=5.0)
    first = None
    for _ in range(30):
        opt.zero_grad()
        S = Fn.apply(p, dl, mp, CMP)
        loss = -S.mean()
        if first is None:
            first = float(S.detach().mean())
        loss.backward()
        opt.step()
        with torch.no_grad():
            p.clamp_(lo[0], hi[0])
    last = float(Fn.apply(p, dl, mp, CMP).detach().mean())
    assert last > first


def test_cpp_semblance_op_forward_backward():
    """The C++ SemblanceOp (libhss_host.so): forward gives the gradient kernel's S and
    backward gives grad_S * dS/dpoints, for the three models."""
    import ctypes
    import os
    import torch
    from highperformanceseismicstacking_amd import HOST_LIB_PATH
    L = ctypes.CDLL(HOST_LIB_PATH)
    L.hss_host_semblance_op.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_int] + [ctypes.c_void_p] * 5
    for model in (CMP, ZOCRS, OCT):
        line, tau, b, c, m0, h0, apm, lo, hi = problem(model)
        dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau)
        mp = capi.DeviceMidpoints(b, c, m0, h0=h0, apm=apm)
        pts = random_points(model, len(b), line.ns, lo, hi, seed=13)
        rows = torch.from_numpy(pts).cuda()
        points = rows[:, 2:, :].contiguous()
        gS = torch.from_numpy(np.random.default_rng(2).normal(size=(len(b), line.ns))
                              .astype(np.float32)).cuda()
        S = torch.empty((len(b), line.ns), device="cuda")
        gP = torch.empty_like(points)
        stream = torch.cuda.current_stream().cuda_stream
        rc = L.hss_host_semblance_op(ctypes.addressof(dl.struct), ctypes.addressof(mp.struct), model,
                                     points.data_ptr(), gS.data_ptr(), S.data_ptr(), gP.data_ptr(),
                                     stream)
        assert rc == 0
        ref = capi.semblance_gradient(dl, mp, model, rows)
        torch.cuda.synchronize()
        assert torch.equal(S, ref[:, 0, :])
        assert torch.equal(gP, gS.unsqueeze(1) * ref[:, 2:, :])
