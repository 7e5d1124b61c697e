"""OCT traces summed in half-offset order (how the host layer and bench.py run the
non-exact DE variants).

The kernels keep the offset-dependent half of the OCT traveltime (displaced midpoint and
time terms) while consecutive traces share a half offset.  Index lists grouped by offset
make that happen on almost every trace; run through the exact variant, the oracle given
the same list must be matched bit for bit, so the kept terms are the bits a fresh
computation gives.  The natural (m, h) order, where nothing is kept, is covered by the
other OCT tests.
"""
import numpy as np
import pytest

import oracle as O
from highperformanceseismicstacking_amd import OCT, EXACT, FAST, F32, F16, BF16, capi, synth
from helpers import zocrs_window

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _case(h0):
    line, cfg = synth.workload("c4")
    line = line.subset(0, 16 * 24)
    _, _, m0all = line.cmp_ranges()
    picks = m0all[[5, 11, 17]]
    apm = 40.0
    lists = []
    for m in picks:
        b, c = zocrs_window(line, m, apm)
        idx = np.arange(b, b + c)
        lists.append(idx[np.argsort(line.hoff[idx], kind="stable")])
    counts = np.array([len(x) for x in lists])
    begins = np.r_[0, np.cumsum(counts)[:-1]]
    return line, lists, begins, counts, picks, cfg["lower"], cfg["upper"], h0, apm


def _gather(samples, line, idx, tau, m0, h0, apm):
    return O.Gather(samples[idx], line.mid[idx], line.hoff[idx], line.dt, tau, m0=m0, h0=h0,
                    apm=apm)


@pytest.mark.parametrize("storage", [F32, F16])
@pytest.mark.parametrize("h0", [0.0, 150.0])
def test_de_exact_in_offset_order_reproduces_oracle_trajectory(storage, h0):
    import torch
    line, lists, begins, counts, m0, lo, hi, h0, apm = _case(h0)
    tau = synth.tau_index(0.02, 4000)
    NP, G, seed, base = 8, 5, 0xC0FFEE, 11
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=storage)
    mp = capi.DeviceMidpoints(begins, counts, m0, index=np.concatenate(lists), h0=h0, apm=apm)
    nu = torch.zeros(mp.n, dtype=torch.int64, device="cuda")
    res, px, pfx = capi.de_semblance(dl, mp, OCT, lo, hi, NP, G, seed, stream_base=base,
                                     variant=EXACT, population=True, not_used=nu)
    torch.cuda.synchronize()
    res, px, pfx = res.cpu().numpy(), px.cpu().numpy(), pfx.cpu().numpy()
    nu = nu.cpu().numpy()
    samples = dl.decoded if storage != F32 else line.samples
    for g in range(len(m0)):
        og = _gather(samples, line, lists[g], tau, float(m0[g]), h0, apm)
        r, x, fx, rnu = O.de_f32(OCT, og, lo, hi, NP, G, seed, stream_id=base + g,
                                 population=True, not_used=True)
        assert np.array_equal(_bits(px[g]), _bits(x)), g
        assert np.array_equal(_bits(pfx[g]), _bits(fx)), g
        assert np.array_equal(_bits(res[g]), _bits(r)), g
        assert rnu > 0 and int(nu[g]) == rnu, g


@pytest.mark.parametrize("storage", [F32, F16])
@pytest.mark.parametrize("h0", [0.0, 150.0])
def test_de_fast_start_population_rejections_equal_oracle(storage, h0):
    """FAST DE walks each lane's kept traces (the compacted walk) and counts the traces the
    aperture rejects from the decisions of its first pass.  With G = 1 the individuals are
    the start population -- the same Philox draws as the oracle's -- and the decisions come
    from the IEEE traveltime chain in both, so the counts are the oracle's exactly."""
    import torch
    line, lists, begins, counts, m0, lo, hi, h0, apm = _case(h0)
    tau = synth.tau_index(0.02, 4000)
    NP, seed, base = 24, 0xBEEF, 5
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=storage)
    mp = capi.DeviceMidpoints(begins, counts, m0, index=np.concatenate(lists), h0=h0, apm=apm)
    nu = torch.zeros(mp.n, dtype=torch.int64, device="cuda")
    capi.de_semblance(dl, mp, OCT, lo, hi, NP, 1, seed, stream_base=base, variant=FAST,
                      not_used=nu)
    torch.cuda.synchronize()
    nu = nu.cpu().numpy()
    samples = dl.decoded if storage != F32 else line.samples
    for g in range(len(m0)):
        og = _gather(samples, line, lists[g], tau, float(m0[g]), h0, apm)
        _, rnu = O.de_f32(OCT, og, lo, hi, NP, 1, seed, stream_id=base + g, not_used=True)
        assert rnu > 0 and int(nu[g]) == rnu, g


@pytest.mark.parametrize("h0", [0.0, 150.0])
def test_grid_and_stretch_free_exact_in_offset_order_equal_oracle(h0):
    import torch
    line, lists, begins, counts, m0, lo, hi, h0, apm = _case(h0)
    tau = synth.tau_index(0.02, 4000)
    Gv, Gs = 24, 7
    steps = [float(np.float32(np.float32(hi[k] - lo[k]) / np.float32(G)))
             for k, G in enumerate((Gv, Gs))]
    params = capi.build_parameter_array(OCT, lo, steps, [Gv, Gs])
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau)
    mp = capi.DeviceMidpoints(begins, counts, m0, index=np.concatenate(lists), h0=h0, apm=apm)
    res, sem, _, nu = capi.grid_semblance(dl, mp, OCT, params, Gv * Gs, variant=EXACT, full=True,
                                          not_used=True)
    vpar = res[:, 2:4, :].contiguous()
    nvals = np.arange(-2, 3, dtype=np.float32)
    sf, _, _, _ = capi.stretch_free(dl, mp, OCT, vpar, nvals, variant=EXACT)
    torch.cuda.synchronize()
    res, sem, sf, vp = res.cpu().numpy(), sem.cpu().numpy(), sf.cpu().numpy(), vpar.cpu().numpy()
    nu = nu.cpu().numpy()
    pnp = params.cpu().numpy()
    for g in range(len(m0)):
        og = _gather(line.samples, line, lists[g], tau, float(m0[g]), h0, apm)
        rres, rS, _, _, rnu = O.grid_f32(OCT, og, pnp, full=True)
        assert np.array_equal(_bits(res[g]), _bits(rres)), g
        assert np.array_equal(_bits(sem[g]), _bits(rS)), g
        assert int(nu[g]) == int(rnu), g
        r_sf, _, _ = O.stretch_free_f32(OCT, og, vp[g].reshape(-1), nvals)
        assert np.array_equal(_bits(sf[g]), _bits(r_sf)), g


def test_de_fast_in_offset_order_reports_semblance_of_its_parameters():
    """FAST DE over offset-ordered lists (the benchmark's configuration): the reported
    semblance is the oracle's at the reported parameters, summed in the same order."""
    import torch
    line, lists, begins, counts, m0, lo, hi, h0, apm = _case(0.0)
    tau = synth.tau_index(0.022, 4000)
    NP, G, seed = 16, 8, 4242
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=F16)
    mp = capi.DeviceMidpoints(begins, counts, m0, index=np.concatenate(lists), h0=h0, apm=apm)
    res, _, _ = capi.de_semblance(dl, mp, OCT, lo, hi, NP, G, seed, variant=FAST)
    torch.cuda.synchronize()
    res = res.cpu().numpy()
    for g in range(len(m0)):
        og = _gather(dl.decoded, line, lists[g], tau, float(m0[g]), h0, apm)
        for s in range(0, line.ns, 17):
            v = np.float32(res[g, 2, s])
            if v == 0:
                continue
            prm = [np.float32(4.0) / (v * v), res[g, 3, s]]
            S, _, _, _ = O.eval_f32(OCT, og, s, prm)
            assert abs(S - res[g, 0, s]) <= 1e-5 * abs(S) + 1e-6, (g, s, S, res[g, 0, s])


def _sf_case(h0, storage):
    import torch
    line, lists, begins, counts, m0, lo, hi, h0, apm = _case(h0)
    tau = synth.tau_index(0.02, 4000)
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=storage)
    rng = np.random.default_rng(17)
    vp = np.stack([rng.uniform(lo[0], hi[0], (len(m0), line.ns)),
                   rng.uniform(lo[1], hi[1], (len(m0), line.ns))], axis=1).astype(np.float32)
    vpar = torch.from_numpy(vp).cuda()
    nvals = np.arange(-10, 11, dtype=np.float32)
    return line, lists, begins, counts, m0, h0, apm, tau, dl, vp, vpar, nvals


@pytest.mark.parametrize("storage", [F16, BF16])
@pytest.mark.parametrize("h0", [0.0, 150.0])
def test_stretch_free_exact_in_offset_order_equal_oracle(storage, h0):
    """Stretch-free over offset-grouped lists keeps the OCT offset terms per lane (each
    lane's parameters come from its own shifted sample): with 16-bit storage (BF16 summed
    in the FP32 steps, F16 factored) the exact variant is the oracle's bit for bit on the
    decoded samples, and the rejection count is the one of the same traces in line order
    (one aperture decision per trace and parameter pair, whatever the order)."""
    import torch
    line, lists, begins, counts, m0, h0, apm, tau, dl, vp, vpar, nvals = _sf_case(h0, storage)
    grouped = capi.DeviceMidpoints(begins, counts, m0, index=np.concatenate(lists), h0=h0, apm=apm)
    natural = capi.DeviceMidpoints(begins, counts, m0, index=np.concatenate([np.sort(x) for x in lists]),
                                   h0=h0, apm=apm)
    res, sem, stk, nu = capi.stretch_free(dl, grouped, OCT, vpar, nvals, variant=EXACT, full=True,
                                          not_used=True)
    _, _, _, nu_nat = capi.stretch_free(dl, natural, OCT, vpar, nvals, variant=EXACT, not_used=True)
    torch.cuda.synchronize()
    res, sem, stk = res.cpu().numpy(), sem.cpu().numpy(), stk.cpu().numpy()
    nu, nu_nat = nu.cpu().numpy(), nu_nat.cpu().numpy()
    for g in range(len(m0)):
        og = _gather(dl.decoded, line, lists[g], tau, float(m0[g]), h0, apm)
        r, rS, rT = O.stretch_free_f32(OCT, og, vp[g].reshape(-1), nvals, full=True)
        assert np.array_equal(_bits(res[g]), _bits(r)), g
        assert np.array_equal(_bits(sem[g]), _bits(rS)), g
        assert np.array_equal(_bits(stk[g]), _bits(rT)), g
        assert nu[g] > 0 and nu[g] == nu_nat[g], g


def test_stretch_free_fast_in_offset_order_within_tolerance():
    """FAST stretch-free over offset-grouped lists (how the host layer runs it): every
    shift's coherence within 1e-5 of the oracle's over the same list, and the picked shift
    the oracle's wherever its best has no rival within 1e-6."""
    import torch
    line, lists, begins, counts, m0, h0, apm, tau, dl, vp, vpar, nvals = _sf_case(150.0, F32)
    mp = capi.DeviceMidpoints(begins, counts, m0, index=np.concatenate(lists), h0=h0, apm=apm)
    res, sem, _, _ = capi.stretch_free(dl, mp, OCT, vpar, nvals, variant=FAST, full=True)
    torch.cuda.synchronize()
    res, sem = res.cpu().numpy(), sem.cpu().numpy()
    checked = 0
    for g in range(len(m0)):
        og = _gather(line.samples, line, lists[g], tau, float(m0[g]), h0, apm)
        r, rS, _ = O.stretch_free_f32(OCT, og, vp[g].reshape(-1), nvals, full=True)
        a, b = np.nan_to_num(sem[g], nan=-2.0), np.nan_to_num(rS, nan=-2.0)
        assert np.all(np.abs(a - b) <= 1e-5 * np.abs(b) + 1e-6), g
        near = (b >= (r[0] - 1e-6)[:, None]).sum(axis=1)
        single = near == 1
        assert np.array_equal(res[g, 2][single], r[2][single]), g
        checked += int(single.sum())
    assert checked > 0


def test_grid_and_stretch_free_mixed_orders_in_one_launch():
    """Each workgroup decides from its own midpoint's list whether to keep the offset
    terms: a launch whose midpoints alternate between offset-grouped and line-order lists
    runs both trace loops side by side, and the exact variant matches the oracle on every
    midpoint bit for bit (grid search and stretch-free)."""
    import torch
    line, lists, begins, counts, m0, lo, hi, h0, apm = _case(150.0)
    lists = [x if g % 2 == 0 else np.sort(x) for g, x in enumerate(lists)]
    for g, x in enumerate(lists):
        h = line.hoff[x].view(np.uint32)
        runs = 1 + np.count_nonzero(h[1:] != h[:-1])
        assert (4 * runs <= len(x)) == (g % 2 == 0), g
    tau = synth.tau_index(0.02, 4000)
    Gv, Gs = 16, 5
    steps = [float(np.float32(np.float32(hi[k] - lo[k]) / np.float32(G)))
             for k, G in enumerate((Gv, Gs))]
    params = capi.build_parameter_array(OCT, lo, steps, [Gv, Gs])
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=F16)
    mp = capi.DeviceMidpoints(begins, counts, m0, index=np.concatenate(lists), h0=h0, apm=apm)
    res, sem, _, _ = capi.grid_semblance(dl, mp, OCT, params, Gv * Gs, variant=EXACT, full=True)
    vpar = res[:, 2:4, :].contiguous()
    nvals = np.arange(-4, 5, dtype=np.float32)
    sf, ssem, _, _ = capi.stretch_free(dl, mp, OCT, vpar, nvals, variant=EXACT, full=True)
    torch.cuda.synchronize()
    res, sem, sf, ssem = res.cpu().numpy(), sem.cpu().numpy(), sf.cpu().numpy(), ssem.cpu().numpy()
    vp = vpar.cpu().numpy()
    pnp = params.cpu().numpy()
    for g in range(len(m0)):
        og = _gather(dl.decoded, line, lists[g], tau, float(m0[g]), h0, apm)
        rres, rS, _, _, _ = O.grid_f32(OCT, og, pnp, full=True)
        assert np.array_equal(_bits(res[g]), _bits(rres)), g
        assert np.array_equal(_bits(sem[g]), _bits(rS)), g
        r_sf, r_ss, _ = O.stretch_free_f32(OCT, og, vp[g].reshape(-1), nvals, full=True)
        assert np.array_equal(_bits(sf[g]), _bits(r_sf)), g
        assert np.array_equal(_bits(ssem[g]), _bits(r_ss)), g
