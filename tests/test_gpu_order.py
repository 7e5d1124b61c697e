"""Launch order (hss_midpoints.order): the schedule changes, the results do not.

Every search entry point, run with a shuffled order and with the XCD-balanced order, gives
the natural order's result bit for bit: each midpoint keeps its result rows, its workspace
rows and its DE random stream.  A DE run split over two launches with different orders
equals one launch.  A slot whose order entry lies outside [0, n) computes nothing and
leaves its midpoint's rows as they were.
"""
import numpy as np
import pytest

from highperformanceseismicstacking_amd import CMP, ZOCRS, EXACT, FAST, F16, BF16, capi, synth
from helpers import zocrs_bounds, zocrs_window

pytestmark = pytest.mark.gpu


def _orders(n, counts):
    rng = np.random.default_rng(n)
    return {"shuffled": rng.permutation(n), "balanced": "balanced",
            "reversed": np.arange(n)[::-1]}


def _cmp_case(storage):
    line = synth.make_line(19, (12, 70), 3000.0, 601, 2000, seed=5, offset_min=50.0)
    tau = synth.tau_index(0.011, line.dt_us)
    b, c, m0 = line.cmp_ranges()
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=storage)
    return line, dl, b, c, m0


def _zocrs_case():
    line, cfg = synth.workload("c3")
    line = line.subset(0, 24 * 20)
    lo, hi = zocrs_bounds(cfg["lower"][0], cfg["upper"][0], cfg["ain"], cfg["v0"], cfg["bpctg"])
    _, _, m0all = line.cmp_ranges()
    picks = m0all[1:-1:2]
    rng = [zocrs_window(line, m, 60.0) for m in picks]
    tau = synth.tau_index(0.022, line.dt_us)
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=BF16)
    return line, dl, [r[0] for r in rng], [r[1] for r in rng], picks, lo, hi


@pytest.mark.parametrize("variant", [EXACT, FAST])
def test_grid_and_stretch_free_results_do_not_depend_on_the_order(variant):
    import torch
    line, dl, b, c, m0 = _cmp_case(F16)
    G = 96
    step = float(np.float32(np.float32(5000.0 - 1400.0) / np.float32(G)))
    params = capi.build_parameter_array(CMP, [1400.0], [step], [G])
    nvals = np.arange(-3, 4, dtype=np.float32)
    mp = capi.DeviceMidpoints(b, c, m0)
    ref, _, _, _ = capi.grid_semblance(dl, mp, CMP, params, G, variant=variant)
    vpar = ref[:, 2:3, :].contiguous()
    ref_sf, _, _, _ = capi.stretch_free(dl, mp, CMP, vpar, nvals, variant=variant)
    torch.cuda.synchronize()
    for name, order in _orders(len(b), c).items():
        mo = capi.DeviceMidpoints(b, c, m0, order=order)
        got, _, _, _ = capi.grid_semblance(dl, mo, CMP, params, G, variant=variant)
        got_sf, _, _, _ = capi.stretch_free(dl, mo, CMP, vpar, nvals, variant=variant)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), name
        assert torch.equal(got_sf, ref_sf), name


def test_de_gradient_and_polish_results_do_not_depend_on_the_order():
    import torch
    line, dl, b, c, m0, lo, hi = _zocrs_case()
    NP, G, seed = 16, 8, 31
    mp = capi.DeviceMidpoints(b, c, m0, apm=60.0)
    ref, _, _ = capi.de_semblance(dl, mp, ZOCRS, lo, hi, NP, G, seed, stream_base=4, variant=FAST)
    ref_gr = capi.semblance_gradient(dl, mp, ZOCRS, ref)
    ref_pol = capi.polish(dl, mp, ZOCRS, lo, hi, 4, ref)
    torch.cuda.synchronize()
    assert (ref[:, 0, :] > 0).float().mean() > 0.5
    for name, order in _orders(len(b), c).items():
        mo = capi.DeviceMidpoints(b, c, m0, apm=60.0, order=order)
        got, _, _ = capi.de_semblance(dl, mo, ZOCRS, lo, hi, NP, G, seed, stream_base=4,
                                      variant=FAST)
        gr = capi.semblance_gradient(dl, mo, ZOCRS, ref)
        pol = capi.polish(dl, mo, ZOCRS, lo, hi, 4, ref)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), name
        assert torch.equal(gr, ref_gr), name
        assert torch.equal(pol, ref_pol), name


def test_de_split_over_launches_with_different_orders_equals_one_launch():
    import torch
    line, dl, b, c, m0, lo, hi = _zocrs_case()
    NP, G, seed = 16, 8, 9
    n = len(b)
    mp = capi.DeviceMidpoints(b, c, m0, apm=60.0)
    ref, _, _ = capi.de_semblance(dl, mp, ZOCRS, lo, hi, NP, G, seed, variant=FAST)
    nbytes = capi.de_workspace_bytes(n, line.ns, ZOCRS, NP)
    ws = torch.empty(max(nbytes, 4) // 4, dtype=torch.float32, device="cuda")
    out = torch.empty_like(ref)
    m1 = capi.DeviceMidpoints(b, c, m0, apm=60.0, order="balanced")
    m2 = capi.DeviceMidpoints(b, c, m0, apm=60.0, order=np.arange(n)[::-1])
    capi.de_semblance(dl, m1, ZOCRS, lo, hi, NP, G, seed, variant=FAST, result=out, workspace=ws,
                      gen_range=(0, 3))
    capi.de_semblance(dl, m2, ZOCRS, lo, hi, NP, G, seed, variant=FAST, result=out, workspace=ws,
                      gen_range=(3, G))
    torch.cuda.synchronize()
    assert torch.equal(out, ref)


def test_order_entry_outside_the_batch_computes_nothing():
    import torch
    line, dl, b, c, m0 = _cmp_case(F16)
    G = 32
    step = float(np.float32(np.float32(5000.0 - 1400.0) / np.float32(G)))
    params = capi.build_parameter_array(CMP, [1400.0], [step], [G])
    n = len(b)
    ref, _, _, _ = capi.grid_semblance(dl, capi.DeviceMidpoints(b, c, m0), CMP, params, G)
    order = np.arange(n)
    order[3] = n + 5   # midpoint 3 is not taken up by any slot
    order[7] = -1      # nor is midpoint 7
    out = torch.full_like(ref, -7.0)
    capi.grid_semblance(dl, capi.DeviceMidpoints(b, c, m0, order=order), CMP, params, G,
                        result=out)
    torch.cuda.synchronize()
    skipped = np.zeros(n, bool)
    skipped[[3, 7]] = True
    assert torch.all(out[torch.from_numpy(skipped)] == -7.0)
    assert torch.equal(out[torch.from_numpy(~skipped)], ref[torch.from_numpy(~skipped)])
