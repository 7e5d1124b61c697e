"""Differential evolution against a brute-force optimum.

DE is judged on its final semblance per sample (its random stream differs from the
reference's), so besides the trajectory parity with the oracle the optimiser itself is
checked: on subsets shaped like the ZO-CRS and OCT benchmark lines, with the benchmark's
population and generation counts, DE's final semblance is compared with the best
semblance of an exhaustive fine grid over the same bounds (the grid search is pinned to
the oracle bit for bit elsewhere).  A sample where DE ends more than 1e-3 below the
grid's best has missed the global optimum.  The reference's DE (F = 0.85, CR = 0.5, a
population of 16 to 48 for 30 to 50 generations) converges early on a multimodal
semblance, so it misses on a share of the samples; the test bounds that share, requires
the median gap to be within 1e-3, and requires the generations to pay: DE's final
population must reach the grid optimum far more often than its start population does.
DE may also end above the grid's best: between grid points, and where the parameters
leave only a few traces inside the trace length (the semblance of a single trace is 1,
in the reference's formula as here), which no grid point of the fine grid may hit.

Each case prints one line:
  DE-OPTIMUM <case> within_1e-3=<fraction> start_within=<fraction of the start
  population's best> median_gap=<S_grid - S_de> worst_gap=<...> best_over_grid=<...>
"""
import numpy as np
import pytest

from highperformanceseismicstacking_amd import CMP, ZOCRS, OCT, F32, F16, BF16, FAST, capi, synth
from helpers import grid_steps, zocrs_bounds, zocrs_window

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "cmp":
        line = synth.make_line(4, (24, 48), 3000.0, 501, 4000, seed=41, offset_min=50.0)
        b, c, m0 = line.cmp_ranges()
        return (line, CMP, b, c, m0, [1500.0], [4500.0], [1024], 0.0, 0.0, F32, 16, 30)
    if name == "zocrs":
        line, cfg = synth.workload("c3")
        line = line.subset(0, 24 * 24)
        lo, hi = zocrs_bounds(cfg["lower"][0], cfg["upper"][0], cfg["ain"], cfg["v0"],
                              cfg["bpctg"])
        _, _, m0all = line.cmp_ranges()
        picks = m0all[[6, 12, 18]]
        rng = [zocrs_window(line, m, 80.0) for m in picks]
        return (line, ZOCRS, np.array([r[0] for r in rng]), np.array([r[1] for r in rng]),
                picks, lo, hi, [96, 40, 40], 0.0, 80.0, BF16, cfg["population"],
                cfg["generations"])
    line, cfg = synth.workload("c4")
    line = line.subset(0, 16 * 30)
    _, _, m0all = line.cmp_ranges()
    picks = m0all[[8, 15, 22]]
    rng = [zocrs_window(line, m, 60.0) for m in picks]
    return (line, OCT, np.array([r[0] for r in rng]), np.array([r[1] for r in rng]), picks,
            cfg["lower"], cfg["upper"], [256, 96], 0.0, 60.0, F16, cfg["population"],
            cfg["generations"])


def _within(Sg, S, live):
    gap = (Sg - S)[live]
    return float(np.mean(gap <= 1e-3)), gap


# min_within: measured share minus a margin (the runs are deterministic: Philox, seed)
@pytest.mark.parametrize("name,min_within", [("cmp", 0.75), ("zocrs", 0.58), ("oct", 0.60)])
def test_de_reaches_the_fine_grid_optimum(name, min_within):
    import torch
    (line, model, b, c, m0, lo, hi, G, h0, apm, storage, NP, gens) = _case(name)
    tau = synth.tau_index(0.02, line.dt_us)
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=storage)
    mp = capi.DeviceMidpoints(b, c, m0, h0=h0, apm=apm)
    params = capi.build_parameter_array(model, lo, grid_steps(lo, hi, G), G)
    P = int(np.prod(G))
    grid, _, _, _ = capi.grid_semblance(dl, mp, model, params, P, variant=FAST)
    seed = 20240601
    de, _, _ = capi.de_semblance(dl, mp, model, lo, hi, NP, gens, seed, variant=FAST)
    start, _, _ = capi.de_semblance(dl, mp, model, lo, hi, NP, 1, seed, variant=FAST)
    torch.cuda.synchronize()
    Sg = grid[:, 0, :].cpu().numpy().astype(np.float64)
    Sd = de[:, 0, :].cpu().numpy().astype(np.float64)
    S0 = start[:, 0, :].cpu().numpy().astype(np.float64)
    live = Sg > 0
    assert live.mean() > 0.5
    within, gap = _within(Sg, Sd, live)
    within0, _ = _within(Sg, S0, live)
    print(f"DE-OPTIMUM {name} within_1e-3={within:.4f} start_within={within0:.4f} "
          f"median_gap={np.median(gap):.2e} worst_gap={gap.max():.3e} "
          f"best_over_grid={-gap.min():.3e}", flush=True)
    assert within >= min_within
    assert np.median(gap) <= 1e-3
    assert within >= within0 + 0.2
    # the population's best never falls below the start population's best (elitist
    # selection: an individual is replaced only by a strictly better trial)
    assert np.all(Sd >= S0)
    assert np.all(Sd <= 1.0 + 1e-6)
