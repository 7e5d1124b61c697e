"""The search launches captured in a hipGraph and replayed (how bench.py runs the grid step).

Graph nodes are replayed with the captured arguments; the entry points record no memset
and no host synchronization once the one-time per-(device, dt) sequence check and the
autotuner have run, which is why a caller makes one eager call before capturing.  Each
test replays three times with eager launches and synchronizations in between and checks
every replay against the eager result bit for bit.
"""
import numpy as np
import pytest

from highperformanceseismicstacking_amd import CMP, ZOCRS, EXACT, FAST, BF16, capi, synth
from helpers import zocrs_bounds, zocrs_window

pytestmark = pytest.mark.gpu


def test_grid_step_graph_replay_equals_eager():
    import torch
    line = synth.make_line(6, (30, 60), 3000.0, 801, 2000, seed=3, offset_min=50.0)
    tau = synth.tau_index(0.011, line.dt_us)
    G = 257
    step = float(np.float32(np.float32(5000.0 - 1400.0) / np.float32(G)))
    b, c, m0 = line.cmp_ranges()
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau)
    mp = capi.DeviceMidpoints(b, c, m0)
    params = capi.build_parameter_array(CMP, [1400.0], [step], [G])
    eager, _, _, _ = capi.grid_semblance(dl, mp, CMP, params, G, variant=FAST)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    out = torch.empty_like(eager)
    capi.grid_semblance(dl, mp, CMP, params, G, variant=FAST, result=out, stream=s)
    s.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        capi.grid_semblance(dl, mp, CMP, params, G, variant=FAST, result=out, stream=s)
    for _ in range(3):
        out.fill_(-7.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager)
        # eager work between the replays
        capi.grid_semblance(dl, mp, CMP, params, G, variant=EXACT)
        torch.cuda.synchronize()


def test_de_graph_replay_equals_eager():
    import torch
    line, cfg = synth.workload("c3")
    line = line.subset(0, 24 * 12)
    lo, hi = zocrs_bounds(cfg["lower"][0], cfg["upper"][0], cfg["ain"], cfg["v0"], cfg["bpctg"])
    _, _, m0all = line.cmp_ranges()
    picks = m0all[[3, 8]]
    rng = [zocrs_window(line, m, 60.0) for m in picks]
    tau = synth.tau_index(0.022, line.dt_us)
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=BF16)
    mp = capi.DeviceMidpoints([r[0] for r in rng], [r[1] for r in rng], picks, apm=60.0)
    NP, G, seed = 16, 6, 77
    eager, _, _ = capi.de_semblance(dl, mp, ZOCRS, lo, hi, NP, G, seed, variant=FAST)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    nbytes = capi.de_workspace_bytes(mp.n, line.ns, ZOCRS, NP)
    ws = torch.empty(max(nbytes, 4) // 4, dtype=torch.float32, device="cuda")
    out = torch.empty_like(eager)
    capi.de_semblance(dl, mp, ZOCRS, lo, hi, NP, G, seed, variant=FAST, result=out, workspace=ws,
                      stream=s)
    s.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        capi.de_semblance(dl, mp, ZOCRS, lo, hi, NP, G, seed, variant=FAST, result=out,
                          workspace=ws, stream=s)
    for _ in range(3):
        out.fill_(-7.0)
        ws.fill_(1e30)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager)
        capi.de_semblance(dl, mp, ZOCRS, lo, hi, NP, G, seed + 1, variant=FAST)
        torch.cuda.synchronize()


def test_first_call_under_capture_then_eager_uses_verified_sequences():
    """A first semblance call for a new dt made under stream capture cannot run the
    one-time sequence check (it would wait on the capturing stream): it takes the library
    operations (same bits) and caches nothing, so the next eager call runs the check and
    the kernels use the verified short sequences from then on."""
    import torch
    dt_us = 3000                      # a sample interval no other test uses
    line = synth.make_line(3, 24, 2000.0, 401, dt_us, seed=12, offset_min=50.0)
    tau = synth.tau_index(0.012, dt_us)
    G = 64
    step = float(np.float32(np.float32(4500.0 - 1500.0) / np.float32(G)))
    b, c, m0 = line.cmp_ranges()
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau)
    mp = capi.DeviceMidpoints(b, c, m0)
    params = capi.build_parameter_array(CMP, [1500.0], [step], [G])
    torch.cuda.synchronize()
    assert capi.index_ops_state(line.dt) == -1
    s = torch.cuda.Stream()
    out = torch.empty((mp.n, 3, line.ns), dtype=torch.float32, device="cuda")
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        capi.grid_semblance(dl, mp, CMP, params, G, variant=FAST, result=out, stream=s)
    assert capi.index_ops_state(line.dt) == -1
    graph.replay()
    torch.cuda.synchronize()
    eager, _, _, _ = capi.grid_semblance(dl, mp, CMP, params, G, variant=FAST)
    torch.cuda.synchronize()
    assert capi.index_ops_state(line.dt) == 1
    assert torch.equal(out, eager)
    for _ in range(2):
        out.fill_(-7.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager)
