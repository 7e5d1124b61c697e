#!/usr/bin/env python3
"""Semblance throughput on one MI355X.

Default step (--workload c5): the whole 2000-midpoint CMP line (fold 24-120, 2001 samples
at 2 ms, 2048 velocities 1400-5000 m/s, 15-sample window) -- grid ("linear") search with
the best-velocity pick fused in, one launch, then the stretch-free re-stack (21 shifts,
-10..10) of the winners, one launch; both captured as hipGraphs and replayed.  `value`
counts the grid and stretch-free evaluations of a step (one eval = one (t0, candidate)
pair) over the step's wall time.  Inputs are resident in HBM before the timed region.
Prints ONE JSON line.

Unless --no-extras, the line also carries, measured in the same run on the same GPU:
the FP16-storage step; the EXACT (reference-bits) grid launch against the plain per-gather
FP32 EXACT kernel (the reference's kernel structure); the precision path (FP16 storage,
compensated window sums, then 10 Gauss-Newton steps of the adjoint polish); the C2 line
(500 midpoints, 1024 velocities); the ZO-CRS DE line time (C3) and the OCT DE line time
(C4), each with a grid search over the same midpoints and traces (a candidate grid across
the DE bounds) and the stretch-free re-stack of the DE winners; and a device-to-device copy
as the measured HBM reference.

Roofline of the dominant kernel (the grid launch; HIP events on its stream, inside the
timed region): the kernel is bound by its FP32 vector instruction stream, so `achieved`
is FP32 flop/s of the flop model fold * ns * P * (6 (w+1) + 8) per launch against the
157.3 TFLOP/s FP32 peak.  Beside it: the gathered-sample stream (fold * ns * P * (w+1) *
sample bytes per launch), the HBM traffic and LDS figures of the rocprofv3 counter
summary committed under profiles/ for the same kernel and workload, and the measured
copy bandwidth.  --workload c3 / c4 run the DE lines alone (line time = DE + polish).
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3      # MI355X FP32 peak (dual-issue / packed FP32 counted)
FP32_FMA_ISSUE_TFLOPS = 78.6  # non-packed FP32 FMAs on 16-lane SIMDs: 256 CU x 4 x 16 x 2 x 2.4 GHz
LDS_PEAK_GBS = 256 * 128 * 2.4  # 128 bytes per clock per CU


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c5", choices=["c2", "c3", "c4", "c5"],
                    help="c5: CMP grid search + stretch-free over a 2000-CMP line (default, "
                         "the headline); c2: CMP grid search over a 500-CMP line; c3: ZO-CRS "
                         "DE + Gauss-Newton polish, BF16 (line time); c4: OCT DE + polish, FP16")
    ap.add_argument("--polish-steps", type=int, default=10,
                    help="Gauss-Newton steps after DE (c3, c4) and in the precision path")
    ap.add_argument("--storage", default=None, choices=["f32", "f16", "bf16"],
                    help="trace storage (default: f32 for c2/c5, bf16 for c3, f16 for c4)")
    ap.add_argument("--variant", default="fast", choices=["fast", "exact", "kahan"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU work budget of the oracle baseline (three equal samples; 0 "
                         "disables it)")
    ap.add_argument("--cpl", type=int, default=0, choices=[0, 1, 2],
                    help="pin the grid parameter tile (candidates per lane); 0 = autotune")
    ap.add_argument("--t0-tiles", type=int, default=1, choices=[1, 2, 4],
                    help="t0 tiles per workgroup when --cpl pins the shape")
    ap.add_argument("--staged", type=int, default=0, choices=[0, 1],
                    help="workgroup-staged traces (CMP, one t0 tile) when --cpl pins the shape")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the step directly instead of replaying captured hipGraphs")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline step (no comparison legs, no other workloads)")
    ap.add_argument("--grid-leg", action="store_true",
                    help="c3 / c4: also time a grid search over the DE line's midpoints and "
                         "traces (config.grid_search; the default run's extras include it)")
    ap.add_argument("--sf-leg", action="store_true",
                    help="c3 / c4: also time the stretch-free re-stack of the DE winners "
                         "(config.stretch_free; the default run's extras include it)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="after the timed steps, run one plain device-to-device copy of this "
                         "many MiB (a known-size reference for rocprofv3 FETCH/WRITE counters)")
    return ap.parse_args()


# ------------------------------------------------------------------ helpers
def progress(msg):
    """One line per finished phase on stderr (the JSON line alone goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def calibration_copy(mb):
    """One plain device-to-device copy of `mb` MiB, the known-size reference that scales
    rocprofv3's FETCH_SIZE / WRITE_SIZE counters (nothing when mb is 0)."""
    if mb <= 0:
        return
    import torch
    nbytes = mb << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    # write 1 GiB elsewhere first, so no part of src is still held in a cache
    flush = torch.zeros(1 << 28, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    del flush
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def copy_bandwidth(mib=2048, reps=5):
    """Measured HBM bandwidth of a device-to-device copy of `mib` MiB (read + write bytes
    over the median time of `reps` copies, HIP events)."""
    import torch
    n = (mib << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    ms = statistics.median(times)
    del src, dst
    return 2.0 * (mib << 20) / (ms * 1e-3) / 1e9


def time_region(fn, steps, warmup, stream):
    """Average ms of fn() over `steps` runs after `warmup`, HIP events on `stream`."""
    import torch
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            fn()
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / max(steps, 1)


def read_profile_figures(kernel, workload, storage):
    """Figures per launch of `kernel` ("grid", "de", "sf") on `workload` from the committed
    rocprofv3 summary profiles/traffic_<kernel>_<workload>_<storage>.txt: HBM bytes
    (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), LDS bytes and LDS bank-conflict
    cycles, and the kernel time the counters were taken at.  {} when none is committed."""
    name = f"traffic_{kernel}_{workload}_{storage}.txt"
    path = os.path.join(ROOT, "profiles", name)
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        if ":" not in line or line.startswith("#"):
            continue
        k, v = line.split(":", 1)
        try:
            out[k.strip()] = float(v)
        except ValueError:
            pass
    out["source"] = f"profiles/{name} (rocprofv3 --pmc of this kernel and workload)"
    return out


def roofline(flops, gathered_bytes, launch_ms, prof, copy_gbs):
    """The roofline object of the dominant kernel.  bound = "valu": the launch is bound by
    its FP32 vector instruction stream (the counter summaries show HBM and LDS far below
    their peaks), so `achieved` is algorithmic FP32 TFLOP/s against the FP32 peak."""
    t = launch_ms * 1e-3
    achieved = flops / t / 1e12
    out = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
           "frac": achieved / FP32_PEAK_TFLOPS,
           "traffic": prof.get("traffic_bytes_per_launch"),
           "flops_per_launch": flops,
           "frac_of_fma_issue_rate": achieved / FP32_FMA_ISSUE_TFLOPS,
           "gathered_stream_gbs": gathered_bytes / t / 1e9,
           "hbm_peak_gbs": HBM_PEAK_GBS,
           "copy_gbs_measured": copy_gbs}
    if prof.get("traffic_bytes_per_launch") is not None:
        out["hbm_traffic_gbs"] = prof["traffic_bytes_per_launch"] / t / 1e9
        out["hbm_frac"] = out["hbm_traffic_gbs"] / HBM_PEAK_GBS
    if prof.get("lds_bytes_per_launch") is not None:
        out["lds_gbs"] = prof["lds_bytes_per_launch"] / t / 1e9
        out["lds_peak_gbs"] = LDS_PEAK_GBS
        out["lds_frac"] = out["lds_gbs"] / LDS_PEAK_GBS
        out["lds_bank_conflict_cycles"] = prof.get("lds_bank_conflict_cycles")
    if prof:
        out["traffic_source"] = prof["source"]
    return out


def grid_flops(counts, ns, P, w, model_extra=0):
    """Flop model of a grid launch: per (t0, candidate, trace) 6 flops per window value
    (difference, interpolation, two sums and a square) plus 8 for the traveltime and
    index, plus the model's extra traveltime flops."""
    import numpy as np
    units = float(np.sum(np.asarray(counts, dtype=np.float64))) * ns * P
    return units * (6 * (w + 1) + 8 + model_extra)


def cpu_baseline(line, tau, params, begins, counts, m0s, seconds):
    """The CPU oracle (C restatement of the reference loops, OpenMP over t0) on the first
    midpoints of the same line: three samples of seconds/3 each, median rate."""
    import numpy as np

    import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))   # the box's CPU share
    O.lib()
    rates, done_total = [], 0
    g = 0
    per = seconds / 3.0
    for _ in range(3):
        evals, t_start = 0, time.perf_counter()
        while True:
            gi = g % len(begins)
            sl = slice(int(begins[gi]), int(begins[gi] + counts[gi]))
            og = O.Gather(line.samples[sl], line.mid[sl], line.hoff[sl], line.dt, tau,
                          m0=float(m0s[gi]))
            O.grid_f32(O.CMP, og, params, nthreads=threads, full=False)
            evals += line.ns * params.size
            g += 1
            if time.perf_counter() - t_start >= per:
                break
        rates.append(evals / (time.perf_counter() - t_start))
        done_total += evals
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": statistics.median(rates), "unit": "evals/s", "cores": threads,
            "kind": "port",
            "sample": f"first {g} of {len(begins)} midpoints of the same line in three timed "
                      f"samples of ~{per:.0f} s (median of {[round(r / 1e6, 1) for r in rates]} "
                      f"M evals/s), FP32 C oracle built -O3 -march=x86-64-v3 without "
                      f"contraction, OpenMP over t0 on {threads} threads",
            "host_cpus": os.cpu_count(), "cpu_model": model}


# ------------------------------------------------------------------ grid lines
class GridLine:
    """A CMP line resident on the device for grid search, midpoints in the XCD-balanced
    launch order (the host layer's batching; results come back per midpoint)."""

    def __init__(self, name, storage, line_cfg=None):
        import numpy as np
        import torch

        from highperformanceseismicstacking_amd import CMP, capi, synth
        self.name = name
        self.line, self.cfg = line_cfg if line_cfg is not None else synth.workload(name)
        line, cfg = self.line, self.cfg
        self.tau = synth.tau_index(cfg["tau"], line.dt_us)
        self.w = 2 * self.tau + 1
        self.lo, self.hi, self.P = cfg["lower"][0], cfg["upper"][0], cfg["granularity"][0]
        self.step = float(np.float32(np.float32(self.hi - self.lo) / np.float32(self.P)))
        b, c, m = line.cmp_ranges()
        self.line_order = (b, c, m)
        order = capi.xcd_balanced_order(c)
        self.begins, self.counts, self.m0s = b[order], c[order], m[order]
        self.storage = storage
        self.dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, self.tau,
                                  storage=storage)
        self.mp = capi.DeviceMidpoints(self.begins, self.counts, self.m0s)
        self.params = capi.build_parameter_array(CMP, [self.lo], [self.step], [self.P])
        self.result = torch.empty((self.mp.n, 3, line.ns), dtype=torch.float32, device="cuda")
        self.vpar = torch.empty((self.mp.n, 1, line.ns), dtype=torch.float32, device="cuda")
        self.nvals = np.arange(-10, 11, dtype=np.float32)
        torch.cuda.synchronize()

    @property
    def grid_evals(self):
        return int(self.mp.n) * self.line.ns * self.P

    @property
    def sf_evals(self):
        return int(self.mp.n) * self.line.ns * self.nvals.size

    def grid(self, variant, stream=None):
        from highperformanceseismicstacking_amd import CMP, capi
        capi.grid_semblance(self.dl, self.mp, CMP, self.params, self.P, variant=variant,
                            result=self.result, stream=stream)

    def sf(self, variant, stream=None):
        # stretch-free re-stack of the grid winners: shifts -10..10 around each sample
        # (StretchFreeAlgorithm.cpp:77-82), velocity per sample from the grid result
        from highperformanceseismicstacking_amd import CMP, capi
        self.vpar.copy_(self.result[:, 2:3, :])
        capi.stretch_free(self.dl, self.mp, CMP, self.vpar, self.nvals, variant=variant,
                          stream=stream)

    def gathered_bytes(self):
        import numpy as np
        return (float(np.sum(self.counts.astype(np.float64))) * self.line.ns * self.P
                * (self.w + 1) * self.dl.sample_bytes())

    def describe(self):
        c = self.counts
        return (f"{self.name}: CMP grid search, {self.mp.n} midpoints, fold {int(c.min())}-"
                f"{int(c.max())} (mean {c.mean():.1f}), ns={self.line.ns} "
                f"dt={float(self.line.dt) * 1e3:g} ms, {self.P} velocities {self.lo:g}-"
                f"{self.hi:g} m/s, w={self.w}")


def time_grid_steps(gl, variant, steps, warmup, with_sf, use_graph):
    """Runs `steps` steps (grid, then stretch-free when with_sf) after `warmup`, the parts
    captured as hipGraphs on one stream; returns (wall ms per step, grid launch ms, sf ms),
    the launch times from HIP events recorded on that stream around each part."""
    import torch
    gl.grid(variant)      # outside any capture: the one-time index-sequence check, autotune
    if with_sf:
        gl.sf(variant)
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    g_grid = g_sf = None
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            gl.grid(variant, stream)
            if with_sf:
                gl.sf(variant, stream)
        stream.synchronize()
        if use_graph:
            g_grid = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g_grid, stream=stream):
                gl.grid(variant, stream)
            if with_sf:
                g_sf = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g_sf, stream=stream):
                    gl.sf(variant, stream)
            g_grid.replay()
            if g_sf is not None:
                g_sf.replay()
        stream.synchronize()

    def run_grid():
        g_grid.replay() if g_grid is not None else gl.grid(variant, stream)

    def run_sf():
        g_sf.replay() if g_sf is not None else gl.sf(variant, stream)

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        for i in range(steps):
            ev[i][0].record(stream)
            run_grid()
            ev[i][1].record(stream)
            if with_sf:
                run_sf()
            ev[i][2].record(stream)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    grid_ms = sum(e[0].elapsed_time(e[1]) for e in ev) / steps
    sf_ms = sum(e[1].elapsed_time(e[2]) for e in ev) / steps if with_sf else 0.0
    return wall_ms, grid_ms, sf_ms


def grid_extras(args, gl, grid_ms):
    """Comparison legs on the headline line: FP16 storage, EXACT against the plain
    per-gather EXACT kernel, and the precision path."""
    import torch
    from highperformanceseismicstacking_amd import CMP, EXACT, FAST, F16, KAHAN, capi
    out = {}
    stream = torch.cuda.Stream()
    # FP16 storage, the same step
    g16 = GridLine(gl.name, F16, (gl.line, gl.cfg))
    _, g16_ms, sf16_ms = time_grid_steps(g16, FAST, 3, 1, True, True)
    progress(f"FP16 step: grid {g16_ms:.1f} ms, stretch-free {sf16_ms:.1f} ms")
    out["f16_storage"] = {"grid_launch_ms": g16_ms, "stretch_free_launch_ms": sf16_ms,
                          "evals_per_s": (g16.grid_evals + g16.sf_evals) / ((g16_ms + sf16_ms) * 1e-3)}
    # its LDS traffic (the wave segments), per launch from the committed counter summary,
    # over this run's launch time
    p16 = read_profile_figures("grid", gl.name, "f16")
    if p16.get("lds_bytes_per_launch") is not None:
        lds_gbs = p16["lds_bytes_per_launch"] / (g16_ms * 1e-3) / 1e9
        out["f16_storage"].update({
            "lds_gbs": lds_gbs, "lds_peak_gbs": LDS_PEAK_GBS, "lds_frac": lds_gbs / LDS_PEAK_GBS,
            "hbm_traffic_bytes_per_launch": p16.get("traffic_bytes_per_launch"),
            "counters_source": p16["source"]})
    # the precision path: FP16 storage, compensated window sums, then the adjoint polish
    # of every winner (Gauss-Newton steps on the semblance gradient)
    kah_ms = time_region(lambda: g16.grid(KAHAN, stream), 2, 1, stream)
    pol = torch.empty_like(g16.result)
    lo, hi = [float(gl.lo)], [float(gl.hi)]

    def polish():
        capi.polish(g16.dl, g16.mp, CMP, lo, hi, args.polish_steps, g16.result, result=pol,
                    stream=stream)

    pol_ms = time_region(polish, 2, 1, stream)
    progress(f"precision path: FP16 Kahan grid {kah_ms:.1f} ms, polish {pol_ms:.1f} ms")
    # EXACT (the reference's operation order, bit for bit) against the plain per-gather
    # kernel (the reference's kernel structure, one launch pair per midpoint, EXACT)
    ex_ms = time_region(lambda: gl.grid(EXACT, stream), 2, 1, stream)
    dl = gl.dl
    scratch = (torch.empty((gl.line.ns, gl.P), dtype=torch.float32, device="cuda"),
               torch.empty((gl.line.ns, gl.P), dtype=torch.float32, device="cuda"))
    res = torch.empty((gl.mp.n, 3, gl.line.ns), dtype=torch.float32, device="cuda")
    b, c, m = gl.begins, gl.counts, gl.m0s
    capi.grid_semblance_plain(dl, CMP, gl.params, gl.P, b[:2], c[:2], m[:2], result=res,
                              scratch=scratch)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    capi.grid_semblance_plain(dl, CMP, gl.params, gl.P, b, c, m, result=res, scratch=scratch)
    e1.record()
    torch.cuda.synchronize()
    plain_ms = e0.elapsed_time(e1)
    progress(f"EXACT grid {ex_ms:.1f} ms, plain per-gather EXACT {plain_ms:.1f} ms")
    del scratch, res
    n_ev = gl.grid_evals
    out["plain_per_gather_fp32_exact"] = {
        "launch_ms": plain_ms, "evals_per_s": n_ev / (plain_ms * 1e-3),
        "note": "the reference's kernel structure (thread = (t0, candidate), one launch pair "
                "per midpoint), IEEE sqrt/division; the reference itself builds with "
                "--use_fast_math (cuda/Makefile:26)"}
    out["exact_grid_launch_ms"] = ex_ms
    out["speedup_exact_vs_plain_exact"] = plain_ms / ex_ms
    out["speedup_fast_vs_plain_exact"] = plain_ms / grid_ms
    out["speedup_f16_fast_vs_plain_exact"] = plain_ms / g16_ms
    out["precision_path"] = {
        "what": f"FP16 storage + compensated (Kahan) window sums, then {args.polish_steps} "
                "Gauss-Newton steps of the adjoint polish of every winner",
        "grid_launch_ms": kah_ms, "polish_launch_ms": pol_ms,
        "evals_per_s": n_ev / ((kah_ms + pol_ms) * 1e-3),
        "speedup_vs_plain_exact": plain_ms / (kah_ms + pol_ms)}
    del g16, pol
    torch.cuda.empty_cache()
    return out


def other_workloads(args):
    """The C2 line (grid) and the C3 / C4 DE line times."""
    import torch
    from highperformanceseismicstacking_amd import FAST, F32, capi
    out = {}
    c2 = GridLine("c2", F32)
    _, c2_ms, _ = time_grid_steps(c2, FAST, 5, 1, False, True)
    progress(f"c2 grid {c2_ms:.2f} ms")
    out["c2_grid"] = {"workload": c2.describe(), "launch_ms": c2_ms,
                      "evals_per_s": c2.grid_evals / (c2_ms * 1e-3),
                      "flops_tflops": grid_flops(c2.counts, c2.line.ns, c2.P, c2.w) / (c2_ms * 1e-3) / 1e12}
    del c2
    torch.cuda.empty_cache()
    for wl, storage in (("c3", "bf16"), ("c4", "f16")):
        de_args = argparse.Namespace(**vars(args))
        de_args.workload, de_args.storage, de_args.variant = wl, storage, "fast"
        de_args.steps, de_args.warmup, de_args.calibrate_copy_mb = 3, 1, 0
        de_args.grid_leg = de_args.sf_leg = True
        de = run_de(de_args)
        key = "zocrs_de" if wl == "c3" else "oct_de"
        progress(f"{wl} line {de['config']['line_time_s']:.3f} s, grid "
                 f"{de['config']['grid_search']['launch_ms']:.1f} ms, stretch-free "
                 f"{de['config']['stretch_free']['launch_ms']:.1f} ms")
        out[key] = {"workload": de["config"]["workload"],
                    "line_time_s": de["config"]["line_time_s"],
                    "de_launch_ms": de["config"]["de_launch_ms"],
                    "polish_launch_ms": de["config"]["polish_launch_ms"],
                    "evals_per_s": de["value"],
                    "grid_search": de["config"]["grid_search"],
                    "stretch_free": de["config"]["stretch_free"]}
        torch.cuda.empty_cache()
    return out


def run_grid(args):
    import numpy as np

    from highperformanceseismicstacking_amd import CMP, EXACT, FAST, KAHAN, F32, F16, BF16, capi
    storage = {"f32": F32, "f16": F16, "bf16": BF16}[args.storage]
    variant = {"fast": FAST, "exact": EXACT, "kahan": KAHAN}[args.variant]
    gl = GridLine(args.workload, storage)
    if args.cpl:
        capi.grid_set_shape(CMP, gl.w, variant, storage, args.cpl, args.t0_tiles, args.staged)
    with_sf = args.workload == "c5"
    progress(f"{args.workload} line resident ({gl.mp.n} midpoints)")
    wall_ms, grid_ms, sf_ms = time_grid_steps(gl, variant, args.steps, args.warmup, with_sf,
                                              not args.no_graph)
    progress(f"timed steps: {wall_ms:.1f} ms per step, grid {grid_ms:.1f} ms, "
             f"stretch-free {sf_ms:.1f} ms")
    calibration_copy(args.calibrate_copy_mb)
    evals = gl.grid_evals + (gl.sf_evals if with_sf else 0)
    config = {
        "workload": gl.describe() + (", then stretch-free re-stack (21 shifts)" if with_sf else ""),
        "storage": args.storage,
        "variant": args.variant,
        "evals_per_step": evals,
        "grid_evals_per_step": gl.grid_evals,
        "grid_launch_ms": grid_ms,
        "stretch_free_launch_ms": sf_ms,
        "graph": not args.no_graph,
        "launch_shape": dict(zip(("candidates_per_lane", "t0_tiles_per_workgroup", "staged"),
                                 capi.grid_tuned_shape(CMP, gl.w, variant, storage))),
        "launch_order": "xcd_balanced",
    }
    copy_gbs = None
    if not args.no_extras:
        copy_gbs = copy_bandwidth()
        progress(f"copy {copy_gbs:.0f} GB/s")
        config.update(grid_extras(args, gl, grid_ms))
        config.update(other_workloads(args))
    prof = read_profile_figures("grid", args.workload, args.storage)
    out = {
        "metric": "semblance evals/s (t0 x param), CMP grid search"
                  + (" + stretch-free" if with_sf else "") + " over a line, 1 GPU",
        "value": evals / (wall_ms * 1e-3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "roofline": roofline(grid_flops(gl.counts, gl.line.ns, gl.P, gl.w), gl.gathered_bytes(),
                             grid_ms, prof, copy_gbs),
        "cpu_baseline": None,
    }
    if args.cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(gl.line, gl.tau, gl.params.cpu().numpy(),
                                           *gl.line_order, args.cpu_seconds)
        progress("CPU baseline done")
    return out


# ------------------------------------------------------------------ DE lines
def run_de(args):
    """c3 / c4: differential evolution over the whole line in one launch, then the
    Gauss-Newton polish of every winner; one step = DE + polish (the line time).  Returns
    the JSON object."""
    import numpy as np
    import torch

    from highperformanceseismicstacking_amd import (F32, F16, BF16, FAST, EXACT, KAHAN, ZOCRS, OCT,
                                                   capi, synth)
    storage = {"f32": F32, "f16": F16, "bf16": BF16}[args.storage]
    variant = {"fast": FAST, "exact": EXACT, "kahan": KAHAN}[args.variant]
    line, cfg = synth.workload(args.workload)
    tau = synth.tau_index(cfg["tau"], line.dt_us)
    w = 2 * tau + 1
    _, _, m0s = line.cmp_ranges()
    f32 = np.float32
    model = ZOCRS if args.workload == "c3" else OCT
    NP, G = cfg["population"], cfg["generations"]
    seed = 20240601
    if model == ZOCRS:
        ain = f32(f32(f32(cfg["ain"]) * f32(np.pi)) / f32(180.0))
        ua = f32(f32(f32(2.0) * f32(np.sin(ain, dtype=np.float32))) / f32(cfg["v0"]))
        ub = f32(f32(f32(cfg["bpctg"]) * f32(4.0)) / f32(f32(cfg["lower"][0]) * f32(cfg["lower"][0])))
        lower = [cfg["lower"][0], float(-ua), float(-ub)]
        upper = [cfg["upper"][0], float(ua), float(ub)]
    else:
        lower, upper = cfg["lower"], cfg["upper"]
    h0 = float(cfg.get("h0", 0.0))
    apm = float(cfg["apm"])
    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=storage)
    stream = torch.cuda.Stream()
    sel_ms = 0.0
    if model == ZOCRS:
        # |m0 - m| <= apm: a contiguous range of the (m, h)-sorted line (zocrs/common.cu:14)
        lo_i = np.searchsorted(line.mid, m0s - f32(apm), side="left")
        begins, counts = [], []
        for g, m0 in enumerate(m0s):
            b = int(lo_i[g])
            while b > 0 and abs(f32(m0) - line.mid[b - 1]) <= f32(apm):
                b -= 1
            e = b
            while e < line.ntr and abs(f32(m0) - line.mid[e]) <= f32(apm):
                e += 1
            begins.append(b)
            counts.append(e - b)
        mp = capi.DeviceMidpoints(begins, counts, m0s, h0=h0, apm=apm, order="balanced")
        counts = np.array(counts)
    else:
        # OCT: the traces some start-population candidate reaches (device selection)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        mask = capi.select_traces(dl, OCT, m0s, apm=apm, h0=h0, oct_mode=capi.OCT_DE_START, NP=NP,
                                  lower=lower, upper=upper, seed=seed, stream_base=0)
        e1.record()
        torch.cuda.synchronize()
        sel_ms = e0.elapsed_time(e1)
        m = mask.cpu().numpy().astype(bool)
        idx = [np.flatnonzero(m[g]) for g in range(len(m0s))]
        if variant != EXACT:
            # grouped by half offset (as the host layer does): the kernel then computes
            # the offset-dependent half of the traveltime once per group
            idx = [x[np.argsort(line.hoff[x], kind="stable")] for x in idx]
        counts = np.array([len(x) for x in idx])
        begins = np.r_[0, np.cumsum(counts)[:-1]]
        mp = capi.DeviceMidpoints(begins, counts, m0s, index=np.concatenate(idx), h0=h0, apm=apm,
                                  order="balanced")
    np_ = 3 if model == ZOCRS else 2
    nbytes = capi.de_workspace_bytes(mp.n, line.ns, model, NP)
    ws = torch.empty(max(nbytes, 4) // 4, dtype=torch.float32, device="cuda")
    res = torch.empty((mp.n, 2 + np_, line.ns), dtype=torch.float32, device="cuda")
    pol = torch.empty_like(res)

    def de():
        capi.de_semblance(dl, mp, model, lower, upper, NP, G, seed, variant=variant, result=res,
                          workspace=ws, stream=stream)

    def pl():
        capi.polish(dl, mp, model, lower, upper, args.polish_steps, res, result=pol, stream=stream)

    # warm-up outside timing: the first launch runs the one-time index-sequence check
    de()
    pl()
    torch.cuda.synchronize()
    de_ms = time_region(de, args.steps, args.warmup, stream)
    pol_ms = time_region(pl, args.steps, args.warmup, stream) if args.polish_steps > 0 else 0.0
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        for _ in range(args.steps):
            de()
            if args.polish_steps > 0:
                pl()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    resn = res.cpu().numpy()
    poln = pol.cpu().numpy()
    n_evals = int(mp.n) * line.ns * NP * G
    sample_bytes = dl.sample_bytes()
    # gathered samples: every (t0, individual, generation) reads w+1 samples of each
    # selected trace it keeps; OCT drops the traces outside the displaced-midpoint
    # aperture (notUsed), counted by one more, untimed, launch
    kept = float(np.sum(counts.astype(np.float64))) * line.ns * NP * G
    if model == OCT:
        nu = torch.zeros(mp.n, dtype=torch.int64, device="cuda")
        capi.de_semblance(dl, mp, model, lower, upper, NP, G, seed, variant=variant,
                          workspace=ws, not_used=nu)
        torch.cuda.synchronize()
        kept -= float(nu.sum().item())
    alg_bytes = kept * (w + 1) * sample_bytes
    # flop model: the window and index work of the grid model plus the traveltime
    # (ZO-CRS: 6 more flops; OCT: about 40 more flops and 4 square roots per trace)
    flops = kept * (6 * (w + 1) + 8 + (6 if model == ZOCRS else 40))
    live = resn[:, 2, :] > 0
    prof = read_profile_figures("de", args.workload, args.storage)
    out = {
        "metric": ("semblance evals/s (t0 x individual x generation), "
                   f"{'ZO-CRS' if model == ZOCRS else 'OCT'} differential evolution over a line, 1 GPU"),
        "value": n_evals / (de_ms * 1e-3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1000.0,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {'ZO-CRS' if model == ZOCRS else 'OCT'} DE, {mp.n} "
                        f"midpoints, traces per midpoint {int(counts.min())}-{int(counts.max())} "
                        f"(mean {counts.mean():.0f}), ns={line.ns} dt={float(line.dt)*1e3:g} ms, "
                        f"NP={NP}, G={G}, w={w}, then {args.polish_steps} Gauss-Newton steps",
            "storage": args.storage,
            "variant": args.variant,
            "line_time_s": wall,
            "de_launch_ms": de_ms,
            "polish_launch_ms": pol_ms,
            "selection_ms": sel_ms,
            "launch_order": "xcd_balanced",
            "evals_per_step": n_evals,
            "mean_semblance_de": float(resn[:, 0, :][live].mean()),
            "mean_semblance_polished": float(poln[:, 0, :][live].mean()),
        },
        "roofline": roofline(flops, alg_bytes, de_ms, prof, None),
        "cpu_baseline": None,
    }
    if getattr(args, "grid_leg", False):
        # grid search over the same midpoints and selected traces: a candidate grid that
        # spans the DE bounds (the variant and storage of the DE run)
        grid_counts = (16, 8, 8) if model == ZOCRS else (64, 16)
        grid_steps = [float(f32(f32(upper[k] - lower[k]) / f32(grid_counts[k])))
                      for k in range(np_)]
        P = int(np.prod(grid_counts))
        gparams = capi.build_parameter_array(model, lower, grid_steps, list(grid_counts))
        gres = torch.empty((mp.n, 2 + np_, line.ns), dtype=torch.float32, device="cuda")

        def gr():
            capi.grid_semblance(dl, mp, model, gparams, P, variant=variant, result=gres,
                                stream=stream)

        gr()  # the first launch of at least 4096 workgroups tunes the launch shape
        torch.cuda.synchronize()
        grid_ms = time_region(gr, max(2, args.steps), 1, stream)
        out["config"]["grid_search"] = {
            "candidates": " x ".join(str(x) for x in grid_counts) + " over the DE bounds",
            "P": P,
            "launch_ms": grid_ms,
            "evals_per_s": int(mp.n) * line.ns * P / (grid_ms * 1e-3),
        }
    if getattr(args, "sf_leg", False):
        out["config"]["stretch_free"] = stretch_free_leg(args, line, dl, mp, model, res, m0s, h0,
                                                         apm, variant, stream)
    calibration_copy(args.calibrate_copy_mb)
    return out


def stretch_free_leg(args, line, dl, mp, model, res, m0s, h0, apm, variant, stream):
    """The stretch-free re-stack (21 shifts, -10..10) of a DE line's winners over the same
    midpoints.  ZO-CRS keeps the DE ranges; OCT selects each midpoint's traces with the
    winners' per-sample (velocity, slope), as the host layer does, grouped by half offset
    for the non-exact variants."""
    import numpy as np
    import torch

    from highperformanceseismicstacking_amd import EXACT, OCT, capi
    np_ = res.shape[1] - 2
    vpar = res[:, 2:2 + np_, :].contiguous()
    nvals = np.arange(-10, 11, dtype=np.float32)
    smp = mp
    counts = None
    if model == OCT:
        mask = capi.select_traces(dl, OCT, m0s, apm=apm, h0=h0, oct_mode=capi.OCT_PER_SAMPLE,
                                  vparams=vpar)
        torch.cuda.synchronize()
        m = mask.cpu().numpy().astype(bool)
        del mask
        idx = [np.flatnonzero(m[g]) for g in range(len(m0s))]
        if variant != EXACT:
            idx = [x[np.argsort(line.hoff[x], kind="stable")] for x in idx]
        counts = np.array([len(x) for x in idx])
        begins = np.r_[0, np.cumsum(counts)[:-1]]
        smp = capi.DeviceMidpoints(begins, counts, m0s, index=np.concatenate(idx), h0=h0, apm=apm,
                                   order="balanced")

    def sf():
        capi.stretch_free(dl, smp, model, vpar, nvals, variant=variant, stream=stream)

    sf()
    torch.cuda.synchronize()
    sf_ms = time_region(sf, max(2, args.steps), 1, stream)
    out = {"shifts": int(nvals.size), "launch_ms": sf_ms,
           "evals_per_s": int(smp.n) * line.ns * int(nvals.size) / (sf_ms * 1e-3)}
    prof = read_profile_figures("sf", args.workload, args.storage)
    if "traffic_bytes_per_launch" in prof:
        out["traffic"] = prof["traffic_bytes_per_launch"]
        out["traffic_source"] = prof["source"]
    if counts is not None:
        out["traces_per_midpoint"] = f"{int(counts.min())}-{int(counts.max())} (mean {counts.mean():.0f})"
    return out


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (a line would "
              "shard by midpoint ranges, one process per GPU, with no communication during "
              "the search)", file=sys.stderr)
        sys.exit(2)
    if args.storage is None:
        args.storage = {"c3": "bf16", "c4": "f16"}.get(args.workload, "f32")
    if args.workload in ("c3", "c4"):
        out = run_de(args)
    else:
        out = run_grid(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
