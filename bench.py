#!/usr/bin/env python3
"""Semblance throughput on one MI355X: grid ("linear") search over a whole CMP line.

One step = one pass of the hot path over the line: every (midpoint, t0, velocity)
semblance of the 500-midpoint line (fold 48-96, 1501 samples at 2 ms, 1024 velocities
1400-5000 m/s, 11-sample window) with the best-velocity pick fused in, in one launch.
Inputs are resident in HBM before the timed region starts.  Prints ONE JSON line.
Unless --no-extras, the line also carries the plain per-gather FP32 baseline, the FP16
storage timing and the second part of the metric, the ZO-CRS differential-evolution line
time (--workload c3 on its own gives that workload's full line).

Roofline: the algorithmic bytes of a launch are the gathered-sample stream,
sum over (midpoint, t0, candidate) of fold * (w+1) * sample_bytes, divided by the
launch's average duration (HIP events on the launch stream).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5"],
                    help="c2: CMP grid search over a 500-CMP line (the headline metric); "
                         "c3: ZO-CRS DE + Gauss-Newton polish, BF16 (line time); c4: OCT DE + "
                         "polish, FP16; c5: CMP grid search + stretch-free, 2000 CMPs")
    ap.add_argument("--polish-steps", type=int, default=10,
                    help="Gauss-Newton steps after DE (c3, c4)")
    ap.add_argument("--storage", default=None, choices=["f32", "f16", "bf16"],
                    help="trace storage (default: f32 for c2/c5, bf16 for c3, f16 for c4)")
    ap.add_argument("--variant", default="fast", choices=["fast", "exact", "kahan"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU work budget of the oracle baseline sample (0 disables it)")
    ap.add_argument("--cpl", type=int, default=0, choices=[0, 1, 2],
                    help="pin the grid parameter tile (candidates per lane); 0 = autotune")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the step directly instead of replaying a captured hipGraph")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the plain per-gather baseline and the FP16-storage timing")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="after the timed steps, run one plain device-to-device copy of this "
                         "many MiB (a known-size reference for rocprofv3 FETCH/WRITE counters)")
    return ap.parse_args()


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


def roofline(achieved, traffic, launch_ms):
    """The roofline object: `achieved` is the gathered-sample stream (algorithmic bytes
    per launch / launch time, GB/s) against the HBM peak; `traffic` the HBM bytes per
    launch the counters measured (None when no summary is committed), and
    `traffic_gbs` that traffic over the launch time.  The gathered stream is served
    mostly from registers, L1, L2 and LDS, so `frac` can exceed 1; the kernels are bound
    by the vector ALU, not by HBM."""
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic}
    if traffic is not None:
        out["traffic_gbs"] = traffic / (launch_ms * 1e-3) / 1e9
    out["binding"] = "vector ALU"
    return out


def read_traffic(kernel, workload, storage):
    """HBM bytes per launch of `kernel` ("grid", "de") measured by rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), when a summary for this workload is
    committed."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}_{storage}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(line, tau, params, begins, counts, m0s, seconds):
    """The CPU oracle (C restatement, OpenMP over t0) on the first midpoints of the same
    line until `seconds` of work are done."""
    import numpy as np

    import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    O.lib()
    evals = 0
    t_start = time.perf_counter()
    done = 0
    for g in range(len(begins)):
        sl = slice(int(begins[g]), int(begins[g] + counts[g]))
        og = O.Gather(line.samples[sl], line.mid[sl], line.hoff[sl], line.dt, tau,
                      m0=float(m0s[g]))
        O.grid_f32(O.CMP, og, params, nthreads=threads, full=False)
        evals += line.ns * (params.size)
        done += 1
        if time.perf_counter() - t_start >= seconds:
            break
    el = time.perf_counter() - t_start
    return {"value": evals / el, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"first {done} of {len(begins)} midpoints of the same line "
                      f"({evals} evals, {el:.1f} s), FP32 C oracle, OpenMP over t0"}


def measure_extras(args, line, tau, model, params, G, begins, counts, m0s, mp, variant,
                   launch_ms):
    """Same line, same GPU: (1) the plain per-gather FP32 kernel (the reference's kernel
    structure, one launch pair per midpoint), timed once over the whole line; (2) the
    batched kernel on FP16-storage traces (per-trace power-of-two scale)."""
    import torch
    from highperformanceseismicstacking_amd import F16, capi
    out = {}
    n_evals = int(mp.n) * line.ns * G
    if args.storage == "f32":
        dl32 = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau)
        scratch = (torch.empty((line.ns, G), dtype=torch.float32, device="cuda"),
                   torch.empty((line.ns, G), dtype=torch.float32, device="cuda"))
        res = torch.empty((mp.n, 3, line.ns), dtype=torch.float32, device="cuda")
        capi.grid_semblance_plain(dl32, model, params, G, begins[:2], counts[:2], m0s[:2],
                                  result=res, scratch=scratch)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        capi.grid_semblance_plain(dl32, model, params, G, begins, counts, m0s, result=res,
                                  scratch=scratch)
        e1.record()
        torch.cuda.synchronize()
        plain_ms = e0.elapsed_time(e1)
        out["plain_fp32_per_gather_ms"] = plain_ms
        out["plain_fp32_per_gather_evals_per_s"] = n_evals / (plain_ms * 1e-3)
        out["speedup_vs_plain_fp32"] = plain_ms / launch_ms
        del scratch, res, dl32
    dl16 = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=F16)
    res = torch.empty((mp.n, 3, line.ns), dtype=torch.float32, device="cuda")
    capi.grid_semblance(dl16, mp, model, params, G, variant=variant, result=res)
    torch.cuda.synchronize()
    reps = 3
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        capi.grid_semblance(dl16, mp, model, params, G, variant=variant, result=res)
    e1.record()
    torch.cuda.synchronize()
    f16_ms = e0.elapsed_time(e1) / reps
    out["f16_storage_launch_ms"] = f16_ms
    out["f16_storage_evals_per_s"] = n_evals / (f16_ms * 1e-3)
    if "plain_fp32_per_gather_ms" in out:
        out["f16_speedup_vs_plain_fp32"] = out["plain_fp32_per_gather_ms"] / f16_ms
    return out


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
    return e0.elapsed_time(e1) / steps


def bench_de(args):
    print(json.dumps(run_de(args)), flush=True)


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
                          workspace=ws)

    def pl():
        capi.polish(dl, mp, model, lower, upper, args.polish_steps, res, result=pol)

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
    achieved = alg_bytes / (de_ms * 1e-3) / 1e9
    live = resn[:, 2, :] > 0
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
        "roofline": roofline(achieved, read_traffic("de", args.workload, args.storage),
                             de_ms),
        "cpu_baseline": None,
    }
    calibration_copy(args.calibrate_copy_mb)
    return out


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (a line would "
              "shard by midpoint ranges, one process per GPU, with no communication during "
              "the search)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from highperformanceseismicstacking_amd import CMP, EXACT, FAST, KAHAN, F32, F16, BF16, capi, synth

    if args.storage is None:
        args.storage = {"c3": "bf16", "c4": "f16"}.get(args.workload, "f32")
    if args.workload in ("c3", "c4"):
        return bench_de(args)
    storage = {"f32": F32, "f16": F16, "bf16": BF16}[args.storage]
    variant = {"fast": FAST, "exact": EXACT, "kahan": KAHAN}[args.variant]
    line, cfg = synth.workload(args.workload)
    tau = synth.tau_index(cfg["tau"], line.dt_us)
    w = 2 * tau + 1
    lo, hi, G = cfg["lower"][0], cfg["upper"][0], cfg["granularity"][0]
    if args.workload != "c2":
        args.cpu_seconds = 0.0  # the CPU baseline is quoted on the headline workload
    step = float(np.float32(np.float32(hi - lo) / np.float32(G)))
    begins, counts, m0s = line.cmp_ranges()
    # the whole line in one launch, midpoints in the order that balances the XCDs (the
    # host layer's batching; results come back per midpoint)
    line_order = (begins, counts, m0s)
    order = capi.xcd_balanced_order(counts)
    begins, counts, m0s = begins[order], counts[order], m0s[order]

    dl = capi.DeviceLine(line.samples, line.mid, line.hoff, line.dt, tau, storage=storage)
    mp = capi.DeviceMidpoints(begins, counts, m0s)
    params = capi.build_parameter_array(CMP, [lo], [step], [G])
    result = torch.empty((mp.n, 3, line.ns), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    if args.cpl:
        capi.grid_set_config(CMP, w, variant, storage, args.cpl)

    def one_step():
        capi.grid_semblance(dl, mp, CMP, params, G, variant=variant, result=result)

    # warm-up outside any capture: the first launch for a sample interval runs the
    # one-time check of the short sqrt / division sequences (a synchronous call)
    one_step()
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    graph = None
    with torch.cuda.stream(stream):
        for _ in range(args.warmup):
            one_step()
        stream.synchronize()
        if not args.no_graph:
            # the step as one hipGraph (captured on `stream`; replays go to the same stream)
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, stream=stream):
                one_step()
            graph.replay()
        stream.synchronize()

    def run_step():
        if graph is not None:
            graph.replay()
        else:
            one_step()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        for i in range(args.steps):
            starts[i].record(stream)
            run_step()
            ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / args.steps

    extras = {}
    if not args.no_extras and args.workload == "c2":
        extras = measure_extras(args, line, tau, CMP, params, G, begins, counts, m0s, mp, variant,
                                launch_ms)
        # the metric's second part: the ZO-CRS DE line time (c3: 200 midpoints, NP 32,
        # G 50, BF16, then 10 Gauss-Newton steps), measured as `--workload c3` does
        de_args = argparse.Namespace(**vars(args))
        de_args.workload, de_args.storage, de_args.variant = "c3", "bf16", "fast"
        de_args.steps, de_args.warmup, de_args.calibrate_copy_mb = 3, 1, 0
        de = run_de(de_args)
        extras["zocrs_de_line_time_s"] = de["config"]["line_time_s"]
        extras["zocrs_de_launch_ms"] = de["config"]["de_launch_ms"]
        extras["zocrs_de_evals_per_s"] = de["value"]
        extras["zocrs_de_workload"] = de["config"]["workload"]
    if args.workload == "c5":
        # stretch-free re-stack of the grid winners: shifts -10..10 around each sample
        # (StretchFreeAlgorithm.cpp:77-82), velocity per sample from the grid result
        vpar = result[:, 2:3, :].contiguous()
        nvals = np.arange(-10, 11, dtype=np.float32)

        def sf():
            capi.stretch_free(dl, mp, CMP, vpar, nvals, variant=variant)

        sf_ms = time_region(sf, max(1, args.steps // 2), 1, stream)
        extras["stretch_free_launch_ms"] = sf_ms
        extras["stretch_free_evals_per_s"] = int(mp.n) * line.ns * nvals.size / (sf_ms * 1e-3)

    calibration_copy(args.calibrate_copy_mb)

    n_evals = int(mp.n) * line.ns * G
    sample_bytes = dl.sample_bytes()
    alg_bytes = float(np.sum(counts.astype(np.float64))) * line.ns * G * (w + 1) * sample_bytes
    ms_per_step = wall * 1000.0 / args.steps
    achieved = alg_bytes / (launch_ms * 1e-3) / 1e9
    traffic = read_traffic("grid", args.workload, args.storage)

    out = {
        "metric": ("semblance evals/s (t0 x param), CMP grid search over a line, 1 GPU"
                   + (" (then stretch-free)" if args.workload == "c5" else "")),
        "value": n_evals / (ms_per_step * 1e-3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: CMP grid search, {mp.n} midpoints, fold "
                        f"{int(counts.min())}-{int(counts.max())} (mean {counts.mean():.1f}), "
                        f"ns={line.ns} dt={float(line.dt)*1e3:g} ms, {G} velocities "
                        f"{lo:g}-{hi:g} m/s, w={w}",
            "storage": args.storage,
            "variant": args.variant,
            "evals_per_step": n_evals,
            "launch_ms": launch_ms,
            "graph": graph is not None,
            "parameter_tile": capi.grid_tuned_config(CMP, w, variant, storage),
            **extras,
        },
        "roofline": roofline(achieved, traffic, launch_ms),
        "cpu_baseline": None,
    }
    if args.cpu_seconds > 0:
        pnp = params.cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(line, tau, pnp, *line_order, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
