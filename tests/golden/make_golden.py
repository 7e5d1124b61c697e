#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the reference-built checker.

Runs in the build container only (it needs oracle/_ref, compiled from the reference's own
sources by oracle/ref/Makefile).  The fixtures are data: seeded inputs and the outputs the
reference's code produced for them.
  kernel_eval.npz   semblance/stack/used/notUsed of single (t0, candidate) evaluations
                    through the reference's COMPUTE_SEMBLANCE / REDUCE_SEMBLANCE_STACK and
                    OCT computeDisplacedMidpoint / computeTime, for CMP, ZO-CRS and OCT;
  oct_functions.npz OCT displaced midpoint / traveltime with their error codes;
  gather_*.su/.npz  SU files and what the reference's Gather::readGatherFromFile keeps;
  traveltime.npz    ZO-CRS / OCT parameter bounds from TraveltimeBuilder;
  dump_<gather>_<model>/  the output files the reference's own Dumper (with ResultSet and the
                    result classes) writes for a gather_*.su file, a traveltime model and a
                    seeded result set (values.f32 [cdp][result][ns], stats.f32 [cdp][5],
                    args.json), in the order SingleHostRunner::main writes them.
"""
import json
import shutil
import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from highperformanceseismicstacking_amd import synth  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
f32 = np.float32


def ref_lib():
    L = ctypes.CDLL(os.path.join(REF, "libref_kernel.so"))
    vp, fl, u, i = ctypes.c_void_p, ctypes.c_float, ctypes.c_uint, ctypes.c_int
    L.ref_eval.argtypes = [i, vp, vp, vp, vp, u, u, fl, i, i, fl, fl, fl, u, vp, vp, vp, vp, vp]
    L.ref_oct_displaced_midpoint.argtypes = [fl] * 6 + [ctypes.POINTER(fl)]
    L.ref_oct_time.argtypes = [fl] * 8 + [ctypes.POINTER(fl)]
    return L


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def kernel_cases(rng):
    """Small gathers with candidates chosen to hit the edges: window at the first and last
    admissible sample, ZO-CRS negative radicands, OCT rejections and error branches."""
    cases = []
    for model in (0, 1, 2):
        for trial in range(6):
            ntr = int(rng.integers(1, 9))
            ns = int(rng.integers(24, 80))
            dt = float(synth.dt_seconds([2000, 4000][trial % 2]))
            tau = int(rng.integers(0, 6))
            w = 2 * tau + 1
            samples = rng.normal(0, 1, (ntr, ns)).astype(np.float32)
            if trial == 5:
                samples[:, :] = 0.0                # all-zero windows: S = 0/0 = NaN
            m0 = f32(1000.0)
            mid = (m0 + rng.uniform(-60, 60, ntr)).astype(np.float32) if model else \
                np.full(ntr, m0, np.float32)
            hoff = rng.uniform(-400, 400, ntr).astype(np.float32)
            if model == 2 and trial == 4:
                hoff[0] = f32(0.0)                  # |h| == |h0| branch with h0 = 0
            hoff2 = (hoff * hoff).astype(np.float32)
            h0 = f32(0.0) if trial % 3 == 0 else f32(rng.uniform(-200, 200))
            apm = f32(rng.uniform(5, 80))
            evals = []
            for _ in range(40):
                s = int(rng.integers(0, ns))
                v = f32(rng.uniform(1200, 5000))
                c = f32(f32(4.0) / f32(v * v))
                if model == 0:
                    prm = [c, 0, 0]
                elif model == 1:
                    prm = [c, f32(rng.uniform(-2e-3, 2e-3)), f32(rng.uniform(-3e-5, 3e-5))]
                else:
                    prm = [c, f32(rng.uniform(-2e-3, 2e-3)), 0]
                evals.append((s, prm))
            cases.append(dict(model=model, samples=samples, mid=mid, hoff=hoff, hoff2=hoff2,
                              dt=f32(dt), tau=tau, w=w, m0=m0, h0=h0, apm=apm, evals=evals))
    return cases


def make_kernel_eval(L, rng):
    rows = {k: [] for k in ["case", "s", "prm", "S", "stack", "used", "notused"]}
    arrays = {}
    for ci, c in enumerate(kernel_cases(rng)):
        for name in ("samples", "mid", "hoff", "hoff2"):
            arrays[f"c{ci}_{name}"] = c[name]
        arrays[f"c{ci}_scalars"] = np.array([c["model"], c["tau"], c["w"]], np.int32)
        arrays[f"c{ci}_fscalars"] = np.array([c["dt"], c["m0"], c["h0"], c["apm"]], np.float32)
        ntr, ns = c["samples"].shape
        for s, prm in c["evals"]:
            prm = np.array(prm, np.float32)
            S, T = np.zeros(1, np.float32), np.zeros(1, np.float32)
            used, nu = np.zeros(1, np.uint32), np.zeros(1, np.uint32)
            L.ref_eval(c["model"], p(c["samples"]), p(c["mid"]), p(c["hoff"]), p(c["hoff2"]), ntr,
                       ns, float(c["dt"]), c["tau"], c["w"], float(c["m0"]), float(c["h0"]),
                       float(c["apm"]), s, p(prm), p(S), p(T), p(used), p(nu))
            rows["case"].append(ci)
            rows["s"].append(s)
            rows["prm"].append(prm)
            rows["S"].append(S[0])
            rows["stack"].append(T[0])
            rows["used"].append(used[0])
            rows["notused"].append(nu[0])
    out = {k: np.array(v) for k, v in rows.items()}
    out.update(arrays)
    out["ncases"] = np.array(len(arrays) // 6)
    np.savez_compressed(os.path.join(HERE, "kernel_eval.npz"), **out)
    return out


def make_oct_functions(L, rng):
    n = 3000
    h = rng.uniform(-1500, 1500, n).astype(np.float32)
    h0 = rng.choice([0.0, 100.0, -250.0, 800.0], n).astype(np.float32)
    h[::17] = h0[::17]                     # |h| == |h0|
    h[::23] = -h0[::23]
    t0 = (rng.integers(0, 1000, n) * f32(0.004)).astype(np.float32)
    t0[::29] = 0.0                          # division-by-zero branches
    m0 = np.full(n, 5000.0, np.float32)
    v = rng.uniform(1400, 5000, n).astype(np.float32)
    c = (f32(4.0) / (v * v)).astype(np.float32)
    slope = rng.uniform(-1e-3, 1e-3, n).astype(np.float32)
    slope[::31] = 0.0
    m = (m0 + rng.uniform(-300, 300, n)).astype(np.float32)
    err1 = np.zeros(n, np.int32)
    mh = np.zeros(n, np.float32)
    err2 = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    for i in range(n):
        x = ctypes.c_float(0)
        err1[i] = L.ref_oct_displaced_midpoint(h[i], h0[i], t0[i], m0[i], c[i], slope[i],
                                               ctypes.byref(x))
        mh[i] = x.value if err1[i] == 0 else np.nan
        y = ctypes.c_float(0)
        if err1[i] == 0:
            err2[i] = L.ref_oct_time(h[i], h0[i], t0[i], m0[i], m[i], mh[i], c[i], slope[i],
                                     ctypes.byref(y))
            t[i] = y.value if err2[i] == 0 else np.nan
        else:
            err2[i] = -1
            t[i] = np.nan
    np.savez_compressed(os.path.join(HERE, "oct_functions.npz"), h=h, h0=h0, t0=t0, m0=m0, c=c,
                        slope=slope, m=m, err1=err1, mh=mh, err2=err2, t=t)


def make_gather_fixture(name, line, aph, apm, tau, az, lowest, highest, shuffle_seed,
                        bad_ns_rows=()):
    """An SU file with the traces shuffled (the reader sorts them), optionally some rows of
    a different ns, and the reference reader's view of it."""
    su = os.path.join(HERE, f"gather_{name}.su")
    rng = np.random.default_rng(shuffle_seed)
    order = rng.permutation(line.ntr)
    shuffled = synth.Line(line.samples[order], line.sx[order], line.gx[order], line.scalco,
                          line.ns, line.dt_us, line.cdp[order])
    synth.write_su(su, shuffled)
    if bad_ns_rows:
        # rewrite the chosen records with a shorter trace: same header but ns - 4 samples
        hdr, data = synth.read_su(su, line.ns)
        recs = []
        raw = open(su, "rb").read()
        rec = 240 + 4 * line.ns
        for r in range(line.ntr):
            blob = bytearray(raw[r * rec:(r + 1) * rec])
            if r in bad_ns_rows:
                blob[114:116] = np.array([line.ns - 4], "<u2").tobytes()
                blob = blob[: 240 + 4 * (line.ns - 4)]
            recs.append(bytes(blob))
        open(su, "wb").write(b"".join(recs))
    prefix = os.path.join("/tmp", f"golden_{name}")
    subprocess.run([os.path.join(REF, "ref_host"), "gather", su, str(aph), str(apm), str(tau),
                    str(az), lowest, highest, prefix], check=True)
    info = {}
    cdps, tm, th = [], [], []
    for ln in open(prefix + ".txt"):
        k, *vals = ln.split()
        if k == "cdp":
            cdps.append(float.fromhex(vals[0]))
        elif k == "trace":
            tm.append(float.fromhex(vals[0]))
            th.append(float.fromhex(vals[1]))
        elif k == "dt":
            info[k] = float.fromhex(vals[0])
        else:
            info[k] = int(vals[0])
    samples = np.fromfile(prefix + ".bin", dtype=np.float32).reshape(len(tm), info["ns"]) \
        if tm else np.zeros((0, info["ns"]), np.float32)
    np.savez_compressed(os.path.join(HERE, f"gather_{name}.npz"),
                        params=np.array([aph, apm, tau, az], np.float64),
                        lowest=np.array(lowest), highest=np.array(highest),
                        ns=info["ns"], dt=np.float32(info["dt"]), tau_index=info["tau"],
                        w=info["w"], cdps=np.array(cdps, np.float32),
                        mid=np.array(tm, np.float32), hoff=np.array(th, np.float32),
                        samples=samples)


def make_traveltime():
    rows = []
    for vmin, vmax, v0, ain, bp in [(1581.139, 5000, 2000, 90, 0.625), (1500, 4500, 2000, 60, 0.1),
                                    (1400, 5000, 1800, 30, 0.25), (2000, 6000, 2500, 45, 1.0)]:
        out = subprocess.run([os.path.join(REF, "ref_host"), "zocrs", str(vmin), str(vmax),
                              str(v0), str(ain), str(bp)], check=True, capture_output=True,
                             text=True).stdout.split("\n")
        vals = [float.fromhex(x) for ln in out[:3] for x in ln.split()[1:]]
        rows.append([vmin, vmax, v0, ain, bp] + vals)
    np.savez_compressed(os.path.join(HERE, "traveltime.npz"),
                        zocrs=np.array(rows, np.float64))


DUMP_CASES = [
    # (gather, read args (aph, apm, tau, azimuth), model, lower, upper, v0, ain, bpctg, h0, method)
    ("basic", (500.0, 0.0, 0.02, 90.0), "cmp", [1500.0], [4500.0], 0.0, 0.0, 0.0, 0.0, "greedy"),
    ("basic", (500.0, 0.0, 0.02, 90.0), "zocrs", [1600.0], [4800.0], 2000.0, 60.0, 0.5, 0.0, "de"),
    ("basic", (500.0, 0.0, 0.02, 90.0), "oct", [1500.0, -5e-4], [4500.0, 5e-4], 0.0, 0.0, 0.0,
     100.0, "greedy"),
    ("filters", (400.0, 0.0, 0.022, 90.0), "cmp", [1400.0], [5000.0], 0.0, 0.0, 0.0, 0.0, "de"),
]
NRESULTS = {"cmp": 3, "zocrs": 5, "oct": 4}


def make_dump_fixtures():
    """Runs the reference's Dumper on each case with the SU file named relative to
    tests/golden (so gather.txt reads 'Input file = gather_<name>.su')."""
    for ci, (gname, (aph, apm, tau, az), model, lo, hi, v0, ain, bp, h0, method) in enumerate(DUMP_CASES):
        su = f"gather_{gname}.su"
        cdps = np.load(os.path.join(HERE, f"gather_{gname}.npz"))["cdps"]
        ns = int(np.load(os.path.join(HERE, f"gather_{gname}.npz"))["ns"])
        rng = np.random.default_rng(700 + ci)
        values = rng.normal(0.0, 1.0, (len(cdps), NRESULTS[model], ns)).astype(np.float32)
        # statistics of the magnitudes a run produces (efficiency, interpolations/s, traces,
        # seconds), so that the CSV text takes fixed and scientific forms
        stats = (rng.uniform(0.0, 1.0, (len(cdps), 5)) *
                 np.array([1.0, 3e9, 120.0, 2e-3, 0.25])).astype(np.float32)
        stats[:, 2] = np.round(stats[:, 2])
        out = os.path.join(HERE, f"dump_{gname}_{model}")
        shutil.rmtree(out, ignore_errors=True)
        os.makedirs(out)
        values.tofile(os.path.join(out, "values.f32"))
        stats.tofile(os.path.join(out, "stats.f32"))
        tmp = os.path.join("/tmp", f"golden_dump_{ci}")
        shutil.rmtree(tmp, ignore_errors=True)
        os.makedirs(tmp)
        args = ["dump", su, str(aph), str(apm), str(tau), str(az), model, str(len(lo))]
        args += [repr(x) for x in lo] + [repr(x) for x in hi]
        args += [str(v0), str(ain), str(bp), str(h0), tmp, f"gather_{gname}", method,
                 os.path.join(out, "values.f32"), os.path.join(out, "stats.f32")]
        res = subprocess.run([os.path.join(REF, "ref_host")] + args, cwd=HERE, check=True,
                             capture_output=True, text=True)
        written = res.stdout.strip()
        for f in sorted(os.listdir(written)):
            shutil.copy(os.path.join(written, f), os.path.join(out, f))
        json.dump({"gather": su, "aph": aph, "apm": apm, "tau": tau, "azimuth": az,
                   "model": model, "lower": lo, "upper": hi, "v0": v0, "ain": ain, "bpctg": bp,
                   "h0": h0, "method": method, "data_name": f"gather_{gname}",
                   "directory": os.path.basename(written.rstrip("/"))[:-len("YYYYMMDD_HHMM")],
                   "files": sorted(os.listdir(written))},
                  open(os.path.join(out, "args.json"), "w"), indent=1)


def main():
    if not os.path.exists(os.path.join(REF, "libref_kernel.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle", "ref")], check=True)
    L = ref_lib()
    rng = np.random.default_rng(1234)
    make_kernel_eval(L, rng)
    make_oct_functions(L, rng)
    line = synth.make_line(5, (3, 6), 1200.0, 32, 4000, seed=5, offset_min=0.0)
    # traces with h beyond aph and midpoints outside the allowed range are dropped by the
    # reader; the CDP entries are created before those filters
    make_gather_fixture("basic", line, aph=500.0, apm=0.0, tau=0.02, az=90.0, lowest="-",
                        highest="-", shuffle_seed=3)
    make_gather_fixture("filters", line, aph=400.0, apm=0.0, tau=0.022, az=90.0,
                        lowest="1010", highest="1040", shuffle_seed=4, bad_ns_rows=(2, 7))
    make_traveltime()
    make_dump_fixtures()
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
