// One (t0, candidate) semblance evaluation by one lane, shared by the grid, differential
// evolution and stretch-free kernels.
//
// The lane walks the midpoint's traces in order -- the order of the reference's
// per-thread loop (cmp/linear_search.cu:55-76 and twins) -- and computes the traveltime,
// t/dt, the sample index and the interpolation weight with the reference's IEEE
// operations in both variants: one ulp of t/dt is up to 6e-5 of interpolation weight
// at t/dt ~ 1000, so the index chain must round as the reference's does for the results
// to agree within 1e-5.  HSS_EXACT then reproduces an IEEE FP32 evaluation of the
// reference text bit for bit; HSS_FAST contracts the window sums (window.hpp).
#pragma once

#include "device_common.hpp"
#include "window.hpp"

namespace hss {

struct EvalCtx {
    const void *samples;
    const float *scale;
    const float *mid;
    const float *hoff;
    const float *hoff2;
    const int *midx;  // NULL: traces beg .. beg+cnt-1
    int ntr;          // traces in the whole array
    int nsamp_odd;    // (ntr * ns) & 1
    int left_cap;     // 2 + (kSeg + 2) / ns: traces past which a segment never reaches
    int beg, cnt;
    int ns, tau, w;
    float dt, inv_dt;  // inv_dt = RN(1/dt)
    IndexBounds ib;    // index_bounds(ns, tau)
    float m0, h0, apm;
};

template <int MODEL>
__device__ __forceinline__ TraceGeom load_geom(const float *__restrict__ mid,
                                               const float *__restrict__ hoff,
                                               const float *__restrict__ hoff2, int ti) {
    TraceGeom g{0.0f, 0.0f, 0.0f};
    if constexpr (MODEL == HSS_CMP) {
        g.h2 = hoff2[ti];
    } else if constexpr (MODEL == HSS_ZOCRS) {
        g.m = mid[ti];
        g.h2 = hoff2[ti];
    } else {
        g.m = mid[ti];
        g.h = hoff[ti];
    }
    return g;
}


// Wave segments.  For one trace, the 64 lanes of a wave (consecutive t0 samples, one
// candidate or nearby ones) read windows that overlap almost entirely: with moveout
// dt/dt0 = t0/t <= 1 the window starts of a wave span at most 64 samples.  Instead of
// (w+1) samples per lane through the vector memory path, the wave reads the kSeg samples
// from the lowest window start on with one 4- or 8-byte load per lane, decodes and scales
// them once into its own LDS slot, and each lane reads its window from there.  A trace
// whose windows do not all fit one segment takes a second pass from the lowest remaining
// start (same values, same order of traces per lane).  (Storing the sample differences
// next to the samples saves one subtraction per window value but doubles the LDS bytes
// each window reads, which costs more than it saves.)

// Orders the wave's LDS accesses around cross-lane reads and writes: LDS instructions of
// one wave execute in program order, so keeping the compiler from moving them suffices.
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Loads samples [g0, g0 + kSeg) of the flat sample array into seg[0..kSeg); `lim` is the
// largest lane offset that keeps both samples of a lane inside the array (16-bit storage:
// inside the array's last aligned word; then lim is even), at most 2*kWave.  Positions
// past the end of the array are left as they are; no window reads them.  16-bit storage
// needs g0 even (4-byte aligned words).  Every lane of the wave must call it.
template <int STORAGE>
__device__ __forceinline__ void load_segment(const EvalCtx &e, size_t g0, int lim, float sc,
                                             bool scaled, LdsFloat *seg, int lane) {
    const unsigned o = min((unsigned)(2 * lane), (unsigned)lim);
    float v0, v1;
    if constexpr (STORAGE == HSS_F32) {
        const float *p = (const float *)e.samples + g0;
        v0 = p[o];
        v1 = p[o + 1];
    } else {
        const uint32_t *p = (const uint32_t *)((const uint16_t *)e.samples + g0);
        const uint32_t word = p[o >> 1];
        if constexpr (STORAGE == HSS_F16) {
            v0 = (float)__builtin_bit_cast(_Float16, (uint16_t)(word & 0xffffu));
            v1 = (float)__builtin_bit_cast(_Float16, (uint16_t)(word >> 16));
        } else {
            v0 = __uint_as_float(word << 16);
            v1 = __uint_as_float(word & 0xffff0000u);
        }
    }
    if (STORAGE != HSS_F32 || scaled) {
        // sc is 1 when the line has no scales (exact)
        v0 = v0 * sc;
        v1 = v1 * sc;
    }
    wave_lds_order();  // the previous pass's windows have been read
    seg[o] = v0;
    seg[o + 1] = v1;
    wave_lds_order();
}

// SHIFT: stretch-free evaluation, t = t_n + n*dt (cmp/stretch_free.cu:54-56); `shift` is
// that n*dt.  Returns S and stack through the pointers and adds the wave's number of model
// rejections of active lanes (OCT notUsedCount) to *rejected of its lowest active lane (a
// caller sums *rejected over the wave's active lanes).
// The compiled window sizes (W > 0) compute sqrt and t/dt with the verified short
// sequences (sqrt_rn, div_rn); the runtime-w instantiation (W == 0) uses the library
// operations and is also what the launchers pick when the check failed for this dt.
// CPL candidates per lane share the trace loop, the geometry broadcast and, with SEG,
// the segment passes; `on[c]` false leaves candidate c out (its results are those of an
// empty window sum).
// SEG (W > 0 only): read windows through the wave's LDS segment `seg` (kSeg floats,
// private to the wave).  Every lane of the wave must call evaluate().  The segment pays
// for 16-bit storage, whose per-lane path would read each sample with its own 2-byte
// load; FP32 windows come in three 16-byte loads per lane, and there the segment's extra
// instructions cost more than the vector memory path it relieves.
// OCT_KEEP (OCT only): keep the offset-dependent half of the traveltime across
// consecutive traces with the same half offset (for lists grouped by offset; in the
// reference's order it only adds register pressure).
// FACTORED16 (16-bit per-lane windows): Window::add16's form; false only for a line
// stored without scales (the caller checks e.scale == nullptr).
template <int MODEL, int W, int VARIANT, int STORAGE, bool SHIFT, int CPL,
          bool SEG = (STORAGE != HSS_F32), bool OCT_KEEP = false, bool FACTORED16 = true>
__device__ __forceinline__ void evaluate_n(const EvalCtx &e, float t0, float t0sq,
                                           const Candidate (&k)[CPL], const bool (&on)[CPL],
                                           float shift, bool active, float *seg_generic,
                                           float (&S)[CPL], float (&stack)[CPL],
                                           unsigned *rejected) {
    LdsFloat *seg = (LdsFloat *)seg_generic;
    using Reader = SampleReader<STORAGE>;
    constexpr bool USE_SEG = SEG && W > 0;
    constexpr bool FASTIDX = W != 0;
    const bool scaled = (STORAGE != HSS_F32) && e.scale != nullptr;
    Reader rd{(decltype(Reader::p))e.samples};
    Window<W, VARIANT> acc[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) acc[c].reset();
    unsigned rej = 0;  // the wave's rejections (wave-uniform)
    const int lane = threadIdx.x & (kWave - 1);
    OctCand oc[CPL];
    if constexpr (MODEL == HSS_OCT) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) oc[c] = oct_cand(t0, e.h0, k[c]);
    }
    // OCT_KEEP: the displaced midpoint and the time terms of each candidate depend on the
    // trace's half offset, not on its midpoint, so they are kept while consecutive traces
    // share the offset (its bit pattern; a wave-uniform test).  Lists grouped by offset
    // then cost one aperture test and one multiply-add per trace for most traces; a
    // candidate's time terms are computed only when some lane keeps the trace for it.
    // ZO-CRS: (t0 + A*dm)^2 + B*dm*dm depends on the trace's midpoint, not its offset; it
    // is kept while consecutive traces share the midpoint (its bit pattern; a wave-uniform
    // test), which in the reference's (m, h) order is a whole CMP's run of traces.
    float zmid[CPL];
    uint32_t zkey = 0;
    bool zvalid = false;
    OctMid om[CPL];
    OctTime otm[CPL];
    uint32_t mid_key = 0;
    bool mid_valid = false;
    bool time_valid[CPL];
    // Traces go by in chunks of 64: lane i loads trace j0+i's index, geometry and scale
    // with one coalesced access, and the trace loop broadcasts them with v_readlane, so
    // no memory latency sits inside the per-trace loop.
    for (int j0 = 0; j0 < e.cnt; j0 += kWave) {
        const int mine = j0 + lane;
        const bool have = mine < e.cnt;
        const int ti_l = have ? (e.midx ? e.midx[e.beg + mine] : e.beg + mine) : 0;
        const TraceGeom g_l = have ? load_geom<MODEL>(e.mid, e.hoff, e.hoff2, ti_l)
                                   : TraceGeom{0.0f, 0.0f, 0.0f};
        const float sc_l = (scaled && have) ? e.scale[ti_l] : 1.0f;
        // OCT: the trace's own traveltime terms, once per trace (lane i: trace j0+i)
        OctTrace o_l{};
        if constexpr (MODEL == HSS_OCT) o_l = oct_trace<FASTIDX>(g_l.h, g_l.m, e.h0);
        const int nchunk = min(kWave, e.cnt - j0);
        // HSS_KAHAN: the trace loop runs in blocks of kKahanChunk traces (j0 is a multiple
        // of kWave, itself a multiple of the chunk), each block's sums flushed into the
        // compensated totals before the next (wave-uniform)
        constexpr int JB = VARIANT == HSS_KAHAN ? kKahanChunk : kWave;
        for (int jb = 0; jb < nchunk; jb += JB) {
        if constexpr (VARIANT == HSS_KAHAN) {
            if (j0 + jb > 0) {
#pragma unroll
                for (int c = 0; c < CPL; ++c) acc[c].flush(e.w);
            }
        }
        const int jend = min(jb + JB, nchunk);
        for (int jj = jb; jj < jend; ++jj) {
            const int ti = __builtin_amdgcn_readlane(ti_l, jj);
            TraceGeom tg{0.0f, 0.0f, 0.0f};
            OctTrace ot{};
            if constexpr (MODEL == HSS_OCT) {
                ot.h = bcast(o_l.h, jj);
                ot.m = bcast(o_l.m, jj);
                // OCT_KEEP broadcasts the offset-dependent terms only where it computes
                // with them
                if constexpr (!OCT_KEEP) {
                    bcast_oct_mid_terms(ot, o_l, jj);
                    bcast_oct_time_terms(ot, o_l, jj);
                }
            } else {
                tg.m = bcast(g_l.m, jj);
                tg.h2 = bcast(g_l.h2, jj);
            }
            const float sc = scaled ? bcast(sc_l, jj) : 1.0f;
            int kk[CPL];
            float f[CPL];
            bool use[CPL];
            if constexpr (MODEL == HSS_ZOCRS) {
                const uint32_t key = __float_as_uint(tg.m);
                if (!zvalid || key != zkey) {
                    const float dm = tg.m - e.m0;
#pragma unroll
                    for (int c = 0; c < CPL; ++c) zmid[c] = zocrs_mid_part(t0, dm, k[c]);
                    zkey = key;
                    zvalid = true;
                }
            }
            if constexpr (MODEL == HSS_OCT && OCT_KEEP) {
                const uint32_t key = __float_as_uint(ot.h);
                if (!mid_valid || key != mid_key) {
                    bcast_oct_mid_terms(ot, o_l, jj);
#pragma unroll
                    for (int c = 0; c < CPL; ++c) {
                        om[c] = oct_mid<FASTIDX>(ot, oc[c], e.m0);
                        time_valid[c] = false;
                    }
                    mid_key = key;
                    mid_valid = true;
                }
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    if (time_valid[c]) continue;
                    // exactly the complement of oct_traveltime_at's rejection (a NaN
                    // displaced midpoint is kept, as in the reference)
                    const bool need = !om[c].bad & !(fabsf(ot.m - om[c].mh) > e.apm) & on[c];
                    if (__builtin_amdgcn_ballot_w64(need & active) != 0) {
                        // this trace has the kept terms' offset, so its terms are theirs
                        bcast_oct_time_terms(ot, o_l, jj);
                        otm[c] = oct_time_terms<FASTIDX>(ot, oc[c], e.m0, om[c]);
                        time_valid[c] = true;
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                kk[c] = 0;
                f[c] = 0.0f;
                float t = 0.0f;  // defined for rejected traces too (its index is ignored)
                bool r;
                bool ok;
                if constexpr (MODEL == HSS_OCT && OCT_KEEP)
                    // until candidate c's time terms exist for this offset no active lane
                    // keeps the trace for it, so `ok` and `t` are not used
                    ok = oct_traveltime_at(om[c], otm[c], ot.m, e.apm, &t, &r);
                else if constexpr (MODEL == HSS_OCT)
                    ok = oct_traveltime<FASTIDX>(ot, oc[c], e.m0, e.apm, &t, &r);
                else if constexpr (MODEL == HSS_ZOCRS) {
                    r = false;
                    ok = zocrs_time_from<FASTIDX>(zmid[c], k[c].c * tg.h2, &t);
                } else
                    ok = traveltime<MODEL, FASTIDX>(tg, t0, t0sq, e.m0, e.h0, e.apm, k[c], &t, &r);
                // counted per wave (the callers add the count of active lanes only)
                rej += (unsigned)__builtin_popcountll(__builtin_amdgcn_ballot_w64(r & on[c] & active));
                if constexpr (SHIFT) t = t + shift;
                // (non-short-circuit: the index of a rejected trace is computed and ignored,
                // which keeps the lane predicates as masks instead of branches)
                const bool in = window_k<FASTIDX>(t, e.dt, e.inv_dt, e.ib, &kk[c], &f[c]);
                use[c] = ok & in & active & on[c];
            }
            const size_t trace0 = (size_t)ti * (size_t)e.ns;
            if constexpr (USE_SEG) {
                // Segment passes: a pass starts at the lowest window start among the
                // outermost pending lanes and serves every pending window that lies inside
                // the segment.  The first pass is written out on its own (normally it
                // serves everything); further passes loop until nothing is pending.  The
                // window that set the start always fits, so every pass makes progress.
                bool any = false;
#pragma unroll
                for (int c = 0; c < CPL; ++c) any = any | use[c];
                unsigned long long pending = __builtin_amdgcn_ballot_w64(any);
                if (pending == 0) continue;
                // samples from the start of trace ti to the end of the array, capped so
                // the 32-bit products below cannot overflow (only the cap's lower bound
                // matters: past kSeg + 2 samples the clamp is inactive)
                const int left = min(e.ntr - ti, e.left_cap);
                const int avail = left * e.ns;
                // one pass from window start k0 + tau, serving the windows in `serve`;
                // returns which of them did not fit
                auto pass = [&](int kmin_lane_value, unsigned long long live, const bool (&serve)[CPL],
                                bool (&rest)[CPL]) {
                    const int lo = __builtin_ctzll(live);
                    const int hi = 63 - __builtin_clzll(live);
                    // first sample of the segment: the lowest window start k - tau (>= 0)
                    const int b = min(__builtin_amdgcn_readlane(kmin_lane_value, lo),
                                      __builtin_amdgcn_readlane(kmin_lane_value, hi)) - e.tau;
                    const size_t gb = trace0 + (size_t)b;
                    const int par = STORAGE == HSS_F32 ? 0 : (int)(gb & 1);
                    // largest even (16-bit) lane offset with both samples in the array
                    const int last = STORAGE == HSS_F32 ? avail - 2 - b
                                                        : avail - 2 + e.nsamp_odd - b + par;
                    load_segment<STORAGE>(e, gb - (size_t)par, min(last, 2 * kWave), sc, scaled,
                                          seg, lane);
#pragma unroll
                    for (int c = 0; c < CPL; ++c) {
                        const int rel = kk[c] - e.tau - b + par;
                        const bool fits = (unsigned)rel <= (unsigned)(kSeg - 1 - W);
                        if (serve[c] & fits) acc[c].add_lds(seg + rel, f[c], e.w);
                        rest[c] = serve[c] & !fits;
                    }
                };
                int kmin = 0x7fffffff;
#pragma unroll
                for (int c = 0; c < CPL; ++c) kmin = use[c] ? min(kmin, kk[c]) : kmin;
                bool todo[CPL];
                pass(kmin, pending, use, todo);
                bool more = false;
#pragma unroll
                for (int c = 0; c < CPL; ++c) more = more | todo[c];
                pending = __builtin_amdgcn_ballot_w64(more);
                while (pending) {
                    int km = 0x7fffffff;
#pragma unroll
                    for (int c = 0; c < CPL; ++c) km = todo[c] ? min(km, kk[c]) : km;
                    bool nxt[CPL];
                    pass(km, pending, todo, nxt);
                    more = false;
#pragma unroll
                    for (int c = 0; c < CPL; ++c) {
                        todo[c] = nxt[c];
                        more = more | nxt[c];
                    }
                    pending = __builtin_amdgcn_ballot_w64(more);
                }
            } else if constexpr (STORAGE == HSS_F32) {
                // the trace's samples shifted by -tau: window k - tau .. k + tau + 1
                const float *row = (const float *)e.samples + trace0 - (size_t)e.tau;
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    if (use[c]) acc[c].add_row(row, (unsigned)kk[c], f[c], e.w);
                }
            } else if constexpr (W > 0) {
                // 16-bit samples, per-lane reads of whole words (add16) from the trace's
                // samples shifted by -tau, as a word row and its parity
                const size_t row = trace0 - (size_t)e.tau;
                const unsigned rpar = (unsigned)row & 1u;
                const uint32_t *roww = (const uint32_t *)((const uint16_t *)e.samples + (row - rpar));
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    if (use[c])
                        acc[c].template add16<STORAGE, FACTORED16>(roww, rpar, (unsigned)kk[c],
                                                                   f[c], sc);
                }
            } else {
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    if (use[c])
                        acc[c].add(rd, trace0 + (size_t)(kk[c] - e.tau), f[c], sc, scaled, e.w);
                }
            }
        }  // jj
        }  // jb
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) acc[c].reduce(e.w, &S[c], &stack[c]);
    // the wave's count goes to its lowest active lane
    const unsigned long long act = __builtin_amdgcn_ballot_w64(active);
    *rejected += (act != 0 && lane == __builtin_ctzll(act)) ? rej : 0u;
}

// One candidate per lane.
template <int MODEL, int W, int VARIANT, int STORAGE, bool SHIFT, bool SEG = (STORAGE != HSS_F32),
          bool OCT_KEEP = false, bool FACTORED16 = true>
__device__ __forceinline__ void evaluate(const EvalCtx &e, float t0, float t0sq,
                                         const Candidate &k, float shift, bool active,
                                         float *seg_generic, float *S, float *stack,
                                         unsigned *rejected) {
    const Candidate kk[1] = {k};
    const bool on[1] = {true};
    float s1[1], t1[1];
    evaluate_n<MODEL, W, VARIANT, STORAGE, SHIFT, 1, SEG, OCT_KEEP, FACTORED16>(
        e, t0, t0sq, kk, on, shift, active, seg_generic, s1, t1, rejected);
    *S = s1[0];
    *stack = t1[0];
}

// Context for midpoint g of a batch (fields that do not depend on g come from `base`).
// The midpoint a workgroup works on: launch slot `slot` (logical block / tiles per
// midpoint) takes midpoint order[slot] when the caller gives a launch order, else
// midpoint slot.  An order entry outside [0, n) gives -1: the workgroup has nothing to do.
// The midpoint, not the slot, indexes everything else (its traces, result rows, workspace
// rows and DE random stream), so the order changes the schedule and nothing else.
__device__ __forceinline__ int launch_midpoint(const int *order, int n, int slot) {
    if (order == nullptr) return slot;
    const int g = order[slot];
    return (unsigned)g < (unsigned)n ? g : -1;
}

__device__ __forceinline__ EvalCtx midpoint_ctx(const EvalCtx &base, const int *mbeg,
                                                const int *mcnt, const float *mm0, int g) {
    EvalCtx e = base;
    e.beg = mbeg[g];
    e.cnt = mcnt[g];
    e.m0 = mm0[g];
    return e;
}

// Whether midpoint e's trace list comes in runs of one half offset long enough for
// evaluate_n's OCT_KEEP to pay: four traces or more per run on average (lists grouped by
// offset; in the reference's (m, h) order runs are single traces, where keeping only adds
// work).  Runs are told apart by the offset's bit pattern, the test OCT_KEEP itself makes.
// Every lane of the wave must call it; the result is wave-uniform.
__device__ __forceinline__ bool oct_offset_runs_long(const EvalCtx &e) {
    const int lane = threadIdx.x & (kWave - 1);
    unsigned changes = 0;
    for (int j0 = 0; j0 < e.cnt; j0 += kWave) {
        const int j = j0 + lane;
        bool change = false;
        if (j < e.cnt) {
            const int ti = e.midx ? e.midx[e.beg + j] : e.beg + j;
            change = j == 0;
            if (j > 0) {
                const int tp = e.midx ? e.midx[e.beg + j - 1] : e.beg + j - 1;
                change = __float_as_uint(e.hoff[ti]) != __float_as_uint(e.hoff[tp]);
            }
        }
        changes += (unsigned)__builtin_popcountll(__builtin_amdgcn_ballot_w64(change));
    }
    return 4u * changes <= (unsigned)e.cnt;
}

}  // namespace hss
