// CMP grid search with the traces staged in LDS for the whole workgroup.
//
// The grid kernel (grid_impl.hpp) lets every wave read its own windows: FP32 lanes with
// 16-byte loads (the texture-address unit is then the busiest unit after the VALU), 16-bit
// storage through a per-wave LDS segment that each wave loads and decodes for itself.
// Here the four waves of a workgroup (one 64-sample t0 tile of a midpoint) take their
// candidates in rounds -- round r: wave v evaluates the CPL candidates starting at
// (4r + v)*CPL, the same candidates in the same order as the grid kernel's wave v -- and
// walk the midpoint's traces together, a block at a time.  For each trace of a block the
// workgroup first loads, decodes and scales into LDS the samples that any window of the
// round can read, then every lane reads its windows from there.  (Blocks of 32 traces;
// HSS_KAHAN: 16.)
//
// The staged span of a trace follows from monotonicity (CMP: t = sqrt(t0^2 + c h^2) and
// every operation of the index chain are nondecreasing in t0 and in c): the smallest
// window start of the round is at the tile's first t0 and the round's smallest c, the
// largest at its last t0 and largest c.  A block with a trace whose span exceeds the stage
// takes direct reads instead.  Each lane adds its traces in the same order with the same
// operations on the same values as the grid kernel, so the results are the same bits
// (HSS_KAHAN's chunk boundaries are the same 16-trace blocks).
#pragma once

#include "grid_impl.hpp"

namespace hss {

// Samples staged per trace: 128, loaded by the first two waves (one sample per thread).
// The C5 line's spans (a 64-sample t0 tile plus the window and the moveout across the
// round's candidates) fit; a 256-sample stage was 4.7% slower (more LDS per workgroup, one
// workgroup fewer per CU).
constexpr int kStageCap = 128;
static_assert(kStageCap % kWave == 0 && kStageCap <= kGridThreads, "whole waves load the stage");
// Traces per stage: 32 (16 KB of LDS; 48 was no faster, 16 was 3% slower on the C5 line),
// and for HSS_KAHAN its compensation chunk, since its sums are flushed at those fixed
// trace blocks.
template <int VARIANT>
constexpr int stage_traces() { return VARIANT == HSS_KAHAN ? kKahanChunk : 32; }

// Register budget: five waves per SIMD (96 VGPRs), 4% faster than the four the compiler
// picks by itself on the C5 line and faster than six; HSS_KAHAN keeps the grid kernel's
// budget.  At 96 registers nothing may be spilled: a spill slot is per-lane memory that
// stays in L2 beside the traces (5 resident midpoints per XCD at five waves), and the C5
// launch then read 63 GB and wrote 46 GB instead of about 3 GB.  Hence the running best in
// LDS, the uniform values in scalar registers and thread_index_here().
template <int VARIANT, int CPL>
constexpr int staged_min_waves() {
    return VARIANT == HSS_KAHAN ? grid_min_waves<VARIANT, CPL>() : 5;
}

// The sample index (int)(t/dt) of the index chain, clamped into [-1, 2^24] first (the
// span only needs the order; lanes use indices inside (lo, hi) of index_bounds()).
template <bool FAST>
__device__ __forceinline__ int span_index(float t, float dt, float inv) {
    float ti;
    if constexpr (FAST) ti = div_rn(t, dt, inv);
    else ti = t / dt;
    ti = fminf(fmaxf(ti, -1.0f), 16777216.0f);  // a NaN becomes -1
    return (int)ti;
}

// The running best of each thread, in LDS.
struct RunBest {
    float best[kGridThreads];
    float stack[kGridThreads];
    int idx[kGridThreads];
};

template <int W, int VARIANT, int STORAGE, int CPL>
__global__ __launch_bounds__(kGridThreads)
__attribute__((amdgpu_waves_per_eu(staged_min_waves<VARIANT, CPL>(), 8))) void grid_staged_kernel(
    GridArgs a) {
    using Reader = SampleReader<STORAGE>;
    constexpr bool FASTIDX = W != 0;
    const int lane = threadIdx.x & (kWave - 1);
    // wave-uniform values go through readfirstlane so that they live in scalar registers
    // (the round's candidates are then scalar loads; gfx950 has no scalar float ALU, so a
    // uniform float otherwise takes a vector register for the whole kernel)
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
    const int blk = logical_block(blockIdx.x, gridDim.x);
    const int slot = blk / a.ntiles;
    const int tile = blk - slot * a.ntiles;
    const int g = launch_midpoint(a.morder, a.nmid, slot);
    if (g < 0) return;  // (the whole workgroup)
    const int s = tile * kWave + lane;
    const int ns = a.e.ns;
    const bool active = s < ns;
    const float t0 = (float)s * a.e.dt;
    const float t0sq = t0 * t0;
    const EvalCtx e = midpoint_ctx(a.e, a.mbeg, a.mcnt, a.mm0, g);
    const bool scaled = (STORAGE != HSS_F32) && e.scale != nullptr;
    Reader rd{(decltype(Reader::p))e.samples};
    // the tile's first and last t0 (the span's corners)
    const float t0a = uniform_f((float)(tile * kWave) * e.dt);
    const float t0b = uniform_f((float)min(tile * kWave + kWave - 1, ns - 1) * e.dt);
    const float t0a_sq = uniform_f(t0a * t0a), t0b_sq = uniform_f(t0b * t0b);
    // the midpoint's [ns][P] slabs of the optional cubes (the host keeps ns * P below 2^32)
    float *const sembl_g = a.sembl ? a.sembl + (size_t)g * ns * a.P : nullptr;
    float *const stack_g = a.stack ? a.stack + (size_t)g * ns * a.P : nullptr;

    constexpr int kStageTraces = stage_traces<VARIANT>();
    static_assert(kStageTraces <= kWave, "a block's traces are one wave's lanes");
    __shared__ float stage[kStageTraces][kStageCap];
    // the lane's running best (first strict maximum over its candidates in order), kept in
    // LDS rather than in registers, which the window sums of two candidates use up; each
    // thread reads and writes only its own slots
    __shared__ RunBest run;
    run.best[threadIdx.x] = -1.0f;
    run.stack[threadIdx.x] = 0.0f;
    run.idx[threadIdx.x] = -1;

    const int per_round = kGridWaves * CPL;
    const int rounds = (a.P + per_round - 1) / per_round;
    for (int r = 0; r < rounds; ++r) {
        const int q0 = (r * kGridWaves + wave) * CPL;
        Candidate k[CPL];
        bool on[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            on[c] = q0 + c < a.P;
            k[c] = load_candidate<HSS_CMP>(a.params, on[c] ? q0 + c : min(q0, a.P - 1));
        }
        // the round's range of c (all waves: the same candidates, the same values)
        float cmin = INFINITY, cmax = -INFINITY;
        const int qa = r * per_round, qb = min(a.P, qa + per_round);
        for (int q = qa; q < qb; ++q) {
            const float c = a.params[q];
            cmin = fminf(cmin, c);
            cmax = fmaxf(cmax, c);
        }
        Window<W, VARIANT> acc[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[c].reset();

        for (int j0 = 0; j0 < e.cnt; j0 += kStageTraces) {
            const int nj = min(kStageTraces, e.cnt - j0);
            // lane jj < nj: trace j0 + jj's geometry, scale and staged span
            const int ln = thread_index_here() & (kWave - 1);
            const bool have = ln < nj;
            const int ti_l = have ? (e.midx ? e.midx[e.beg + j0 + ln] : e.beg + j0 + ln) : 0;
            const float h2_l = have ? e.hoff2[ti_l] : 0.0f;
            const float sc_l = (scaled && have) ? e.scale[ti_l] : 1.0f;
            int base_l = 0, fit_l = 0;
            {
                TraceGeom tg{0.0f, 0.0f, 0.0f};
                tg.h2 = h2_l;
                Candidate lo{cmin, 0.0f, 0.0f}, hi{cmax, 0.0f, 0.0f};
                float ta, tb;
                bool r0;
                traveltime<HSS_CMP, FASTIDX>(tg, t0a, t0a_sq, e.m0, e.h0, e.apm, lo, &ta, &r0);
                traveltime<HSS_CMP, FASTIDX>(tg, t0b, t0b_sq, e.m0, e.h0, e.apm, hi, &tb, &r0);
                const int ka = span_index<FASTIDX>(ta, e.dt, e.inv_dt);
                const int kb = span_index<FASTIDX>(tb, e.dt, e.inv_dt);
                base_l = max(ka - e.tau, 0);
                // windows run from k - tau to k + tau + 1
                fit_l = (have && kb + e.tau + 2 - base_l <= kStageCap) ? 1 : 0;
            }
            if constexpr (VARIANT == HSS_KAHAN) {
                if (j0 > 0) {
#pragma unroll
                    for (int c = 0; c < CPL; ++c) acc[c].flush(e.w);
                }
            }
            // The block is staged when every one of its traces fits; otherwise it takes the
            // direct reads (one read path per loop, chosen for the whole workgroup: every
            // wave computes the same spans).
            const unsigned long long want =
                nj >= kWave ? ~0ull : ((1ull << nj) - 1ull);
            const bool staged = (__builtin_amdgcn_ballot_w64(fit_l != 0) & want) == want;
            if (staged && STORAGE == HSS_F32) {
                // FP32: the samples go from global memory straight into LDS
                // (global_load_lds_dword: lane l of wave v < kStageCap/64 writes
                // stage[jj][64 v + l]), no registers in between; a lane past the trace's end
                // loads nothing (its slot is never read)
                __syncthreads();  // every wave has read the previous staged block
                for (int jj = 0; jj < nj; ++jj) {
                    const int base = __builtin_amdgcn_readlane(base_l, jj);
                    const int ti = __builtin_amdgcn_readlane(ti_l, jj);
                    const int i = base + (int)threadIdx.x;
                    if (i < ns && wave < kStageCap / kWave)
                        __builtin_amdgcn_global_load_lds(
                            (const float *)e.samples + (size_t)ti * (size_t)ns + (size_t)i,
                            (__attribute__((address_space(3))) void *)&stage[jj][wave * kWave], 4,
                            0, 0);
                }
                __syncthreads();  // (the compiler waits for the loads before the barrier)
            } else if (staged) {
                // 16-bit: all of the block's loads first, then the decoded and scaled values
                // into LDS (one wait, not one per trace)
                float vals[kStageTraces];
#pragma unroll
                for (int jj = 0; jj < kStageTraces; ++jj) {
                    vals[jj] = 0.0f;
                    if (jj >= nj) continue;
                    const int base = __builtin_amdgcn_readlane(base_l, jj);
                    const int ti = __builtin_amdgcn_readlane(ti_l, jj);
                    const int i = base + (int)threadIdx.x;  // within the trace: no window reads past ns
                    if (i < ns && (int)threadIdx.x < kStageCap) vals[jj] = rd((size_t)ti * (size_t)ns + (size_t)i);
                }
                __syncthreads();  // every wave has read the previous staged block
#pragma unroll
                for (int jj = 0; jj < kStageTraces; ++jj) {
                    if (jj >= nj) continue;
                    float v = vals[jj];
                    // as the wave segments do (a sample past the trace stays 0 and is never read)
                    if (STORAGE != HSS_F32 || scaled) v = v * bcast(sc_l, jj);
                    if ((int)threadIdx.x < kStageCap) stage[jj][threadIdx.x] = v;
                }
                __syncthreads();
            }
            // the lane's sample index, weight and inclusion for trace jj of the block
            auto index = [&](int jj, int (&kk)[CPL], float (&f)[CPL], bool (&use)[CPL]) {
                TraceGeom tg{0.0f, 0.0f, 0.0f};
                tg.h2 = bcast(h2_l, jj);
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    float t = 0.0f;
                    bool rj;
                    const bool ok = traveltime<HSS_CMP, FASTIDX>(tg, t0, t0sq, e.m0, e.h0, e.apm,
                                                                 k[c], &t, &rj);
                    const bool in = window_k<FASTIDX>(t, e.dt, e.inv_dt, e.ib, &kk[c], &f[c]);
                    use[c] = ok & in & active & on[c];
                }
            };
            if (staged) {
                for (int jj = 0; jj < nj; ++jj) {
                    // the staged row shifted so that sample index k reads its window start
                    // k - tau (a workgroup-uniform base: one address operation per window)
                    const int shift = __builtin_amdgcn_readlane(base_l, jj) + e.tau;
                    const uint32_t row = (uint32_t)__builtin_amdgcn_readfirstlane(
                        (int)((uint32_t)(size_t)((const LdsFloat *)&stage[jj][0]) -
                              4u * (uint32_t)shift));
                    int kk[CPL];
                    float f[CPL];
                    bool use[CPL];
                    index(jj, kk, f, use);
#pragma unroll
                    for (int c = 0; c < CPL; ++c) {
                        if (use[c])
                            acc[c].add_lds((const LdsFloat *)(size_t)(row + 4u * (uint32_t)kk[c]),
                                           f[c], e.w);
                    }
                }
            } else {
                for (int jj = 0; jj < nj; ++jj) {
                    const size_t trace0 = (size_t)__builtin_amdgcn_readlane(ti_l, jj) * (size_t)ns;
                    const float sc = bcast(sc_l, jj);
                    int kk[CPL];
                    float f[CPL];
                    bool use[CPL];
                    index(jj, kk, f, use);
#pragma unroll
                    for (int c = 0; c < CPL; ++c) {
                        if (!use[c]) continue;
                        if constexpr (STORAGE == HSS_F32)
                            acc[c].add_row((const float *)e.samples + trace0 - (size_t)e.tau,
                                           (unsigned)kk[c], f[c], e.w);
                        else
                            acc[c].add(rd, trace0 + (size_t)(kk[c] - e.tau), f[c], sc, true, e.w);
                    }
                }
            }
        }
        float S[CPL], stk[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[c].reduce(e.w, &S[c], &stk[c]);
        if (active) {
            const int t = thread_index_here();
            const uint32_t cube_row = (uint32_t)(tile * kWave + (t & (kWave - 1))) * (uint32_t)a.P;
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                if (!on[c]) continue;
                const int q = q0 + c;
                if (sembl_g) sembl_g[cube_row + (uint32_t)q] = S[c];
                if (stack_g) stack_g[cube_row + (uint32_t)q] = stk[c];
                if (S[c] > run.best[t]) {
                    run.best[t] = S[c];
                    run.stack[t] = stk[c];
                    run.idx[t] = q;
                }
            }
        }
    }

    // the four waves' bests of each t0, as merge_bests merges them (the largest value, the
    // smallest candidate index among equal values)
    __syncthreads();
    if (wave != 0 || !active) return;
    const int ln = thread_index_here() & (kWave - 1);
    float B = -1.0f, BS = 0.0f;
    int BI = -1;
#pragma unroll
    for (int v = 0; v < kGridWaves; ++v) {
        const int iv = run.idx[v * kWave + ln];
        const float sv = run.best[v * kWave + ln];
        if (iv >= 0 && (sv > B || (sv == B && iv < BI))) {
            B = sv;
            BS = run.stack[v * kWave + ln];
            BI = iv;
        }
    }
    const int so = tile * kWave + ln;  // the lane's sample
    float *res = a.result + (size_t)g * 3 * ns;
    res[so] = B;
    res[ns + so] = BI >= 0 ? BS : 0.0f;
    res[2 * ns + so] = BI >= 0 ? 2 / sqrtf(load_candidate<HSS_CMP>(a.params, BI).c) : 0.0f;
}

// Host launcher: one workgroup per (launch slot, t0 tile) (a.tt must be 1).
int launch_grid_cmp_staged(const GridArgs &a, int n_slots, int variant, int storage, int cpl,
                           hipStream_t stream);

}  // namespace hss
