// Differential evolution over the traveltime parameters, one population per (midpoint, t0),
// every generation inside one launch.
//
// Reference behaviour (common/differential_evolution.cu:19-155 driven by
// DifferentialEvolutionAlgorithm.cpp:21-81): start x = lo + U*(hi-lo); evaluate; then
// for G-1 generations mutate per parameter from donors floor(U*(NP-1)) with F = 0.85 and
// clamp, cross over with CR = 0.5 and a forced index l = floor(U*(D-1)), evaluate the
// trials and keep a trial on strict '>' of its semblance; finally report the first best
// individual (cmp|zocrs|oct/differential_evolution.cu selectBestIndividuals*).  The
// reference draws from curand seeded with time(); here the uniforms come from
// Philox4x32-10 with counter {s, stream, individual, gen<<4 | block} and key = seed, the
// layout the CPU oracle uses, so runs are reproducible and checkable.
//
// MI355X mapping: one workgroup = (midpoint, 64 consecutive t0); lane l of every wave
// owns population s = tile*64 + l; wave v of V handles individuals v, v+V, ...  The
// populations live in a caller-provided workspace, double buffered (trials of generation
// g read generation g's individuals only), lane-fastest so every access is a 256-byte
// coalesced row; one workgroup barrier separates generations.
#pragma once

#include "eval_compact.hpp"
#include "grid_impl.hpp"
#include "launch_util.hpp"

namespace hss {

// Launch shape, chosen by timing a ZO-CRS line (200 midpoints, ~930 traces each, NP 32,
// G 50, BF16) and an OCT line (300 midpoints, ~2150 traces each, NP 48, G 40, FP16):
//  * de_cpl: individuals per lane in one trace loop.  Two interleave two independent
//    traveltime/window chains, which hides the latency of the long dependent OCT and
//    ZO-CRS sequences better than more resident waves do (OCT -10%, ZO-CRS -6%; three
//    were slower for ZO-CRS and no better for OCT).
//  * de_min_occupancy: the register budget, as waves per SIMD the compiler must fit
//    (OCT 6: 80 VGPRs; ZO-CRS 8: 64; CMP 8: 64).  Since the traveltime terms are kept
//    across traces (offset groups, midpoint runs) the trace loop is short and waits on
//    memory, and more resident waves pay for the spills of the per-generation code:
//    OCT 5/6/7/8 waves measured 2435/2334/2356/2479 ms, ZO-CRS 6/7/8 559/557/536 ms.
//  * de_waves: waves per workgroup, so that whole workgroups fill a CU's wave slots at
//    that occupancy (OCT 8 x 3 workgroups = 24 per CU; ZO-CRS and CMP 8 x 4 = 32).  For OCT
//    8 waves were 1.4% faster than 4 x 6 workgroups: a workgroup's generations run half as
//    long, so the launch's tail (the last workgroups on each XCD, about a quarter of a
//    second each at 4 waves) is shorter.
// Kahan windows (twice the accumulators), windows above 11 samples and the runtime-w
// kernel keep one individual per lane and the compiler's own register budget.
template <int MODEL>
constexpr int de_waves() { return 8; }
template <int MODEL, int W, int VARIANT>
constexpr bool de_tuned() { return VARIANT != HSS_KAHAN && W != 0 && W <= 11; }
template <int MODEL, int W, int VARIANT>
constexpr int de_cpl() { return de_tuned<MODEL, W, VARIANT>() && MODEL != HSS_CMP ? 2 : 1; }
template <int MODEL, int W, int VARIANT>
constexpr int de_min_occupancy() {
    if (!de_tuned<MODEL, W, VARIANT>()) return 1;
    return MODEL == HSS_OCT ? 6 : 8;
}
constexpr int kDeWavesMax = 8;
// Every lane carries its own individual, so a trace's windows spread over far more than
// one wave segment; DE reads 16-bit windows per lane (Window::add16) instead.
constexpr bool kDeSegments = false;
// OCT: DE keeps the offset-dependent half of the traveltime across consecutive traces
// with the same half offset (the host layer groups the lists of the non-exact variants by
// offset).  Any order gives the bits of computing every trace afresh; in the reference's
// (m, h) order, where nothing is shared, the launch is about 11% slower than a kernel that
// does not keep them (measured on the C4 line).
constexpr bool kDeOctKeep = true;
// OCT: the aperture test of a run of traces first, then each lane's kept traces only
// (eval_compact.hpp); same bits as evaluate_n.  Not for HSS_KAHAN.
template <int MODEL, int VARIANT>
constexpr bool de_compact() { return MODEL == HSS_OCT && kDeOctKeep && VARIANT != HSS_KAHAN; }
constexpr float kDeF = 0.85f;   // F_FAC, common/include/gpu/interface.h:5
constexpr float kDeCR = 0.5f;   // CR, common/include/gpu/interface.h:6

struct DeArgs {
    EvalCtx e;
    const int *mbeg;
    const int *mcnt;
    const float *mm0;
    const int *morder;  // launch order (launch_midpoint) or NULL
    int nmid;
    int NP, G;
    int gen_begin, gen_end;  // generations run by this launch; the workspace carries the rest
    float lo[3], hi[3];
    uint32_t key0, key1, stream_base;
    int ntiles;
    float *ws;        // [n*ntiles][2][NP][D+2][64]
    float *result;    // [n][2+D][ns]
    float *pop_x;     // [n][ns][NP][D] or NULL
    float *pop_fx;    // [n][ns][NP][2] or NULL
    unsigned long long *not_used;  // [n] OCT rejections, accumulated, or NULL
};

// Philox4x32-10 (Salmon et al., SC'11).
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// 24-bit uniform in (0, 1].
__device__ __forceinline__ float u01(uint32_t x) {
    return ((float)(x >> 8) + 1.0f) * 5.9604644775390625e-08f;
}

// The first nq uniforms of (s, stream, i, gen).
template <int NQ>
__device__ __forceinline__ void de_uniforms(const DeArgs &a, uint32_t s, uint32_t stream, int i,
                                            int gen, float *u) {
#pragma unroll
    for (int b = 0; b * 4 < NQ; ++b) {
        uint32_t r[4];
        philox4x32_10(s, stream, (uint32_t)i, ((uint32_t)gen << 4) | (uint32_t)b, a.key0, a.key1, r);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (b * 4 + q < NQ) u[b * 4 + q] = u01(r[q]);
    }
}

// Storage-form candidate of an individual in v space (cmp|zocrs|oct/differential_evolution.cu:
// c = 4/v^2, the other parameters as they are).
template <int MODEL>
__device__ __forceinline__ Candidate de_candidate(const float *x) {
    Candidate k{0.0f, 0.0f, 0.0f};
    const float v = x[0];
    k.c = 4.0f / (v * v);
    if constexpr (MODEL == HSS_ZOCRS) {
        k.p1 = x[1];
        k.p2 = x[2];
    } else if constexpr (MODEL == HSS_OCT) {
        k.p1 = x[1];
    }
    return k;
}

template <int MODEL, int W, int VARIANT, int STORAGE>
__global__ __launch_bounds__(kDeWavesMax * kWave)
__attribute__((amdgpu_waves_per_eu(de_min_occupancy<MODEL, W, VARIANT>(), 8))) void de_semblance_kernel(DeArgs a) {
    constexpr int kDeWaves = de_waves<MODEL>();
    constexpr int D = MODEL == HSS_CMP ? 1 : (MODEL == HSS_ZOCRS ? 3 : 2);
    constexpr int REC = D + 2;  // D parameters, S, stack
    const int lane = threadIdx.x & (kWave - 1);
    // The wave index is uniform (readfirstlane), so the individual loops, their bounds and
    // the donors' rows are scalar work.  What outlives an evaluation -- the trials, the
    // lane's workspace column, its sample -- is kept in LDS or recomputed from the thread
    // index where it is used (thread_index_here), so that the registers of the budget go
    // to the trace walk and nothing is spilled: a spill slot is per-lane memory that takes
    // L2 from the traces (this kernel waits on its trace reads).
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
    const int blk = logical_block(blockIdx.x, gridDim.x);
    const int slot = blk / a.ntiles;
    const int tile = blk - slot * a.ntiles;
    const int g = launch_midpoint(a.morder, a.nmid, slot);
    if (g < 0) return;
    const int s = tile * kWave + lane;
    const int ns = a.e.ns;
    const bool active = s < ns;
    const float t0 = (float)s * a.e.dt;
    const float t0sq = t0 * t0;
    const EvalCtx e = midpoint_ctx(a.e, a.mbeg, a.mcnt, a.mm0, g);
    const uint32_t stream = a.stream_base + (uint32_t)g;
    const int NP = a.NP;
    float *const base = a.ws + ((size_t)g * a.ntiles + tile) * 2 * NP * REC * kWave;
    auto rec = [&](int buf, int i, int f) -> float * {
        return base + ((size_t)(buf * NP + i) * REC + f) * kWave +
               (thread_index_here() & (kWave - 1));
    };
    // the lane's sample, as the counter of its random stream
    auto sample_here = [&]() {
        return (uint32_t)(tile * kWave + (thread_index_here() & (kWave - 1)));
    };

    __shared__ float seg_lds[kDeWaves][kSegFloats];
    __shared__ OctTraceRec trace_recs[kDeWaves][kWave];
    constexpr int CPL = de_cpl<MODEL, W, VARIANT>();
    // the individuals a lane evaluates together, from their drawing to their selection
    __shared__ float held[CPL * D][kDeWaves * kWave];
    uint64_t rejected = 0;  // the wave's rejections (uniform)
    auto eval = [&](const Candidate (&kc)[CPL], const bool (&on)[CPL], float (&S)[CPL],
                    float (&stk)[CPL]) {
        unsigned rej = 0;  // on the wave's lowest active lane
        if constexpr (de_compact<MODEL, VARIANT>())
            evaluate_oct_compact<W, VARIANT, STORAGE, false, CPL>(
                e, t0, kc, on, 0.0f, active, trace_recs[wave], S, stk, &rej);
        else if (STORAGE == HSS_BF16 && e.scale == nullptr)
            // a BF16 line stored without scales: its windows summed as FP32 windows (one
            // form per code path, chosen outside the trace loop)
            evaluate_n<MODEL, W, VARIANT, STORAGE, false, CPL, kDeSegments, kDeOctKeep, false>(
                e, t0, t0sq, kc, on, 0.0f, active, seg_lds[wave], S, stk, &rej);
        else
            evaluate_n<MODEL, W, VARIANT, STORAGE, false, CPL, kDeSegments, kDeOctKeep>(
                e, t0, t0sq, kc, on, 0.0f, active, seg_lds[wave], S, stk, &rej);
        if constexpr (MODEL == HSS_OCT) {
            // the wave's total, summed over the lanes that hold a count (readlane: uniform)
            for (uint64_t m = __builtin_amdgcn_ballot_w64(rej != 0u); m; m &= m - 1)
                rejected += (uint32_t)__builtin_amdgcn_readlane((int)rej, __builtin_ctzll(m));
        }
    };
    // Generation g's population lives in buffer g & 1 of the workspace, so a run split
    // over several launches ([gen_begin, gen_end) each) gives the same result as one.
    // generation 0: the start population (startPopulations) and its fitness
    // A wave evaluates CPL individuals i0, i0+V, ... at a time (one trace loop for all);
    // a slot past NP is evaluated on individual i0's data and discarded.
    for (int i0 = wave; a.gen_begin == 0 && i0 < NP; i0 += kDeWaves * CPL) {
        Candidate kc[CPL];
        bool on[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const int i = i0 + c * kDeWaves;
            on[c] = i < NP;
            float u[4];
            de_uniforms<D>(a, sample_here(), stream, on[c] ? i : i0, 0, u);
            float x[D];
            const int t = thread_index_here();
#pragma unroll
            for (int d = 0; d < D; ++d) {
                x[d] = a.lo[d] + u[d] * (a.hi[d] - a.lo[d]);
                held[c * D + d][t] = x[d];
            }
            kc[c] = de_candidate<MODEL>(x);
        }
        float S[CPL], stk[CPL];
        eval(kc, on, S, stk);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            if (!on[c]) continue;
            const int i = i0 + c * kDeWaves;
            const int t = thread_index_here();
#pragma unroll
            for (int d = 0; d < D; ++d) *rec(0, i, d) = held[c * D + d][t];
            *rec(0, i, D) = S[c];
            *rec(0, i, D + 1) = stk[c];
        }
    }
    __syncthreads();

    for (int gen = a.gen_begin > 1 ? a.gen_begin : 1; gen < a.gen_end; ++gen) {
        const int cur = (gen - 1) & 1;
        for (int i0 = wave; i0 < NP; i0 += kDeWaves * CPL) {
            Candidate kc[CPL];
            bool on[CPL];
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const int i1 = i0 + c * kDeWaves;
                on[c] = i1 < NP;
                const int i = on[c] ? i1 : i0;
                constexpr int NQ = 4 * D + 1;
                float u[NQ];
                de_uniforms<NQ>(a, sample_here(), stream, i, gen, u);
                int np1i = NP - 1;  // donors are drawn from NP - 1 (converted where used)
                __asm__ volatile("" : "+s"(np1i));
                const float np1 = (float)np1i;
                // mutatePopulations (common/differential_evolution.cu:43-89)
                float mutant[D];
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const unsigned r1 = (unsigned)(u[3 * d] * np1);
                    const unsigned r2 = (unsigned)(u[3 * d + 1] * np1);
                    const unsigned r3 = (unsigned)(u[3 * d + 2] * np1);
                    float nv = *rec(cur, r1, d) + kDeF * (*rec(cur, r2, d) - *rec(cur, r3, d));
                    // the bounds as opaque scalars: canonicalized at each use (one operation)
                    // rather than once into vector registers held through the evaluation
                    int hb = __builtin_amdgcn_readfirstlane(__float_as_int(a.hi[d]));
                    int lb = __builtin_amdgcn_readfirstlane(__float_as_int(a.lo[d]));
                    __asm__ volatile("" : "+s"(hb), "+s"(lb));
                    nv = fminf(nv, __int_as_float(hb));
                    nv = fmaxf(nv, __int_as_float(lb));
                    mutant[d] = nv;
                }
                // crossoverPopulations (common/differential_evolution.cu:91-122)
                const unsigned l = (unsigned)(u[3 * D] * (float)(D - 1));
                float trial[D];
                const int t = thread_index_here();
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const float r = u[3 * D + 1 + d];
                    trial[d] = (r > kDeCR && l != (unsigned)d) ? *rec(cur, i, d) : mutant[d];
                    held[c * D + d][t] = trial[d];
                }
                kc[c] = de_candidate<MODEL>(trial);
            }
            float S[CPL], stk[CPL];
            eval(kc, on, S, stk);
            // nextGeneration (common/differential_evolution.cu:124-155): strict '>'
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                if (!on[c]) continue;
                const int i = i0 + c * kDeWaves;
                const bool take = S[c] > *rec(cur, i, D);
                const int t = thread_index_here();
#pragma unroll
                for (int d = 0; d < D; ++d)
                    *rec(cur ^ 1, i, d) = take ? held[c * D + d][t] : *rec(cur, i, d);
                *rec(cur ^ 1, i, D) = take ? S[c] : *rec(cur, i, D);
                *rec(cur ^ 1, i, D + 1) = take ? stk[c] : *rec(cur, i, D + 1);
            }
        }
        __syncthreads();
    }

    if (MODEL == HSS_OCT && a.not_used) {
        if ((thread_index_here() & (kWave - 1)) == 0 && rejected)
            atomicAdd(a.not_used + g, (unsigned long long)rejected);
    }
    if (a.gen_end < a.G || wave != 0 || !active) return;
    const int cur = (a.G - 1) & 1;
    // selectBestIndividuals*: first strict maximum of S over the population
    float best = -1.0f, bstack = 0.0f;
    int bi = -1;
    for (int i = 0; i < NP; ++i) {
        const float S = *rec(cur, i, D);
        if (S > best) {
            best = S;
            bstack = *rec(cur, i, D + 1);
            bi = i;
        }
    }
    float *res = a.result + (size_t)g * (2 + D) * ns;
    res[s] = best;
    res[ns + s] = bi >= 0 ? bstack : 0.0f;
#pragma unroll
    for (int d = 0; d < D; ++d) res[(2 + d) * ns + s] = bi >= 0 ? *rec(cur, bi, d) : 0.0f;
    if (a.pop_x) {
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int d = 0; d < D; ++d)
                a.pop_x[(((size_t)g * ns + s) * NP + i) * D + d] = *rec(cur, i, d);
    }
    if (a.pop_fx) {
        for (int i = 0; i < NP; ++i) {
            a.pop_fx[(((size_t)g * ns + s) * NP + i) * 2] = *rec(cur, i, D);
            a.pop_fx[(((size_t)g * ns + s) * NP + i) * 2 + 1] = *rec(cur, i, D + 1);
        }
    }
}

template <int MODEL>
int launch_de(const DeArgs &a, int n_mid, int variant, int storage, hipStream_t stream) {
    int st = HSS_OK;
    const unsigned blocks = (unsigned)n_mid * (unsigned)a.ntiles;
    if (blocks == 0) return HSS_OK;
    int rc = dispatch_storage(storage, [&](auto S) {
        st = dispatch_variant(variant, [&](auto V) {
            dispatch_window(a.e.w, fast_index_verified(a.e.dt, a.e.tau, stream), [&](auto Wc) {
                hipLaunchKernelGGL((de_semblance_kernel<MODEL, decltype(Wc)::value,
                                                        decltype(V)::value, decltype(S)::value>),
                                   dim3(blocks), dim3(de_waves<MODEL>() * kWave), 0, stream, a);
            });
        });
    });
    if (rc != HSS_OK) return rc;
    if (st != HSS_OK) return st;
    return hip_status(hipGetLastError());
}

int launch_de_cmp(const DeArgs &a, int n_mid, int variant, int storage, hipStream_t stream);
int launch_de_zocrs(const DeArgs &a, int n_mid, int variant, int storage, hipStream_t stream);
int launch_de_oct(const DeArgs &a, int n_mid, int variant, int storage, hipStream_t stream);

}  // namespace hss
