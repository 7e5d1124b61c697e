// Stretch-free re-stack with the best-shift pick fused in.
//
// Reference behaviour (cmp/stretch_free.cu:4-66, zocrs/stretch_free.cu:4-85,
// oct/stretch_free.cu:5-97, then common/stretch_free.cu:3-35): for every t0 sample s and
// every shift n of nArray (-10..10, StretchFreeAlgorithm.cpp:77-82) take the parameters
// picked earlier at sample s_n = s - n (velocity as v, c_n = 4/v^2), evaluate the
// traveltime at t0_n = s_n*dt and shift it by n*dt; a shift whose s_n falls outside the
// trace gives S = stack = 0.  The first strict maximum over the shifts wins and the
// result is (S, stack, n).
//
// MI355X mapping: as the grid kernel -- one workgroup per (midpoint, 64 t0), lane = t0,
// wave v takes shifts v, v+4, ...; here the candidate differs per lane (it comes from
// sample s_n), and the four per-wave bests merge through LDS.
#pragma once

#include "eval.hpp"
#include "eval_compact.hpp"
#include "grid_impl.hpp"
#include "launch_util.hpp"

namespace hss {

constexpr int kSfMaxShifts = 64;
// Each lane's candidate comes from another sample (s - n), so a trace's windows spread
// over more than one wave segment; 16-bit windows are read per lane (Window::add16),
// which measured twice as fast as segments on a 2000-midpoint CMP line.
constexpr bool kSfSegments = false;

struct SfArgs {
    EvalCtx e;
    const int *mbeg;
    const int *mcnt;
    const float *mm0;
    const int *morder;  // launch order (launch_midpoint) or NULL
    int nmid;
    const float *vparams;  // [n][np][ns]: velocity, then A,B or slope, per sample
    float nvals[kSfMaxShifts];
    int nN;
    int ntiles;
    float *result;  // [n][3][ns]
    float *sembl;   // [n][ns][nN] or NULL
    float *stack;
    unsigned long long *not_used;
};

// The budget of OCT's kernel, which holds both trace loops (with and without the kept
// offset terms): six waves per SIMD for windows of up to 11 samples, where both loops fit
// without spills (the C4 line's w = 11: the loop in the reference's order would otherwise
// take 81 registers, five waves); wider and generic windows at the compiler's own count.
template <int MODEL, int W, int VARIANT>
constexpr int sf_min_waves() {
    return MODEL == HSS_OCT && VARIANT != HSS_KAHAN && W > 0 && W <= 11 ? 6 : 1;
}

template <int MODEL, int W, int VARIANT, int STORAGE>
__global__ __launch_bounds__(kGridThreads)
__attribute__((amdgpu_waves_per_eu(sf_min_waves<MODEL, W, VARIANT>(), 8))) void sf_semblance_kernel(SfArgs a) {
    constexpr int NP = MODEL == HSS_CMP ? 1 : (MODEL == HSS_ZOCRS ? 3 : 2);
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int blk = logical_block(blockIdx.x, gridDim.x);
    const int slot = blk / a.ntiles;
    const int tile = blk - slot * a.ntiles;
    const int g = launch_midpoint(a.morder, a.nmid, slot);
    if (g < 0) return;
    const int s = tile * kWave + lane;
    const int ns = a.e.ns;
    const bool active = s < ns;
    const EvalCtx e = midpoint_ctx(a.e, a.mbeg, a.mcnt, a.mm0, g);
    const float *vp = a.vparams + (size_t)g * NP * ns;

    __shared__ float seg_lds[kGridWaves][kSegFloats];
    __shared__ OctTraceRec trace_recs[kGridWaves][kWave];
    float best = -1.0f, bstack = 0.0f;
    int bidx = -1;
    unsigned long long rejected = 0;
    // OCT: keep the offset-dependent traveltime terms across runs of traces with one half
    // offset when the midpoint's list is grouped by offset (wave-uniform); each lane's
    // candidate comes from its own sample, and the kept terms are per lane all the same
    bool keep = false;
    if constexpr (MODEL == HSS_OCT) keep = oct_offset_runs_long(e);
    // a BF16 line stored without scales sums its windows as FP32 windows; both choices are
    // made outside the trace loop, one code path each
    const bool unscaled16 = STORAGE == HSS_BF16 && e.scale == nullptr;
    for (int q = wave; q < a.nN; q += kGridWaves) {
        const int n = (int)a.nvals[q];
        const int sn = s - n;
        const bool valid = active && sn >= 0 && sn < ns;
        const int sc = valid ? sn : 0;
        Candidate k{0.0f, 0.0f, 0.0f};
        const float v = vp[sc];
        k.c = 4.0f / (v * v);
        if constexpr (MODEL == HSS_ZOCRS) {
            k.p1 = vp[ns + sc];
            k.p2 = vp[2 * ns + sc];
        } else if constexpr (MODEL == HSS_OCT) {
            k.p1 = vp[ns + sc];
        }
        const float t0n = (float)sc * e.dt;
        float S = 0.0f, stk = 0.0f;
        unsigned rej = 0;
        auto eval = [&](auto keep_c, auto factored_c) {
            evaluate<MODEL, W, VARIANT, STORAGE, true, kSfSegments, decltype(keep_c)::value,
                     decltype(factored_c)::value>(e, t0n, t0n * t0n, k, (float)n * e.dt, valid,
                                                  seg_lds[wave], &S, &stk, &rej);
        };
        using T = std::true_type;
        using F = std::false_type;
        bool walked = false;
        if constexpr (MODEL == HSS_OCT && VARIANT != HSS_KAHAN) {
            if (keep) {
                // each lane walks only the traces its own aperture keeps (the compacted
                // walk of differential evolution; the same sums in the same order)
                const Candidate kc[1] = {k};
                const bool on[1] = {true};
                float s1[1], t1[1];
                evaluate_oct_compact<W, VARIANT, STORAGE, true, 1>(
                    e, t0n, kc, on, (float)n * e.dt, valid, trace_recs[wave], s1, t1, &rej);
                S = s1[0];
                stk = t1[0];
                walked = true;
            }
        }
        if (walked) {
        } else if (MODEL == HSS_OCT && keep) {
            if (unscaled16)
                eval(T{}, F{});
            else
                eval(T{}, T{});
        } else if (unscaled16) {
            eval(F{}, F{});
        } else {
            eval(F{}, T{});
        }
        if (!valid) {
            S = 0.0f;
            stk = 0.0f;
        }
        if (active) {
            if (valid) rejected += rej;
            if (a.sembl) a.sembl[((size_t)g * ns + s) * a.nN + q] = S;
            if (a.stack) a.stack[((size_t)g * ns + s) * a.nN + q] = stk;
            if (S > best) {
                best = S;
                bstack = stk;
                bidx = q;
            }
        }
    }
    if (MODEL == HSS_OCT && a.not_used) {
        unsigned long long tot = wave_sum_u64(rejected);
        if (lane == 0 && tot) atomicAdd(a.not_used + g, tot);
    }
    float B, BS;
    int BI;
    merge_bests<kGridWaves>(best, bstack, bidx, wave, lane, 1, &B, &BS, &BI);
    if (wave != 0 || !active) return;
    float *res = a.result + (size_t)g * 3 * ns;
    res[s] = B;
    res[ns + s] = BI >= 0 ? BS : 0.0f;
    res[2 * ns + s] = BI >= 0 ? a.nvals[BI] : 0.0f;
}

template <int MODEL>
int launch_sf(const SfArgs &a, int n_mid, int variant, int storage, hipStream_t stream) {
    int st = HSS_OK;
    const unsigned blocks = (unsigned)n_mid * (unsigned)a.ntiles;
    if (blocks == 0) return HSS_OK;
    int rc = dispatch_storage(storage, [&](auto S) {
        st = dispatch_variant(variant, [&](auto V) {
            dispatch_window(a.e.w, fast_index_verified(a.e.dt, a.e.tau, stream), [&](auto Wc) {
                hipLaunchKernelGGL((sf_semblance_kernel<MODEL, decltype(Wc)::value,
                                                        decltype(V)::value, decltype(S)::value>),
                                   dim3(blocks), dim3(kGridThreads), 0, stream, a);
            });
        });
    });
    if (rc != HSS_OK) return rc;
    if (st != HSS_OK) return st;
    return hip_status(hipGetLastError());
}

int launch_sf_cmp(const SfArgs &a, int n_mid, int variant, int storage, hipStream_t stream);
int launch_sf_zocrs(const SfArgs &a, int n_mid, int variant, int storage, hipStream_t stream);
int launch_sf_oct(const SfArgs &a, int n_mid, int variant, int storage, hipStream_t stream);

}  // namespace hss
