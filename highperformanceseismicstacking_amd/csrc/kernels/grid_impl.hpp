// Grid ("linear") search semblance with the best-candidate pick fused in.
//
// Reference behaviour: computeSemblancesFor<Model> (cmp/linear_search.cu:19-81,
// zocrs/linear_search.cu:36-109, oct/linear_search.cu:30-114) evaluates one thread per
// (t0, candidate), and selectBestSemblancesFor<Model> (cmp/linear_search.cu:83-115 and
// twins) scans the candidates of each t0 for the first strict maximum.
//
// MI355X mapping: one workgroup = (midpoint, tile of 64 consecutive t0 samples), four
// wavefronts.  Lane l of every wave owns t0 sample tile*64 + l; wave v evaluates
// candidates v, v+4, v+8, ... in increasing order.  The candidate is wave-uniform (its
// parameters sit in SGPRs), each lane walks the midpoint's traces in order and reads a
// (w+1)-sample window that is nearly contiguous across the 64 lanes, so the loads
// coalesce.  Each lane keeps its running best (strict '>', so the lowest index wins
// among equal values and NaN never wins); the four per-wave bests are merged through LDS
// (larger value, then lower index), which reproduces the reference's sequential scan.
#pragma once

#include "eval.hpp"
#include "launch_util.hpp"

namespace hss {

constexpr int kGridWaves = 4;
constexpr int kGridThreads = kGridWaves * kWave;

// 16-bit storage reads windows through the wave's LDS segments; per-lane reads from the
// trace row (Window::add16, what DE and stretch-free use) measured 45% slower for the grid
// on the C5 line (FP16, 2048 velocities), FAST and Kahan alike.
template <int STORAGE>
constexpr bool grid_segments() { return STORAGE != HSS_F32; }

struct GridArgs {
    EvalCtx e;  // per-midpoint fields are filled in by the kernel
    const int *mbeg;
    const int *mcnt;
    const float *mm0;
    const int *morder;  // launch order (launch_midpoint) or NULL
    int nmid;
    const float *params;
    int P;
    int ntiles;      // 64-sample t0 tiles of a midpoint
    int tt;          // t0 tiles per workgroup (1, 2 or 4; the waves split into tt x 4/tt)
    int wg_per_mid;  // workgroups per midpoint, ceil(ntiles / tt)
    float *result;
    float *sembl;
    float *stack;
    unsigned long long *not_used;
};

template <int MODEL>
__device__ __forceinline__ Candidate load_candidate(const float *__restrict__ params, int q) {
    Candidate k{0.0f, 0.0f, 0.0f};
    if constexpr (MODEL == HSS_CMP) {
        k.c = params[q];
    } else if constexpr (MODEL == HSS_ZOCRS) {
        k.c = params[3 * q];
        k.p1 = params[3 * q + 1];
        k.p2 = params[3 * q + 2];
    } else {
        k.c = params[2 * q];
        k.p1 = params[2 * q + 1];
    }
    return k;
}

// Merges the running bests of the waves that share a t0 tile (waves tloc, tloc + tt, ...;
// larger value first, then lower index); returns the winner in the lanes of wave tloc.
template <int NWAVES>
__device__ __forceinline__ void merge_bests(float best, float bstack, int bidx, int wave,
                                            int lane, int tt, float *B, float *BS, int *BI) {
    __shared__ float sh_best[NWAVES][kWave];
    __shared__ float sh_stack[NWAVES][kWave];
    __shared__ int sh_idx[NWAVES][kWave];
    sh_best[wave][lane] = best;
    sh_stack[wave][lane] = bstack;
    sh_idx[wave][lane] = bidx;
    __syncthreads();
    *B = -1.0f;
    *BS = 0.0f;
    *BI = -1;
    const int tloc = wave % tt;
    for (int v = tloc; v < NWAVES; v += tt) {
        const int iv = sh_idx[v][lane];
        const float sv = sh_best[v][lane];
        if (iv >= 0 && (sv > *B || (sv == *B && iv < *BI))) {
            *B = sv;
            *BS = sh_stack[v][lane];
            *BI = iv;
        }
    }
}

// CPL: candidates per lane (the parameter tile): wave group v of the 4/tt groups takes the
// candidates [q0, q0 + CPL) with q0 = CPL*v, CPL*(v + 4/tt), ...; within a lane candidates
// are taken in increasing order, so the running best is still the first strict maximum.
// Register budget of the compensated variant, as waves per SIMD the compiler must fit:
// its running totals and their compensation double the window state, which at one
// candidate per lane would leave 5 waves; 7 fit without spilling and measured 12% faster
// on the C5 line (FP16).  (Two candidates per lane do not fit well at any budget; the
// autotuner picks one.)
template <int VARIANT, int CPL>
constexpr int grid_min_waves() { return VARIANT == HSS_KAHAN && CPL == 1 ? 7 : 1; }
// The grid kernel's budget: OCT, which holds both trace loops (with and without the kept
// offset terms), at five waves per SIMD -- the loop without them fits as before (93
// registers), the one with them spills a few run-level values; at the compiler's own
// choice (107 registers, four waves) the reference-order line ran 6% slower.
template <int MODEL, int VARIANT, int CPL>
constexpr int grid_waves_budget() {
    return MODEL == HSS_OCT && VARIANT != HSS_KAHAN ? 5 : grid_min_waves<VARIANT, CPL>();
}

template <int MODEL, int W, int VARIANT, int STORAGE, int CPL>
__global__ __launch_bounds__(kGridThreads)
__attribute__((amdgpu_waves_per_eu(grid_waves_budget<MODEL, VARIANT, CPL>(), 8))) void grid_semblance_kernel(
    GridArgs a) {
    constexpr int NP = MODEL == HSS_CMP ? 1 : (MODEL == HSS_ZOCRS ? 3 : 2);
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    // the workgroup covers tt t0 tiles; its waves split into tt tiles x 4/tt candidate
    // groups (wave = group * tt + tile)
    const int tt = a.tt;
    const int ngroups = kGridWaves / tt;
    const int grp = wave / tt;
    const int blk = logical_block(blockIdx.x, gridDim.x);
    const int slot = blk / a.wg_per_mid;
    const int tile = (blk - slot * a.wg_per_mid) * tt + wave % tt;
    const int g = launch_midpoint(a.morder, a.nmid, slot);
    if (g < 0) return;
    const int s = tile * kWave + lane;
    const int ns = a.e.ns;
    const bool active = s < ns;  // (a tile past the last one has no active lane)
    const float t0 = (float)s * a.e.dt;
    const float t0sq = t0 * t0;
    const EvalCtx e = midpoint_ctx(a.e, a.mbeg, a.mcnt, a.mm0, g);

    __shared__ float seg_lds[kGridWaves][kSegFloats];
    float best = -1.0f, bstack = 0.0f;
    int bidx = -1;
    unsigned long long rejected = 0;
    // OCT: keep the offset-dependent traveltime terms across runs of traces with one half
    // offset when the midpoint's list is grouped by offset (the C4 line's grid search: 1849
    // -> 850 ms).  Counted once per workgroup over the list (wave-uniform).
    bool keep = false;
    if constexpr (MODEL == HSS_OCT) keep = oct_offset_runs_long(e);

    for (int q0 = grp * CPL; q0 < a.P; q0 += ngroups * CPL) {
        Candidate k[CPL];
        bool on[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            on[c] = q0 + c < a.P;
            k[c] = load_candidate<MODEL>(a.params, on[c] ? q0 + c : q0);
        }
        float S[CPL], stk[CPL];
        unsigned rej_q = 0;
        if (MODEL == HSS_OCT && keep)
            evaluate_n<MODEL, W, VARIANT, STORAGE, false, CPL, grid_segments<STORAGE>(), true>(
                e, t0, t0sq, k, on, 0.0f, active, seg_lds[wave], S, stk, &rej_q);
        else
            evaluate_n<MODEL, W, VARIANT, STORAGE, false, CPL, grid_segments<STORAGE>()>(
                e, t0, t0sq, k, on, 0.0f, active, seg_lds[wave], S, stk, &rej_q);
        if (active) {
            rejected += rej_q;
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                if (!on[c]) continue;
                const int q = q0 + c;
                if (a.sembl) a.sembl[((size_t)g * ns + s) * a.P + q] = S[c];
                if (a.stack) a.stack[((size_t)g * ns + s) * a.P + q] = stk[c];
                if (S[c] > best) {
                    best = S[c];
                    bstack = stk[c];
                    bidx = q;
                }
            }
        }
    }

    if (MODEL == HSS_OCT && a.not_used) {
        unsigned long long tot = wave_sum_u64(rejected);
        if (lane == 0 && tot) atomicAdd(a.not_used + g, tot);
    }

    float B, BS;
    int BI;
    merge_bests<kGridWaves>(best, bstack, bidx, wave, lane, tt, &B, &BS, &BI);
    if (grp != 0 || !active) return;
    float *res = a.result + (size_t)g * (2 + NP) * ns;
    res[s] = B;
    res[ns + s] = BI >= 0 ? BS : 0.0f;
    if (BI >= 0) {
        const Candidate k = load_candidate<MODEL>(a.params, BI);
        res[2 * ns + s] = 2 / sqrtf(k.c);
        if constexpr (NP >= 2) res[3 * ns + s] = k.p1;
        if constexpr (NP >= 3) res[4 * ns + s] = k.p2;
    } else {
#pragma unroll
        for (int d = 0; d < NP; ++d) res[(2 + d) * ns + s] = 0.0f;
    }
}

// Host launcher for one model: picks the window/variant/storage instantiation.  n_slots
// launch slots (midpoints taken up in launch order) of a.wg_per_mid workgroups each.
template <int MODEL>
int launch_grid(const GridArgs &a, int n_slots, int variant, int storage, int cpl,
                hipStream_t stream) {
    int st = HSS_OK;
    const unsigned blocks = (unsigned)n_slots * (unsigned)a.wg_per_mid;
    if (blocks == 0) return HSS_OK;
    int rc = dispatch_storage(storage, [&](auto S) {
        st = dispatch_variant(variant, [&](auto V) {
            dispatch_window(a.e.w, fast_index_verified(a.e.dt, a.e.tau, stream), [&](auto Wc) {
                constexpr int WW = decltype(Wc)::value, VV = decltype(V)::value,
                              SS = decltype(S)::value;
                if (cpl == 2 && WW > 0)
                    hipLaunchKernelGGL((grid_semblance_kernel<MODEL, WW, VV, SS, 2>), dim3(blocks),
                                       dim3(kGridThreads), 0, stream, a);
                else
                    hipLaunchKernelGGL((grid_semblance_kernel<MODEL, WW, VV, SS, 1>), dim3(blocks),
                                       dim3(kGridThreads), 0, stream, a);
            });
        });
    });
    if (rc != HSS_OK) return rc;
    if (st != HSS_OK) return st;
    return hip_status(hipGetLastError());
}

int launch_grid_cmp(const GridArgs &a, int n_slots, int variant, int storage, int cpl,
                    hipStream_t stream);
int launch_grid_zocrs(const GridArgs &a, int n_slots, int variant, int storage, int cpl,
                    hipStream_t stream);
int launch_grid_oct(const GridArgs &a, int n_slots, int variant, int storage, int cpl,
                    hipStream_t stream);

}  // namespace hss
