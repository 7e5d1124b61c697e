// Trace selection masks for a batch of midpoints over a whole line.
//
// Reference behaviour:
//   CMP    mask = (m0 == m)                         cmp/common.cu:3-15
//   ZO-CRS mask = |m0 - m| <= apm                   zocrs/common.cu:3-16
//   OCT    mask = some (t0 sample, candidate) has no displaced-midpoint error and
//          |m - mh| <= apm, scanning t0 then candidates with an early exit.  The
//          candidates are
//            grid:          the same (c, slope) pairs for every t0 (1024 random draws,
//                           oct/linear_search.cu:155-192);
//            stretch-free:  the (v, slope) picked at that very t0 (oct/stretch_free.cu:99-138);
//            DE:            the start population of that t0 (oct/differential_evolution.cu:130-168).
//
// Each thread owns one (trace, midpoint); a workgroup's threads share the midpoint, which
// the OCT path uses to take each candidate once per workgroup (select_oct).  For OCT a
// trace whose midpoint distance exceeds apm + sqrt(h^2 + h0^2) cannot pass for any
// candidate: the displaced midpoint
// satisfies |mh - m0| = 2|theta||h^2-h0^2| / sqrt(theta^2 gamma^2 + 2q + 2sqrt(...))
// <= |h^2-h0^2| / sqrt(h^2+h0^2) <= sqrt(h^2+h0^2), because the denominator is at least
// |theta| gamma = 2|theta| sqrt(h^2+h0^2).  The test keeps a wide rounding margin, so it
// only skips traces whose answer is certain.
#include "de_impl.hpp"

namespace hss {

struct SelectArgs {
    const float *mid;
    const float *hoff;
    int ntr;
    int ns;
    float dt;
    const float *m0;
    float apm, h0;
    int model;
    hss_oct_candidates oct;
    uint32_t key0, key1;
    unsigned char *mask;
};

constexpr int kSelectThreads = 256;

// The first test of an OCT trace: a midpoint distance beyond apm + sqrt(h^2 + h0^2)
// (plus the rounding margin) cannot pass for any candidate.
__device__ __forceinline__ bool oct_within_reach(float m, float h, float m0, float h0, float apm) {
    const float reach = sqrtf(h * h + h0 * h0);
    const float margin = 1e-4f * (fabsf(m0) + fabsf(m) + reach + apm) + 1e-3f;
    return fabsf(m - m0) <= apm + reach + margin;
}

// Distinct half offsets whose displaced midpoints a workgroup shares (16 KB of LDS).
constexpr int kSelectOffsets = 16;

// Candidates per t0 sample of one midpoint, in each mode.
__device__ __forceinline__ int oct_candidates_per_sample(const SelectArgs &a) {
    return a.oct.mode == HSS_OCT_PER_SAMPLE ? 1
         : (a.oct.mode == HSS_OCT_DE_START ? a.oct.NP : a.oct.npairs);
}

// Candidate p of t0 sample s of midpoint g as (c, slope).
__device__ __forceinline__ void oct_candidate(const SelectArgs &a, int g, int s, int p, float *c,
                                              float *slope) {
    if (a.oct.mode == HSS_OCT_PAIRS) {
        *c = a.oct.d_pairs[2 * p];
        *slope = a.oct.d_pairs[2 * p + 1];
    } else if (a.oct.mode == HSS_OCT_PER_SAMPLE) {
        const float *vp = a.oct.d_vparams + (size_t)g * 2 * a.ns;
        const float v = vp[s];
        *c = 4.0f / (v * v);
        *slope = vp[a.ns + s];
    } else {
        // the start population of differential evolution (de_impl.hpp)
        uint32_t r[4];
        philox4x32_10((uint32_t)s, a.oct.stream_base + (uint32_t)g, (uint32_t)p, 0u, a.key0,
                      a.key1, r);
        const float v = a.oct.lower[0] + u01(r[0]) * (a.oct.upper[0] - a.oct.lower[0]);
        *c = 4.0f / (v * v);
        *slope = a.oct.lower[1] + u01(r[1]) * (a.oct.upper[1] - a.oct.lower[1]);
    }
}

// OCT: the candidates of a midpoint are the same for every trace, and a workgroup serves
// one midpoint, so it takes each (t0 sample, candidate) pair once -- kSelectThreads at a
// time (drawn, for DE start populations) -- and every thread whose trace is still open
// tests them.  The displaced midpoint depends on the trace's half offset, not on its
// midpoint, so when the workgroup's traces have at most kSelectOffsets distinct offsets
// (their bit patterns) it is computed once per (offset, candidate) into a table, an
// error stored as NaN (which fails the aperture test as the error does), and each thread
// only compares its midpoint with its offset's row.  The mask is the OR over all
// candidates, so neither the order nor the early exit changes it; the workgroup stops
// when no trace is open.  Every thread takes part in the barriers (those past the last
// trace have nothing open).
__device__ void select_oct(const SelectArgs &a, int i, int g) {
    __shared__ float cand_c[kSelectThreads], cand_s[kSelectThreads];
    __shared__ float mh_tab[kSelectOffsets][kSelectThreads];
    __shared__ uint32_t hbits[kSelectThreads];
    __shared__ float off_h[kSelectOffsets];
    __shared__ int slot_of[kSelectThreads];
    __shared__ int row_open[kSelectOffsets];
    __shared__ int n_off;
    const int tid = (int)threadIdx.x;
    const bool live = i < a.ntr;
    const float m0 = a.m0[g];
    const float m = live ? a.mid[i] : 0.0f;
    const float h = live ? a.hoff[i] : 0.0f;
    bool open = live && oct_within_reach(m, h, m0, a.h0, a.apm);
    unsigned char used = 0;

    // the workgroup's distinct offsets: the lowest live thread of each bit pattern takes a
    // table row (in any order: the row numbers change nothing)
    if (tid == 0) n_off = 0;
    hbits[tid] = __float_as_uint(h);
    slot_of[tid] = -1;
    __syncthreads();
    const int nlive = min(kSelectThreads, a.ntr - (int)(blockIdx.x * blockDim.x));
    bool first = live;
    for (int t = 0; first && t < tid; ++t) first = hbits[t] != hbits[tid];
    if (first) {
        const int r = atomicAdd(&n_off, 1);
        slot_of[tid] = r;
        if (r < kSelectOffsets) off_h[r] = h;
    }
    __syncthreads();
    const bool tabled = n_off <= kSelectOffsets;  // (uniform)
    int my_row = -1;
    if (tabled && live) {
        for (int t = 0; t < nlive; ++t)
            if (hbits[t] == hbits[tid]) {
                my_row = slot_of[t];
                break;
            }
    }

    const int NP = oct_candidates_per_sample(a);  // candidates per t0 sample
    const long long total = (long long)a.ns * NP;
    for (long long q0 = 0; q0 < total; q0 += kSelectThreads) {
        if (!__syncthreads_or(open)) break;  // (also the barrier before LDS is rewritten)
        if (tabled) {
            // only the offsets some open trace still has get their row computed
            if (tid < kSelectOffsets) row_open[tid] = 0;
            __syncthreads();
            if (open) row_open[my_row] = 1;
            __syncthreads();
        }
        const long long q = q0 + tid;
        if (q < total) {
            const int s = (int)(q / NP), p = (int)(q - (long long)s * NP);
            float c, slope;
            oct_candidate(a, g, s, p, &c, &slope);
            if (tabled) {
                const float t0 = (float)s * a.dt;
                for (int row = 0; row < n_off; ++row) {
                    if (!row_open[row]) continue;
                    float mh;
                    const int err = oct_displaced_midpoint(off_h[row], a.h0, t0, m0, c, slope, &mh);
                    mh_tab[row][tid] = err == kNoError ? mh : __int_as_float(0x7fc00000);
                }
            } else {
                cand_c[tid] = c;
                cand_s[tid] = slope;
            }
        }
        __syncthreads();
        if (open) {
            const int n = (int)min((long long)kSelectThreads, total - q0);
            if (tabled) {
                for (int j = 0; j < n; ++j)
                    if (fabsf(m - mh_tab[my_row][j]) <= a.apm) {
                        used = 1;
                        open = false;
                        break;
                    }
            } else {
                int s = (int)(q0 / NP), p = (int)(q0 - (long long)s * NP);
                for (int j = 0; j < n; ++j) {
                    float mh;
                    const int err = oct_displaced_midpoint(h, a.h0, (float)s * a.dt, m0, cand_c[j],
                                                           cand_s[j], &mh);
                    if (err == kNoError && fabsf(m - mh) <= a.apm) {
                        used = 1;
                        open = false;
                        break;
                    }
                    if (++p == NP) {
                        p = 0;
                        ++s;
                    }
                }
            }
        }
    }
    if (live) a.mask[(size_t)g * a.ntr + i] = used;
}

__global__ __launch_bounds__(kSelectThreads) void select_traces_kernel(SelectArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = blockIdx.y;
    if (a.model == HSS_OCT) {  // (uniform over the launch)
        select_oct(a, i, g);
        return;
    }
    if (i >= a.ntr) return;
    const float m0 = a.m0[g];
    const float m = a.mid[i];
    unsigned char used = 0;
    if (a.model == HSS_CMP)
        used = (m0 == m);
    else  // HSS_ZOCRS
        used = fabsf(m0 - m) <= a.apm;
    a.mask[(size_t)g * a.ntr + i] = used;
}

}  // namespace hss

using namespace hss;

extern "C" int hss_select_traces(const hss_traces *traces, int model, int32_t n, const float *d_m0,
                                 float apm, float h0, const hss_oct_candidates *oct,
                                 uint8_t *d_mask, void *stream) {
    if (!traces || !traces->mid || !traces->hoff || !d_mask) return HSS_ERR_NULL_POINTER;
    if (model_param_count(model) < 0) return HSS_ERR_INVALID_MODEL;
    if (n < 0 || traces->ntr < 0 || traces->ns <= 0 || n > 65535) return HSS_ERR_INVALID_SIZE;
    if (n > 0 && !d_m0) return HSS_ERR_NULL_POINTER;
    SelectArgs a{};
    a.mid = traces->mid;
    a.hoff = traces->hoff;
    a.ntr = traces->ntr;
    a.ns = traces->ns;
    a.dt = traces->dt;
    a.m0 = d_m0;
    a.apm = apm;
    a.h0 = h0;
    a.model = model;
    a.mask = d_mask;
    if (model == HSS_OCT) {
        if (!oct) return HSS_ERR_NULL_POINTER;
        a.oct = *oct;
        if (oct->mode == HSS_OCT_PAIRS) {
            if (oct->npairs < 0) return HSS_ERR_INVALID_SIZE;
            if (oct->npairs > 0 && !oct->d_pairs) return HSS_ERR_NULL_POINTER;
        } else if (oct->mode == HSS_OCT_PER_SAMPLE) {
            if (!oct->d_vparams) return HSS_ERR_NULL_POINTER;
        } else if (oct->mode == HSS_OCT_DE_START) {
            if (oct->NP <= 0) return HSS_ERR_INVALID_SIZE;
            a.key0 = (uint32_t)oct->seed;
            a.key1 = (uint32_t)(oct->seed >> 32);
        } else {
            return HSS_ERR_INVALID_MODEL;
        }
    }
    if (n == 0 || traces->ntr == 0) return HSS_OK;
    dim3 grid((traces->ntr + kSelectThreads - 1) / kSelectThreads, n);
    hipLaunchKernelGGL(select_traces_kernel, grid, dim3(kSelectThreads), 0, (hipStream_t)stream, a);
    return hip_status(hipGetLastError());
}
