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
// Each thread owns one (trace, midpoint).  For OCT a trace whose midpoint distance
// exceeds apm + sqrt(h^2 + h0^2) cannot pass for any candidate: the displaced midpoint
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

__global__ __launch_bounds__(256) void select_traces_kernel(SelectArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = blockIdx.y;
    if (i >= a.ntr) return;
    const float m0 = a.m0[g];
    const float m = a.mid[i];
    unsigned char used = 0;
    if (a.model == HSS_CMP) {
        used = (m0 == m);
    } else if (a.model == HSS_ZOCRS) {
        used = fabsf(m0 - m) <= a.apm;
    } else {
        const float h = a.hoff[i];
        const float reach = sqrtf(h * h + a.h0 * a.h0);
        const float margin = 1e-4f * (fabsf(m0) + fabsf(m) + reach + a.apm) + 1e-3f;
        if (fabsf(m - m0) <= a.apm + reach + margin) {
            const int npairs = a.oct.mode == HSS_OCT_PER_SAMPLE ? 1
                             : (a.oct.mode == HSS_OCT_DE_START ? a.oct.NP : a.oct.npairs);
            for (int s = 0; !used && s < a.ns; ++s) {
                const float t0 = (float)s * a.dt;
                for (int p = 0; !used && p < npairs; ++p) {
                    float c, slope;
                    if (a.oct.mode == HSS_OCT_PAIRS) {
                        c = a.oct.d_pairs[2 * p];
                        slope = a.oct.d_pairs[2 * p + 1];
                    } else if (a.oct.mode == HSS_OCT_PER_SAMPLE) {
                        const float *vp = a.oct.d_vparams + (size_t)g * 2 * a.ns;
                        const float v = vp[s];
                        c = 4.0f / (v * v);
                        slope = vp[a.ns + s];
                    } else {
                        // the start population of differential evolution (de_impl.hpp)
                        uint32_t r[4];
                        philox4x32_10((uint32_t)s, a.oct.stream_base + (uint32_t)g, (uint32_t)p, 0u,
                                      a.key0, a.key1, r);
                        const float v = a.oct.lower[0] + u01(r[0]) * (a.oct.upper[0] - a.oct.lower[0]);
                        c = 4.0f / (v * v);
                        slope = a.oct.lower[1] + u01(r[1]) * (a.oct.upper[1] - a.oct.lower[1]);
                    }
                    float mh;
                    const int err = oct_displaced_midpoint(h, a.h0, t0, m0, c, slope, &mh);
                    used = (err == kNoError && fabsf(m - mh) <= a.apm);
                }
            }
        }
    }
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
    dim3 grid((traces->ntr + 255) / 256, n);
    hipLaunchKernelGGL(select_traces_kernel, grid, dim3(256), 0, (hipStream_t)stream, a);
    return hip_status(hipGetLastError());
}
