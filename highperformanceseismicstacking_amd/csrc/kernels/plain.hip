// The plain per-gather grid search: the reference's kernel structure on this GPU, kept as
// the baseline the batched kernels are measured against.
//
// Reference behaviour and structure (cmp/linear_search.cu:19-115, zocrs/ and oct/ twins):
// one launch per midpoint; one thread per (t0 sample s, candidate p), 64-thread blocks
// over the flat index s*P + p; the thread walks the midpoint's traces with a 32-entry
// window array indexed at run time, computes sqrt and t/dt with the library operations
// and writes S and stack to [ns][P] arrays; a second launch with one thread per sample
// scans the P candidates for the first strict maximum.  The arithmetic is the reference
// order (HSS_EXACT), so the results equal the batched kernel's HSS_EXACT bit for bit.
#include "launch_util.hpp"
#include "eval.hpp"

namespace hss {

struct PlainArgs {
    EvalCtx e;  // beg/cnt/m0 of the one midpoint
    const float *params;
    int P;
    float *sembl;  // [ns][P]
    float *stack;
};

template <int MODEL>
__global__ __launch_bounds__(64) void plain_semblance_kernel(PlainArgs a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)a.e.ns * a.P) return;
    const int s = (int)(i / a.P);
    const int p = (int)(i - (long long)s * a.P);
    const EvalCtx &e = a.e;
    Candidate k{0.0f, 0.0f, 0.0f};
    if constexpr (MODEL == HSS_CMP) {
        k.c = a.params[p];
    } else if constexpr (MODEL == HSS_ZOCRS) {
        k.c = a.params[3 * p];
        k.p1 = a.params[3 * p + 1];
        k.p2 = a.params[3 * p + 2];
    } else {
        k.c = a.params[2 * p];
        k.p1 = a.params[2 * p + 1];
    }
    const float t0 = (float)s * e.dt;
    const float *samples = (const float *)e.samples;
    float A[kMaxWindow];
    for (int j = 0; j < e.w; ++j) A[j] = 0.0f;
    float lin = 0.0f, den = 0.0f;
    unsigned used = 0;
    for (int j = 0; j < e.cnt; ++j) {
        const int ti = e.midx ? e.midx[e.beg + j] : e.beg + j;
        TraceGeom tg{0.0f, 0.0f, 0.0f};
        if constexpr (MODEL == HSS_CMP) {
            tg.h2 = e.hoff2[ti];
        } else if constexpr (MODEL == HSS_ZOCRS) {
            tg.m = e.mid[ti];
            tg.h2 = e.hoff2[ti];
        } else {
            tg.m = e.mid[ti];
            tg.h = e.hoff[ti];
        }
        float t;
        bool r;
        if (!traveltime<MODEL, false>(tg, t0, t0 * t0, e.m0, e.h0, e.apm, k, &t, &r)) continue;
        int first;
        float f;
        if (!window_index<false>(t, e.dt, e.inv_dt, e.ib, e.tau, &first, &f)) continue;
        const float *y = samples + (size_t)ti * e.ns + first;
        for (int q = 0; q < e.w; ++q) {
            const float u = (y[q + 1] - y[q]) * f + y[q];
            A[q] += u;
            lin += u;
            den += u * u;
        }
        used++;
    }
    float sem = 0.0f, stk = 0.0f;
    if (used > 0) {
        float num = 0.0f;
        for (int q = 0; q < e.w; ++q) num += A[q] * A[q];
        sem = num / ((float)used * den);
        stk = lin / (float)(used * (unsigned)e.w);
    }
    a.sembl[i] = sem;
    a.stack[i] = stk;
}

template <int MODEL>
__global__ __launch_bounds__(64) void plain_select_kernel(const float *sembl, const float *stack,
                                                          const float *params, int P, int ns,
                                                          float *result) {
    constexpr int NP = MODEL == HSS_CMP ? 1 : (MODEL == HSS_ZOCRS ? 3 : 2);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    float best = -1.0f, bstack = 0.0f;
    int bi = -1;
    for (int p = 0; p < P; ++p) {
        const float v = sembl[(size_t)s * P + p];
        if (v > best) {
            best = v;
            bstack = stack[(size_t)s * P + p];
            bi = p;
        }
    }
    result[s] = best;
    result[ns + s] = bi >= 0 ? bstack : 0.0f;
    for (int d = 0; d < NP; ++d) {
        float v = 0.0f;
        if (bi >= 0) {
            v = params[(size_t)bi * NP + d];
            if (d == 0) v = 2 / sqrtf(v);
        }
        result[(2 + d) * ns + s] = v;
    }
}

}  // namespace hss

using namespace hss;

extern "C" int hss_grid_semblance_plain(const hss_traces *traces, int model, const float *d_params,
                                        int32_t P, int32_t n, const int32_t *h_begin,
                                        const int32_t *h_count, const float *h_m0, float h0,
                                        float apm, float *d_sembl, float *d_stack,
                                        float *d_result, void *stream) {
    int rc = check_traces(traces);
    if (rc != HSS_OK) return rc;
    if (model_param_count(model) < 0) return HSS_ERR_INVALID_MODEL;
    if (traces->storage != HSS_F32) return HSS_ERR_INVALID_STORAGE;
    if (!d_params || !d_sembl || !d_stack || !d_result) return HSS_ERR_NULL_POINTER;
    if (P <= 0 || n < 0) return HSS_ERR_INVALID_SIZE;
    if (n > 0 && (!h_begin || !h_count || !h_m0)) return HSS_ERR_NULL_POINTER;
    const unsigned long long threads = (unsigned long long)traces->ns * (unsigned long long)P;
    if (threads / 64 + 1 > 0x7fffffffULL) return HSS_ERR_INVALID_SIZE;
    for (int g = 0; g < n; ++g) {
        if (h_begin[g] < 0 || h_count[g] < 0 || h_begin[g] + h_count[g] > traces->ntr)
            return HSS_ERR_INVALID_SIZE;
    }
    const int np = model_param_count(model);
    hipStream_t st = (hipStream_t)stream;
    PlainArgs a{};
    a.e.samples = traces->samples;
    a.e.mid = traces->mid;
    a.e.hoff = traces->hoff;
    a.e.hoff2 = traces->hoff2;
    a.e.ns = traces->ns;
    a.e.tau = traces->tau;
    a.e.w = traces->w;
    a.e.dt = traces->dt;
    a.e.inv_dt = 1.0f / traces->dt;
    a.e.ib = index_bounds(traces->ns, traces->tau);
    a.e.h0 = h0;
    a.e.apm = apm;
    a.params = d_params;
    a.P = P;
    a.sembl = d_sembl;
    a.stack = d_stack;
    const unsigned blocks = (unsigned)((threads + 63) / 64);
    const unsigned sblocks = (unsigned)((traces->ns + 63) / 64);
    for (int g = 0; g < n; ++g) {
        a.e.beg = h_begin[g];
        a.e.cnt = h_count[g];
        a.e.m0 = h_m0[g];
        float *res = d_result + (size_t)g * (size_t)(2 + np) * (size_t)traces->ns;
        dispatch_model(model, [&](auto M) {
            constexpr int MODEL = decltype(M)::value;
            hipLaunchKernelGGL(plain_semblance_kernel<MODEL>, dim3(blocks), dim3(64), 0, st, a);
            hipLaunchKernelGGL(plain_select_kernel<MODEL>, dim3(sblocks), dim3(64), 0, st, d_sembl,
                               d_stack, d_params, P, traces->ns, res);
        });
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
    }
    return HSS_OK;
}
