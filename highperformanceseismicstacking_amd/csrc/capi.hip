// extern "C" entry points of libhss.so (declared in include/hss.h), except
// hss_select_traces (kernels/select.hip).
#include <hip/hip_runtime.h>

#include "../../include/hss.h"
#include "kernels/adjoint.hpp"
#include "kernels/de_impl.hpp"
#include "kernels/grid_impl.hpp"
#include "kernels/grid_staged.hpp"
#include "kernels/launch_util.hpp"
#include "kernels/sf_impl.hpp"

namespace hss {

// buildParameterArrayFor* (cmp/linear_search.cu:4-17, zocrs/linear_search.cu:4-34,
// oct/linear_search.cu:4-28): v-major decode, storage c = 4/v^2.
struct ParamGridArgs {
    float lo[3];
    float step[3];
    unsigned cnt[3];
    unsigned total;
};

template <int MODEL>
__global__ void build_parameter_array_kernel(ParamGridArgs a, float *out) {
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.total) return;
    if constexpr (MODEL == HSS_CMP) {
        float v = a.lo[0] + (float)k * a.step[0];
        out[k] = 4.0f / (v * v);
    } else if constexpr (MODEL == HSS_ZOCRS) {
        const unsigned iv = (k / (a.cnt[1] * a.cnt[2])) % a.cnt[0];
        const unsigned ia = (k / a.cnt[2]) % a.cnt[1];
        const unsigned ib = k % a.cnt[2];
        float v = a.lo[0] + (float)iv * a.step[0];
        out[3 * k] = 4.0f / (v * v);
        out[3 * k + 1] = a.lo[1] + (float)ia * a.step[1];
        out[3 * k + 2] = a.lo[2] + (float)ib * a.step[2];
    } else {
        const unsigned iv = (k / a.cnt[1]) % a.cnt[0];
        const unsigned is = k % a.cnt[1];
        float v = a.lo[0] + (float)iv * a.step[0];
        out[2 * k] = 4.0f / (v * v);
        out[2 * k + 1] = a.lo[1] + (float)is * a.step[1];
    }
}

// The midpoint-independent part of an evaluation context.
static EvalCtx base_ctx(const hss_traces *t, const hss_midpoints *m) {
    EvalCtx e{};
    e.samples = t->samples;
    e.scale = t->scale;
    e.mid = t->mid;
    e.hoff = t->hoff;
    e.hoff2 = t->hoff2;
    e.midx = m->index;
    e.ntr = t->ntr;
    e.nsamp_odd = (int)(((size_t)t->ntr * (size_t)t->ns) & 1);
    e.left_cap = 2 + (kSeg + 2) / t->ns;
    e.ns = t->ns;
    e.tau = t->tau;
    e.w = t->w;
    e.dt = t->dt;
    e.inv_dt = 1.0f / t->dt;
    e.ib = index_bounds(t->ns, t->tau);
    e.h0 = m->h0;
    e.apm = m->apm;
    return e;
}

static int check_batch(const hss_traces *traces, const hss_midpoints *mids, int model,
                       int variant) {
    int rc = check_traces(traces);
    if (rc != HSS_OK) return rc;
    rc = check_midpoints(mids);
    if (rc != HSS_OK) return rc;
    if (model_param_count(model) < 0) return HSS_ERR_INVALID_MODEL;
    if (variant != HSS_EXACT && variant != HSS_FAST && variant != HSS_KAHAN)
        return HSS_ERR_INVALID_VARIANT;
    return HSS_OK;
}

static int tiles_of(int ns) { return (ns + kWave - 1) / kWave; }

}  // namespace hss

using namespace hss;

// Vector-Jacobian product of S with respect to the parameter points:
// out[g][m][s] = grad_S[g][s] * rows[g][2 + m][s] (rows = hss_semblance_gradient output).
__global__ __launch_bounds__(256) void vjp_kernel(int n, int np, int ns, const float *grad_S,
                                                  const float *rows, float *out) {
    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long total = (unsigned long long)n * np * ns;
    if (i >= total) return;
    const int s = (int)(i % (unsigned long long)ns);
    const unsigned long long gm = i / (unsigned long long)ns;
    const int m = (int)(gm % (unsigned long long)np);
    const int g = (int)(gm / (unsigned long long)np);
    out[i] = grad_S[(size_t)g * ns + s] * rows[((size_t)g * (2 + np) + 2 + m) * ns + s];
}

extern "C" {

int hss_abi_version(void) { return HSS_ABI_VERSION; }

const char *hss_status_string(int status) {
    switch (status) {
        case HSS_OK: return "success";
        case HSS_ERR_INVALID_MODEL: return "invalid traveltime model or mode";
        case HSS_ERR_INVALID_WINDOW: return "invalid window: w must be 2*tau+1 and at most 32";
        case HSS_ERR_INVALID_SIZE: return "invalid size";
        case HSS_ERR_NULL_POINTER: return "required pointer is NULL";
        case HSS_ERR_INVALID_STORAGE: return "invalid sample storage type";
        case HSS_ERR_INVALID_VARIANT: return "invalid arithmetic variant";
        case HSS_ERR_NO_DEVICE: return "no HIP device found";
        case HSS_ERR_SCRATCH_LIMIT:
            return "batch needs more DE workspace than the static scratch holds (HSS_DE_SCRATCH_BYTES)";
        default:
            if (status > 0) return hipGetErrorString((hipError_t)status);
            return "unknown status";
    }
}

int hss_device_count(int *count) {
    if (!count) return HSS_ERR_NULL_POINTER;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return e == hipErrorNoDevice ? HSS_ERR_NO_DEVICE : (int)e;
    }
    *count = n;
    return n > 0 ? HSS_OK : HSS_ERR_NO_DEVICE;
}

int hss_build_parameter_array(int model, const float *lower, const float *step,
                              const int32_t *count, float *d_params, void *stream) {
    const int np = model_param_count(model);
    if (np < 0) return HSS_ERR_INVALID_MODEL;
    if (!lower || !step || !count || !d_params) return HSS_ERR_NULL_POINTER;
    ParamGridArgs a{};
    unsigned long long total = 1;
    for (int d = 0; d < 3; ++d) {
        a.lo[d] = d < np ? lower[d] : 0.0f;
        a.step[d] = d < np ? step[d] : 0.0f;
        a.cnt[d] = d < np ? (unsigned)count[d] : 1u;
        if (d < np && count[d] <= 0) return HSS_ERR_INVALID_SIZE;
        total *= a.cnt[d];
    }
    if (total > 0x7fffffffULL) return HSS_ERR_INVALID_SIZE;
    a.total = (unsigned)total;
    const unsigned threads = 256, blocks = (a.total + threads - 1) / threads;
    hipStream_t st = (hipStream_t)stream;
    dispatch_model(model, [&](auto M) {
        hipLaunchKernelGGL(build_parameter_array_kernel<decltype(M)::value>, dim3(blocks),
                           dim3(threads), 0, st, a, d_params);
    });
    return hip_status(hipGetLastError());
}

int hss_grid_semblance(const hss_traces *traces, const hss_midpoints *mids, int model,
                       const float *d_params, int32_t P, int variant, float *d_result,
                       float *d_sembl, float *d_stack, unsigned long long *d_not_used,
                       void *stream) {
    int rc = check_batch(traces, mids, model, variant);
    if (rc != HSS_OK) return rc;
    if (!d_params || !d_result) return HSS_ERR_NULL_POINTER;
    if (P <= 0) return HSS_ERR_INVALID_SIZE;
    // the kernels address a midpoint's [ns][P] slab of the cubes with 32-bit offsets
    if ((d_sembl || d_stack) && (uint64_t)traces->ns * (uint64_t)P > 0xffffffffULL)
        return HSS_ERR_INVALID_SIZE;
    if (mids->n == 0) return HSS_OK;
    GridArgs a{};
    a.e = base_ctx(traces, mids);
    a.mbeg = mids->begin;
    a.mcnt = mids->count;
    a.mm0 = mids->m0;
    a.morder = mids->order;
    a.nmid = mids->n;
    a.params = d_params;
    a.P = P;
    a.ntiles = tiles_of(traces->ns);
    a.tt = 1;
    a.wg_per_mid = a.ntiles;
    a.result = d_result;
    a.sembl = d_sembl;
    a.stack = d_stack;
    a.not_used = d_not_used;
    if ((unsigned long long)mids->n * (unsigned long long)a.ntiles > 0x7fffffffULL)
        return HSS_ERR_INVALID_SIZE;
    hipStream_t st = (hipStream_t)stream;
    auto run = [&](const GridArgs &ga, GridShape sh, int slots) {
        GridArgs la = ga;
        la.tt = sh.tt;
        la.wg_per_mid = (a.ntiles + sh.tt - 1) / sh.tt;
        switch (model) {
            case HSS_CMP:
                if (sh.staged)
                    return launch_grid_cmp_staged(la, slots, variant, traces->storage, sh.cpl, st);
                return launch_grid_cmp(la, slots, variant, traces->storage, sh.cpl, st);
            case HSS_ZOCRS: return launch_grid_zocrs(la, slots, variant, traces->storage, sh.cpl, st);
            default: return launch_grid_oct(la, slots, variant, traces->storage, sh.cpl, st);
        }
    };
    // timing runs must not add to the accumulating rejection counts
    GridArgs probe = a;
    probe.not_used = nullptr;
    GridShape shape{1, 1, 0};
    rc = tuned_grid_shape(model, traces->w, variant, traces->storage, mids->n, a.ntiles, st,
                          [&](GridShape sh, int slots) { return run(probe, sh, slots); }, &shape);
    if (rc != HSS_OK) return rc;
    return run(a, shape, mids->n);
}

int hss_de_workspace_size(int32_t n_mid, int32_t ns, int model, int32_t NP, uint64_t *bytes) {
    const int np = model_param_count(model);
    if (np < 0) return HSS_ERR_INVALID_MODEL;
    if (!bytes) return HSS_ERR_NULL_POINTER;
    if (n_mid < 0 || ns <= 0 || NP <= 0) return HSS_ERR_INVALID_SIZE;
    *bytes = (uint64_t)n_mid * (uint64_t)tiles_of(ns) * 2ull * (uint64_t)NP * (uint64_t)(np + 2) *
             (uint64_t)kWave * sizeof(float);
    return HSS_OK;
}

int hss_de_semblance(const hss_traces *traces, const hss_midpoints *mids, int model,
                     const float *lower, const float *upper, int32_t NP, int32_t G,
                     int32_t gen_begin, int32_t gen_end, uint64_t seed, uint32_t stream_base,
                     int variant, void *d_workspace, uint64_t workspace_bytes, float *d_result,
                     float *d_pop_x, float *d_pop_fx, unsigned long long *d_not_used,
                     void *stream) {
    int rc = check_batch(traces, mids, model, variant);
    if (rc != HSS_OK) return rc;
    if (!lower || !upper || !d_result || !d_workspace) return HSS_ERR_NULL_POINTER;
    if (NP < 1 || G < 1) return HSS_ERR_INVALID_SIZE;
    if (gen_begin < 0 || gen_begin >= gen_end || gen_end > G) return HSS_ERR_INVALID_SIZE;
    uint64_t need = 0;
    rc = hss_de_workspace_size(mids->n, traces->ns, model, NP, &need);
    if (rc != HSS_OK) return rc;
    if (workspace_bytes < need) return HSS_ERR_INVALID_SIZE;
    if (mids->n == 0) return HSS_OK;
    DeArgs a{};
    a.e = base_ctx(traces, mids);
    a.mbeg = mids->begin;
    a.mcnt = mids->count;
    a.mm0 = mids->m0;
    a.morder = mids->order;
    a.nmid = mids->n;
    a.NP = NP;
    a.G = G;
    a.gen_begin = gen_begin;
    a.gen_end = gen_end;
    a.not_used = d_not_used;
    const int np = model_param_count(model);
    for (int d = 0; d < 3; ++d) {
        a.lo[d] = d < np ? lower[d] : 0.0f;
        a.hi[d] = d < np ? upper[d] : 0.0f;
    }
    a.key0 = (uint32_t)seed;
    a.key1 = (uint32_t)(seed >> 32);
    a.stream_base = stream_base;
    a.ntiles = tiles_of(traces->ns);
    a.ws = (float *)d_workspace;
    a.result = d_result;
    a.pop_x = d_pop_x;
    a.pop_fx = d_pop_fx;
    if ((unsigned long long)mids->n * (unsigned long long)a.ntiles > 0x7fffffffULL)
        return HSS_ERR_INVALID_SIZE;
    hipStream_t st = (hipStream_t)stream;
    switch (model) {
        case HSS_CMP: return launch_de_cmp(a, mids->n, variant, traces->storage, st);
        case HSS_ZOCRS: return launch_de_zocrs(a, mids->n, variant, traces->storage, st);
        default: return launch_de_oct(a, mids->n, variant, traces->storage, st);
    }
}

int hss_stretch_free(const hss_traces *traces, const hss_midpoints *mids, int model,
                     const float *d_vparams, const float *nvals, int32_t nN, int variant,
                     float *d_result, float *d_sembl, float *d_stack,
                     unsigned long long *d_not_used, void *stream) {
    int rc = check_batch(traces, mids, model, variant);
    if (rc != HSS_OK) return rc;
    if (!d_vparams || !nvals || !d_result) return HSS_ERR_NULL_POINTER;
    if (nN <= 0 || nN > kSfMaxShifts) return HSS_ERR_INVALID_SIZE;
    if (mids->n == 0) return HSS_OK;
    SfArgs a{};
    a.e = base_ctx(traces, mids);
    a.mbeg = mids->begin;
    a.mcnt = mids->count;
    a.mm0 = mids->m0;
    a.morder = mids->order;
    a.nmid = mids->n;
    a.vparams = d_vparams;
    for (int q = 0; q < nN; ++q) a.nvals[q] = nvals[q];
    a.nN = nN;
    a.ntiles = tiles_of(traces->ns);
    a.result = d_result;
    a.sembl = d_sembl;
    a.stack = d_stack;
    a.not_used = d_not_used;
    if ((unsigned long long)mids->n * (unsigned long long)a.ntiles > 0x7fffffffULL)
        return HSS_ERR_INVALID_SIZE;
    hipStream_t st = (hipStream_t)stream;
    switch (model) {
        case HSS_CMP: return launch_sf_cmp(a, mids->n, variant, traces->storage, st);
        case HSS_ZOCRS: return launch_sf_zocrs(a, mids->n, variant, traces->storage, st);
        default: return launch_sf_oct(a, mids->n, variant, traces->storage, st);
    }
}


static int adjoint_common(const hss_traces *traces, const hss_midpoints *mids, int model,
                          const float *d_points, float *d_out, const float *lower,
                          const float *upper, int32_t steps, bool polish, void *stream) {
    int rc = check_traces(traces);
    if (rc != HSS_OK) return rc;
    rc = check_midpoints(mids);
    if (rc != HSS_OK) return rc;
    const int np = model_param_count(model);
    if (np < 0) return HSS_ERR_INVALID_MODEL;
    if (!d_points || !d_out) return HSS_ERR_NULL_POINTER;
    AdjointArgs a{};
    a.e = base_ctx(traces, mids);
    a.mbeg = mids->begin;
    a.mcnt = mids->count;
    a.mm0 = mids->m0;
    a.morder = mids->order;
    a.nmid = mids->n;
    a.ntiles = tiles_of(traces->ns);
    a.points = d_points;
    a.out = d_out;
    if (polish) {
        if (!lower || !upper) return HSS_ERR_NULL_POINTER;
        if (steps < 0) return HSS_ERR_INVALID_SIZE;
        for (int d = 0; d < np; ++d) {
            if (!(lower[d] <= upper[d])) return HSS_ERR_INVALID_SIZE;
            if (d == 0 && !(lower[0] > 0.0f)) return HSS_ERR_INVALID_SIZE;
            a.lo[d] = lower[d];
            a.hi[d] = upper[d];
        }
        a.steps = steps;
    }
    if ((unsigned long long)mids->n * (unsigned long long)a.ntiles > 0x7fffffffULL)
        return HSS_ERR_INVALID_SIZE;
    if (mids->n == 0) return HSS_OK;
    hipStream_t st = (hipStream_t)stream;
    switch (model) {
        case HSS_CMP: return launch_adjoint_cmp(a, mids->n, traces->storage, polish, st);
        case HSS_ZOCRS: return launch_adjoint_zocrs(a, mids->n, traces->storage, polish, st);
        default: return launch_adjoint_oct(a, mids->n, traces->storage, polish, st);
    }
}

int hss_semblance_gradient(const hss_traces *traces, const hss_midpoints *mids, int model,
                           const float *d_points, float *d_out, void *stream) {
    return adjoint_common(traces, mids, model, d_points, d_out, nullptr, nullptr, 0, false, stream);
}

int hss_polish(const hss_traces *traces, const hss_midpoints *mids, int model, const float *lower,
               const float *upper, int32_t steps, const float *d_start, float *d_result,
               void *stream) {
    if (d_start == d_result) return HSS_ERR_INVALID_SIZE;
    return adjoint_common(traces, mids, model, d_start, d_result, lower, upper, steps, true, stream);
}


int hss_semblance_vjp(int32_t n, int32_t np, int32_t ns, const float *d_grad_S,
                      const float *d_grad_rows, float *d_grad_points, void *stream) {
    if (n < 0 || ns <= 0 || np < 1 || np > 3) return HSS_ERR_INVALID_SIZE;
    if (n == 0) return HSS_OK;
    if (!d_grad_S || !d_grad_rows || !d_grad_points) return HSS_ERR_NULL_POINTER;
    const unsigned long long total = (unsigned long long)n * (unsigned long long)np * (unsigned long long)ns;
    if (total > 0xffffffffULL) return HSS_ERR_INVALID_SIZE;
    const unsigned threads = 256, blocks = (unsigned)((total + threads - 1) / threads);
    hipLaunchKernelGGL(vjp_kernel, dim3(blocks), dim3(threads), 0, (hipStream_t)stream, n, np, ns,
                       d_grad_S, d_grad_rows, d_grad_points);
    return hip_status(hipGetLastError());
}

}  // extern "C"
