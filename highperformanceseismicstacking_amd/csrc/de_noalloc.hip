// hss_de_semblance_noalloc: differential evolution whose population workspace is a
// file-scope device array (`scratch`) instead of a caller-provided buffer, for callers
// that cannot allocate device memory (code under graph capture, serving loops, harnesses
// that forbid hipMalloc).  It allocates nothing: the array is part of the loaded code
// object, one copy per device, held until the process ends.
//
// The host never uses the array's name as a device address (host code sees only a
// stand-in symbol): on the first call on a device, with that device current,
// hipGetSymbolAddress gives that device's copy, and the address is kept per device.
#include <mutex>

#include <hip/hip_runtime.h>

#include "../../include/hss.h"
#include "kernels/launch_util.hpp"

static __device__ __align__(256) unsigned char scratch[HSS_DE_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void *scratch_ptr[kMaxDevices];  // device address of `scratch` on each device, once known

// The current device's copy of `scratch` (nullptr with *status set on failure).
void *device_scratch(int *status) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *status = (int)e;
        return nullptr;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        *status = HSS_ERR_NO_DEVICE;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *status = e != hipSuccess ? (int)e : HSS_ERR_NO_DEVICE;
            return nullptr;
        }
        scratch_ptr[dev] = p;
    }
    *status = HSS_OK;
    return scratch_ptr[dev];
}

}  // namespace

extern "C" int hss_de_semblance_noalloc(const hss_traces *traces, const hss_midpoints *mids,
                                        int model, const float *lower, const float *upper,
                                        int32_t NP, int32_t G, int32_t gen_begin, int32_t gen_end,
                                        uint64_t seed, uint32_t stream_base, int variant,
                                        float *d_result, float *d_pop_x, float *d_pop_fx,
                                        unsigned long long *d_not_used, void *stream) {
    if (!traces || !mids) return HSS_ERR_NULL_POINTER;
    int rc = hss::check_traces(traces);
    if (rc != HSS_OK) return rc;
    rc = hss::check_midpoints(mids);
    if (rc != HSS_OK) return rc;
    uint64_t need = 0;
    rc = hss_de_workspace_size(mids->n, traces->ns, model, NP, &need);
    if (rc != HSS_OK) return rc;
    if (need > (uint64_t)HSS_DE_SCRATCH_BYTES) return HSS_ERR_SCRATCH_LIMIT;
    void *ws = device_scratch(&rc);
    if (!ws) return rc;
    // the same validation and launch as the caller-workspace entry point; a run split
    // over generation ranges resumes from the populations this device's scratch holds
    return hss_de_semblance(traces, mids, model, lower, upper, NP, G, gen_begin, gen_end, seed,
                            stream_base, variant, ws, (uint64_t)HSS_DE_SCRATCH_BYTES, d_result,
                            d_pop_x, d_pop_fx, d_not_used, stream);
}
