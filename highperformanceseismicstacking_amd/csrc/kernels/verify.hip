// Exhaustive check of the short correctly-rounded sequences of device_common.hpp.
//
// sqrt_rn(x) must equal sqrtf(x) for every non-negative float (and +inf) and be a NaN for
// every negative one, and
// div_rn(t, dt, RN(1/dt)) must equal t/dt for every float t wherever a window with
// tau >= 1 could accept either value: 1 <= t/dt < 2e9 (a window needs k = (int)(t/dt) >= tau,
// window_index()).  Below that range the two may differ -- for |t| < ~2^-103 the residual
// of div_rn is not exact, and -0 comes out as +0 -- but both values are then rejected
// alike, so the sample index and the weight never differ.  The weight of an accepted
// index, t/dt - (float)(int)(t/dt), is taken as v_fract_f32(t/dt); that too is compared
// for every float in [1, 2^24).  The launchers use the short sequences only for tau >= 1.
// One launch walks all 2^32 bit patterns; the result is cached per (device, dt).
#include <map>
#include <mutex>
#include <utility>

#include "launch_util.hpp"
#include "device_common.hpp"

namespace hss {

__device__ unsigned long long verify_mismatches;

__global__ __launch_bounds__(256) void verify_index_ops_kernel(float dt, float inv, int check_sqrt) {
    unsigned long long bad = 0;
    const unsigned stride = gridDim.x * blockDim.x;
    for (unsigned long long b = blockIdx.x * blockDim.x + threadIdx.x; b < (1ull << 32); b += stride) {
        const float x = __uint_as_float((unsigned)b);
        if (x != x) continue;  // NaN inputs never reach either sequence
        if (check_sqrt) {
            // same bits for x >= 0 (and -0); a NaN for x < 0, as sqrtf gives
            const float a = sqrt_rn(x);
            if (x >= 0.0f ? __float_as_uint(a) != __float_as_uint(sqrtf(x)) : a == a) bad++;
        }
        const float e = x / dt;
        const float f = div_rn(x, dt, inv);
        const bool ge = e >= 1.0f && e < 2.0e9f;
        const bool gf = f >= 1.0f && f < 2.0e9f;
        if (ge != gf || (ge && __float_as_uint(e) != __float_as_uint(f))) bad++;
        // the interpolation weight of an accepted index (t/dt in [1, 2^24)): v_fract_f32
        if (check_sqrt && x >= 1.0f && x < 0x1.0p24f) {
            const float w_ref = x - (float)(int)x;
            if (__float_as_uint(__builtin_amdgcn_fractf(x)) != __float_as_uint(w_ref)) bad++;
        }
    }
    if (bad) atomicAdd(&verify_mismatches, bad);
}

namespace {
std::mutex verify_mu;
std::map<std::pair<int, unsigned>, bool> verify_cache;  // completed checks only
}  // namespace

// True when the short sequences may replace sqrtf and t/dt for this (dt, tau) on the
// current device.  The first call per (device, dt) runs the check and waits for it.  A
// call on a capturing stream cannot wait, and a HIP call that fails says nothing about
// the sequences: both answer false (the library operations, the same bits) and cache
// nothing, so the next call outside capture runs the check.
bool fast_index_verified(float dt, int tau, hipStream_t stream) {
    // tau >= 1: no window accepts t/dt < 1; dt >= 2^-40: a traveltime below 2^-48 (a
    // radicand below 2^-96, where sqrt_rn_big may differ from sqrtf) has t/dt < 1
    if (tau < 1 || !(dt >= 0x1.0p-40f)) return false;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    unsigned bits;
    __builtin_memcpy(&bits, &dt, 4);
    std::lock_guard<std::mutex> lock(verify_mu);
    auto it = verify_cache.find({dev, bits});
    if (it != verify_cache.end()) return it->second;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone)
        return false;
    const float inv = 1.0f / dt;
    unsigned long long zero = 0, bad = 1;
    if (hipMemcpyToSymbolAsync(HIP_SYMBOL(verify_mismatches), &zero, sizeof(zero), 0,
                               hipMemcpyHostToDevice, stream) != hipSuccess)
        return false;
    hipLaunchKernelGGL(verify_index_ops_kernel, dim3(8192), dim3(256), 0, stream, dt, inv, 1);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyFromSymbolAsync(&bad, HIP_SYMBOL(verify_mismatches), sizeof(bad), 0,
                                 hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return false;
    const bool ok = bad == 0;
    verify_cache[{dev, bits}] = ok;
    return ok;
}

// 1: the check for dt passed on the current device; 0: it found a mismatch; -1: it has
// not completed there yet.
int index_ops_state(float dt) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    unsigned bits;
    __builtin_memcpy(&bits, &dt, 4);
    std::lock_guard<std::mutex> lock(verify_mu);
    auto it = verify_cache.find({dev, bits});
    return it == verify_cache.end() ? -1 : (it->second ? 1 : 0);
}

}  // namespace hss

// Mismatch count of the check for dt on the current device (tests and diagnostics).
extern "C" int hss_verify_index_ops(float dt, unsigned long long *mismatches, void *stream) {
    using namespace hss;
    if (!mismatches) return HSS_ERR_NULL_POINTER;
    if (!(dt > 0.0f)) return HSS_ERR_INVALID_SIZE;
    hipStream_t st = (hipStream_t)stream;
    const unsigned long long zero = 0;
    hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(verify_mismatches), &zero, sizeof(zero), 0,
                                          hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(verify_index_ops_kernel, dim3(8192), dim3(256), 0, st, dt, 1.0f / dt, 1);
    e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    e = hipMemcpyFromSymbolAsync(mismatches, HIP_SYMBOL(verify_mismatches), sizeof(*mismatches), 0,
                                 hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return (int)e;
    return hip_status(hipStreamSynchronize(st));
}

// Whether the semblance entry points use the short sequences for dt on the current device.
extern "C" int hss_index_ops_state(float dt, int *state) {
    if (!state) return HSS_ERR_NULL_POINTER;
    if (!(dt > 0.0f)) return HSS_ERR_INVALID_SIZE;
    *state = hss::index_ops_state(dt);
    return HSS_OK;
}
