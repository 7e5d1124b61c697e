// The launch-shape autotuner of the grid search.
//
// Knob: the parameter tile, i.e. candidates per lane (CPL = 1 or 2; grid_impl.hpp).  Two
// candidates per lane share the trace loop, the geometry broadcast and (16-bit storage)
// the LDS segment passes, at the price of a second set of window registers; which wins
// depends on the window size, the storage and the variant.  The first launch for a key
// (device, model, w, variant, storage) that is large enough to time reliably runs every
// choice on the caller's own inputs, timed with HIP events on the caller's stream, and
// keeps the fastest; later launches reuse it.  The kernels are pure functions of their
// inputs, so the extra runs write the same outputs (accumulating outputs are withheld
// from them by the caller).  No tuning under stream capture (the default is used), and
// HSS_AUTOTUNE=0 in the environment turns tuning off.
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <tuple>

#include "launch_util.hpp"

namespace hss {

namespace {
std::mutex tune_mu;
std::map<std::tuple<int, int, int, int, int>, int> tune_cache;

bool tuning_enabled() {
    const char *v = std::getenv("HSS_AUTOTUNE");
    return !(v && std::strcmp(v, "0") == 0);
}
}  // namespace

int grid_default_cpl(int storage) { return storage == HSS_F32 ? 1 : 1; }

int tuned_grid_cpl(int model, int w, int variant, int storage, unsigned blocks, hipStream_t stream,
                   const std::function<int(int)> &launch, int *cpl_out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = -1;
    const auto key = std::make_tuple(dev, model, w, variant, storage);
    {
        std::lock_guard<std::mutex> lock(tune_mu);
        auto it = tune_cache.find(key);
        if (it != tune_cache.end()) {
            *cpl_out = it->second;
            return HSS_OK;
        }
    }
    *cpl_out = grid_default_cpl(storage);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone)
        return HSS_OK;
    // only launches with several waves of workgroups per CU time reliably
    if (!tuning_enabled() || blocks < 4096) return HSS_OK;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return HSS_OK;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return HSS_OK;
    }
    float best_ms = 0.0f;
    int best = *cpl_out;
    int rc = HSS_OK;
    for (int cpl = 1; cpl <= 2 && rc == HSS_OK; ++cpl) {
        rc = launch(cpl);  // warm-up (code load, caches)
        if (rc != HSS_OK) break;
        if (hipEventRecord(e0, stream) != hipSuccess) break;
        rc = launch(cpl);
        if (rc != HSS_OK) break;
        if (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) break;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) break;
        if (cpl == 1 || ms < best_ms) {
            best_ms = ms;
            best = cpl;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != HSS_OK) return rc;
    std::lock_guard<std::mutex> lock(tune_mu);
    tune_cache[key] = best;
    *cpl_out = best;
    return HSS_OK;
}

void tuned_grid_store(int dev, int model, int w, int variant, int storage, int cpl) {
    std::lock_guard<std::mutex> lock(tune_mu);
    const auto key = std::make_tuple(dev, model, w, variant, storage);
    if (cpl == 0)
        tune_cache.erase(key);
    else
        tune_cache[key] = cpl;
}

int tuned_grid_lookup(int dev, int model, int w, int variant, int storage) {
    std::lock_guard<std::mutex> lock(tune_mu);
    auto it = tune_cache.find(std::make_tuple(dev, model, w, variant, storage));
    return it == tune_cache.end() ? 0 : it->second;
}

}  // namespace hss

// The parameter tile the autotuner chose for a key on the current device (0: not tuned yet).
extern "C" int hss_grid_tuned_config(int model, int w, int variant, int storage, int32_t *cpl) {
    if (!cpl) return HSS_ERR_NULL_POINTER;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return HSS_ERR_NO_DEVICE;
    *cpl = hss::tuned_grid_lookup(dev, model, w, variant, storage);
    return HSS_OK;
}

// Pins the parameter tile for a key on the current device (cpl 1 or 2; 0 forgets the key,
// so the next large launch tunes it again).
extern "C" int hss_grid_set_config(int model, int w, int variant, int storage, int32_t cpl) {
    if (cpl < 0 || cpl > 2) return HSS_ERR_INVALID_SIZE;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return HSS_ERR_NO_DEVICE;
    hss::tuned_grid_store(dev, model, w, variant, storage, cpl);
    return HSS_OK;
}
