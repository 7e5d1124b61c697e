// The launch-shape autotuner of the grid search.
//
// Three knobs (grid_impl.hpp, grid_staged.hpp), none of which changes a result bit:
//  * the parameter tile, candidates per lane (CPL = 1 or 2): two candidates share the
//    trace loop, the geometry broadcast and (16-bit storage) the LDS segment passes, at
//    the price of a second set of window registers;
//  * the t0 tile, 64-sample tiles per workgroup (TT = 1, 2 or 4): the workgroup's four
//    waves split into TT tiles x 4/TT candidate groups, trading the number of workgroups
//    per midpoint (and so how a midpoint's work spreads over the CUs of its XCD) against
//    the candidates each wave walks;
//  * staging (CMP, one t0 tile per workgroup): the four waves walk the traces together and
//    read their windows from the workgroup's LDS copy of each trace's span, instead of
//    each lane (FP32) or each wave (16-bit) reading its own.
// Which shape wins depends on the model, the window size, the storage and the variant.
// The first launch for a key (device, model, w, variant, storage) that is large enough
// to time reliably runs every shape on a leading part of the caller's own batch (its
// first launch slots, about 8192 workgroups), timed with HIP events on the caller's
// stream, and keeps the fastest (ties within 1% to the preferred shape); later launches
// reuse it.  The kernels are pure functions
// of their inputs, so the timing runs write the same outputs (accumulating outputs are
// withheld from them by the caller).  No tuning under stream capture (the default shape
// is used), and HSS_AUTOTUNE=0 in the environment turns tuning off.
#include <algorithm>
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
std::map<std::tuple<int, int, int, int, int>, GridShape> tune_cache;

bool tuning_enabled() {
    const char *v = std::getenv("HSS_AUTOTUNE");
    return !(v && std::strcmp(v, "0") == 0);
}

// In order of preference: a later shape replaces the best so far only when it is faster by
// more than kTuneMargin, so near-ties (common: two candidates per lane often time within a
// fraction of a percent of one) resolve the same way on every run, to the shape listed
// first -- two candidates per lane first, whose four resident workgroups per CU keep a
// midpoint's traces in the XCD's L2 (one candidate per lane, at six, can fetch many times
// the line past it for the same time).
constexpr GridShape kShapes[] = {{2, 1, 0}, {1, 1, 0}, {2, 2, 0}, {1, 2, 0},
                                 {2, 4, 0}, {1, 4, 0}, {2, 1, 1}, {1, 1, 1}};
constexpr float kTuneMargin = 0.01f;
// workgroups a timing run covers (a leading part of the batch); the warm-up run before it
// (code load) covers a sixteenth of that
constexpr unsigned kTuneBlocks = 8192;
}  // namespace

int tuned_grid_shape(int model, int w, int variant, int storage, int n_slots, int ntiles,
                     hipStream_t stream, const std::function<int(GridShape, int)> &launch,
                     GridShape *shape_out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = -1;
    const auto key = std::make_tuple(dev, model, w, variant, storage);
    {
        std::lock_guard<std::mutex> lock(tune_mu);
        auto it = tune_cache.find(key);
        if (it != tune_cache.end()) {
            *shape_out = it->second;
            return HSS_OK;
        }
    }
    // untuned launches (small, under capture, tuning off): the choice of every tuned line
    // -- two candidates per lane, for CMP with the workgroup-staged traces -- but for the
    // compensated variant, which keeps one candidate and direct reads (its two-candidate
    // and staged forms spill)
    *shape_out = variant == HSS_KAHAN ? GridShape{1, 1, 0}
                                      : GridShape{2, 1, model == HSS_CMP ? 1 : 0};
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone)
        return HSS_OK;
    // only launches with several waves of workgroups per CU time reliably
    const unsigned long long blocks = (unsigned long long)n_slots * (unsigned long long)ntiles;
    if (!tuning_enabled() || blocks < 4096) return HSS_OK;
    const int slots = (int)std::min<unsigned long long>(
        (unsigned long long)n_slots, (kTuneBlocks + ntiles - 1) / (unsigned)ntiles);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return HSS_OK;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return HSS_OK;
    }
    float best_ms = 0.0f;
    GridShape best = *shape_out;
    bool timed = false;
    int rc = HSS_OK;
    for (const GridShape &sh : kShapes) {
        if (sh.staged && model != HSS_CMP) continue;
        rc = launch(sh, std::max(1, slots / 16));  // warm-up (code load)
        if (rc != HSS_OK) break;
        if (hipEventRecord(e0, stream) != hipSuccess) break;
        rc = launch(sh, slots);
        if (rc != HSS_OK) break;
        if (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) break;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) break;
        if (!timed || ms < best_ms * (1.0f - kTuneMargin)) {
            best_ms = ms;
            best = sh;
            timed = true;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != HSS_OK) return rc;
    if (!timed) return HSS_OK;  // a HIP call failed: keep the default, cache nothing
    std::lock_guard<std::mutex> lock(tune_mu);
    tune_cache[key] = best;
    *shape_out = best;
    return HSS_OK;
}

void tuned_grid_store(int dev, int model, int w, int variant, int storage, GridShape shape) {
    std::lock_guard<std::mutex> lock(tune_mu);
    const auto key = std::make_tuple(dev, model, w, variant, storage);
    if (shape.cpl == 0)
        tune_cache.erase(key);
    else
        tune_cache[key] = shape;
}

GridShape tuned_grid_lookup(int dev, int model, int w, int variant, int storage) {
    std::lock_guard<std::mutex> lock(tune_mu);
    auto it = tune_cache.find(std::make_tuple(dev, model, w, variant, storage));
    return it == tune_cache.end() ? GridShape{0, 0, 0} : it->second;
}

}  // namespace hss

// The launch shape the autotuner chose for a key on the current device (all 0: not yet).
extern "C" int hss_grid_tuned_shape(int model, int w, int variant, int storage, int32_t *cpl,
                                    int32_t *t0_tiles, int32_t *staged) {
    if (!cpl || !t0_tiles || !staged) return HSS_ERR_NULL_POINTER;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return HSS_ERR_NO_DEVICE;
    const hss::GridShape sh = hss::tuned_grid_lookup(dev, model, w, variant, storage);
    *cpl = sh.cpl;
    *t0_tiles = sh.tt;
    *staged = sh.staged;
    return HSS_OK;
}

extern "C" int hss_grid_tuned_config(int model, int w, int variant, int storage, int32_t *cpl) {
    int32_t tt = 0, staged = 0;
    return hss_grid_tuned_shape(model, w, variant, storage, cpl, &tt, &staged);
}

// Pins the launch shape for a key on the current device; cpl = 0 forgets the key (the
// next large launch tunes it again).
extern "C" int hss_grid_set_shape(int model, int w, int variant, int storage, int32_t cpl,
                                  int32_t t0_tiles, int32_t staged) {
    if (cpl < 0 || cpl > 2) return HSS_ERR_INVALID_SIZE;
    if (cpl > 0 && t0_tiles != 1 && t0_tiles != 2 && t0_tiles != 4) return HSS_ERR_INVALID_SIZE;
    if (staged != 0 && staged != 1) return HSS_ERR_INVALID_SIZE;
    if (cpl > 0 && staged && t0_tiles != 1) return HSS_ERR_INVALID_SIZE;
    if (cpl > 0 && staged && model != HSS_CMP) return HSS_ERR_INVALID_MODEL;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return HSS_ERR_NO_DEVICE;
    hss::tuned_grid_store(dev, model, w, variant, storage, hss::GridShape{cpl, t0_tiles, staged});
    return HSS_OK;
}

extern "C" int hss_grid_set_config(int model, int w, int variant, int storage, int32_t cpl) {
    return hss_grid_set_shape(model, w, variant, storage, cpl, 1, 0);
}
