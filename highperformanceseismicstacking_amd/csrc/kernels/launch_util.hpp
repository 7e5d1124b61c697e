// Host-side helpers shared by the C-ABI entry points: argument validation and the
// compile-time window-size dispatch.
#pragma once

#include <functional>

#include <hip/hip_runtime.h>

#include "../../../include/hss.h"

namespace hss {

inline int model_param_count(int model) {
    switch (model) {
        case HSS_CMP: return 1;
        case HSS_ZOCRS: return 3;
        case HSS_OCT: return 2;
        default: return -1;
    }
}

inline int check_traces(const hss_traces *t) {
    if (!t || !t->samples || !t->mid || !t->hoff || !t->hoff2) return HSS_ERR_NULL_POINTER;
    if (t->ntr < 0 || t->ns <= 0 || t->ns > (1 << 24)) return HSS_ERR_INVALID_SIZE;
    if (t->tau < 0 || t->w != 2 * t->tau + 1 || t->w > 32) return HSS_ERR_INVALID_WINDOW;
    if (t->storage < HSS_F32 || t->storage > HSS_BF16) return HSS_ERR_INVALID_STORAGE;
    // the kernels read samples in aligned 4-byte words (two 16-bit samples per word)
    if (((uintptr_t)t->samples & 3u) != 0) return HSS_ERR_INVALID_STORAGE;
    if (!(t->dt > 0.0f)) return HSS_ERR_INVALID_SIZE;
    return HSS_OK;
}

inline int check_midpoints(const hss_midpoints *m) {
    if (!m) return HSS_ERR_NULL_POINTER;
    if (m->n < 0) return HSS_ERR_INVALID_SIZE;
    if (m->n > 0 && (!m->begin || !m->count || !m->m0)) return HSS_ERR_NULL_POINTER;
    return HSS_OK;
}

inline int hip_status(hipError_t e) { return e == hipSuccess ? HSS_OK : (int)e; }

// True when sqrt_rn / div_rn / v_fract_f32 give the library's sample index and weight for
// this dt and tau on the current device: tau >= 1, dt >= 2^-40 and the exhaustive check
// of verify.hip found no mismatch (run once per device and dt, waiting on `stream` the
// first time).
bool fast_index_verified(float dt, int tau, hipStream_t stream);

// Grid-search launch shape: candidates per lane (cpl, 1 or 2), 64-sample t0 tiles per
// workgroup (tt, 1, 2 or 4) and, for CMP with tt = 1, traces staged in LDS for the whole
// workgroup (staged, grid_staged.hpp).  Results do not depend on it.
struct GridShape {
    int cpl;
    int tt;
    int staged;
};

// The shape for this key, tuned on the first large launch by timing launch(shape, slots)
// for each shape on the first `slots` launch slots of the batch (tune.hip).
int tuned_grid_shape(int model, int w, int variant, int storage, int n_slots, int ntiles,
                     hipStream_t stream, const std::function<int(GridShape, int)> &launch,
                     GridShape *shape_out);

// Calls f(std::integral_constant<int, W>) for the compiled window sizes; any other w
// (odd, <= 32), or fast = false, goes to the runtime-w instantiation W = 0.
template <class F>
inline void dispatch_window(int w, bool fast, F &&f) {
    if (!fast) {
        f(std::integral_constant<int, 0>{});
        return;
    }
    switch (w) {
        case 5: f(std::integral_constant<int, 5>{}); break;
        case 7: f(std::integral_constant<int, 7>{}); break;
        case 9: f(std::integral_constant<int, 9>{}); break;
        case 11: f(std::integral_constant<int, 11>{}); break;
        case 13: f(std::integral_constant<int, 13>{}); break;
        case 15: f(std::integral_constant<int, 15>{}); break;
        default: f(std::integral_constant<int, 0>{}); break;
    }
}

template <class F>
inline int dispatch_model(int model, F &&f) {
    switch (model) {
        case HSS_CMP: f(std::integral_constant<int, HSS_CMP>{}); return HSS_OK;
        case HSS_ZOCRS: f(std::integral_constant<int, HSS_ZOCRS>{}); return HSS_OK;
        case HSS_OCT: f(std::integral_constant<int, HSS_OCT>{}); return HSS_OK;
        default: return HSS_ERR_INVALID_MODEL;
    }
}

template <class F>
inline int dispatch_storage(int storage, F &&f) {
    switch (storage) {
        case HSS_F32: f(std::integral_constant<int, HSS_F32>{}); return HSS_OK;
        case HSS_F16: f(std::integral_constant<int, HSS_F16>{}); return HSS_OK;
        case HSS_BF16: f(std::integral_constant<int, HSS_BF16>{}); return HSS_OK;
        default: return HSS_ERR_INVALID_STORAGE;
    }
}

template <class F>
inline int dispatch_variant(int variant, F &&f) {
    switch (variant) {
        case HSS_EXACT: f(std::integral_constant<int, HSS_EXACT>{}); return HSS_OK;
        case HSS_FAST: f(std::integral_constant<int, HSS_FAST>{}); return HSS_OK;
        case HSS_KAHAN: f(std::integral_constant<int, HSS_KAHAN>{}); return HSS_OK;
        default: return HSS_ERR_INVALID_VARIANT;
    }
}

}  // namespace hss
