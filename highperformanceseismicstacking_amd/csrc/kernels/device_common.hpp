// Device-side building blocks of the semblance hot path (gfx950).
//
// Every translation unit that includes this header is compiled with
// -ffp-contract=off: a float expression is evaluated as written, one IEEE-rounded
// operation per operator, in the order of the reference source text.  Where a fast
// variant wants a fused multiply-add it calls __builtin_fmaf explicitly.
// sqrtf and '/' are the correctly rounded HIP defaults.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/hss.h"

namespace hss {

constexpr int kWave = 64;
constexpr int kMaxWindow = 32;  // MAX_WINDOW_SIZE, common/include/gpu/interface.h:10

// A float in LDS.  Pointers of this type compile to ds_read/ds_write; a plain float*
// into a __shared__ array is a generic pointer, and code that merges it with a global
// one falls back to flat loads, which go through the vector memory path.
using LdsFloat = __attribute__((address_space(3))) float;

// Wave segment (eval.hpp): kSeg samples per wave in LDS.
constexpr int kSeg = 128;
constexpr int kSegFloats = kSeg;

enum ErrorCode {  // gpu_error_code, common/include/gpu/interface.h:12-18
    kNoError = 0,
    kDivisionByZero = 1,
    kNegativeSquaredRoot = 2,
};

// ---------------------------------------------------------------------------
// Correctly rounded square root and division by the sample period with fewer
// instructions than the library sequences.  Both must return the same bits as sqrtf and
// '/': kernels use them only after verify.hip has compared them with the library
// operations on every float input (for the division: for this dt), and fall back to the
// library operations otherwise (see fast_index_verified()).
//
// sqrt: v_sqrt_f32 is within one ulp; of s and its two neighbours the correctly rounded
// root is picked from the signs of the exact residuals x - s'*s (one FMA each), the
// correction step of the library sequence without its denormal scaling; inputs below
// 2^-96 take sqrtf.
// sqrt_rn_big: the sequence alone, for inputs known to be 0 or >= 2^-96.
__device__ __forceinline__ float sqrt_rn_big(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float s_dn = __int_as_float(__float_as_int(s) - 1);
    const float s_up = __int_as_float(__float_as_int(s) + 1);
    const float r_dn = __builtin_fmaf(-s_dn, s, x);
    const float r_up = __builtin_fmaf(-s_up, s, x);
    float r = r_dn <= 0.0f ? s_dn : s;
    r = r_up > 0.0f ? s_up : r;
    return r;
}

__device__ __forceinline__ float sqrt_rn(float x) {
    if (x < 0x1.0p-96f) return sqrtf(x);
    return sqrt_rn_big(x);
}

// t / dt from inv = RN(1/dt): q = t*inv, then one correction with the exact residual.
__device__ __forceinline__ float div_rn(float t, float dt, float inv) {
    const float q = t * inv;
    const float r = __builtin_fmaf(-q, dt, t);
    return __builtin_fmaf(r, inv, q);
}

// ---------------------------------------------------------------------------
// OCT displaced midpoint (reference: cuda/include/semblance/kernel/oct/common.cuh:7-35).
// FAST: the square roots use sqrt_rn (same bits as sqrtf, see above).
template <bool FAST>
__device__ __forceinline__ float sqrt_of(float x) {
    if constexpr (FAST) return sqrt_rn(x);
    else return sqrtf(x);
}

template <bool FAST = false>
__device__ __forceinline__ int oct_displaced_midpoint(float h, float h0, float t0, float m0,
                                                      float c, float slope, float *mh) {
    float theta = t0 * slope;
    float gamma = 2.0f * sqrt_of<FAST>(h * h + h0 * h0);
    float tn0_sq = t0 * t0 - c * h0 * h0;
    float theta_sq = theta * theta;
    float gamma_sq = gamma * gamma;
    float tn0_quad = tn0_sq * tn0_sq;
    float arg = tn0_quad * tn0_quad + tn0_quad * theta_sq * gamma_sq +
                16.0f * h * h * h0 * h0 * theta_sq * theta_sq;
    if (arg < 0) return kNegativeSquaredRoot;
    arg = theta_sq * gamma_sq + 2.0f * tn0_quad + 2.0f * sqrt_of<FAST>(arg);
    if (arg == 0) return kDivisionByZero;
    *mh = m0 + 2.0f * theta * (h * h - h0 * h0) / sqrt_of<FAST>(arg);
    return kNoError;
}

// OCT traveltime along the trajectory (reference: oct/common.cuh:37-94).
template <bool FAST = false>
__device__ __forceinline__ int oct_time(float h, float h0, float t0, float m0, float m, float mh,
                                        float c, float slope, float *out) {
    float tn0_sq = t0 * t0 - c * h0 * h0;
    float w1 = (h + h0) * (h + h0) - (mh - m0) * (mh - m0);
    float w2 = (h - h0) * (h - h0) - (mh - m0) * (mh - m0);
    if (w1 < 0 || w2 < 0) return kNegativeSquaredRoot;
    float u = sqrt_of<FAST>(w1) + sqrt_of<FAST>(w2);
    float th = t0;
    if (fabsf(h) > fabsf(h0)) {
        if (u == 0.0f) return kDivisionByZero;
        float arg = c * h * h + 4.0f * h * h / (u * u) * tn0_sq;
        if (arg < 0) return kNegativeSquaredRoot;
        th = sqrt_of<FAST>(arg);
    } else if (fabsf(h) < fabsf(h0)) {
        if (h0 == 0.0f) return kDivisionByZero;
        float arg = c * h * h + u * u / (4.0f * h0 * h0) * tn0_sq;
        if (arg < 0) return kNegativeSquaredRoot;
        th = sqrt_of<FAST>(arg);
    }
    float tn_sq = th * th - c * h * h;
    if (th == 0.0f || tn0_sq == 0.0f) return kDivisionByZero;
    float ah = (t0 * tn_sq) / (th * tn0_sq) * slope;
    *out = th + ah * (m - mh);
    return kNoError;
}

// ---------------------------------------------------------------------------
// A candidate in storage form: CMP {c}; ZO-CRS {c, A, B}; OCT {c, slope}.
struct Candidate {
    float c, p1, p2;
};

// Per-trace geometry a traveltime needs.
struct TraceGeom {
    float m, h, h2;
};

// Traveltime of one trace for one candidate at zero-offset time t0.  Returns true when
// the trace enters the window sum; sets *reject when the model rejects the trace the way
// the reference counts in notUsedCount (OCT only).
//   CMP    cmp/linear_search.cu:60-62
//   ZO-CRS zocrs/linear_search.cu:83-90
//   OCT    oct/linear_search.cu:84-98
// t0sq is t0*t0 rounded, hoisted out of the trace loop.  FAST: square roots by sqrt_rn.
// FAST (used only with tau >= 1 and dt >= 2^-40, see fast_index_verified()): the CMP and
// ZO-CRS square root is the traveltime itself, so a radicand below 2^-96 -- the only
// inputs where sqrt_rn_big may differ from sqrtf -- gives t < 2^-48 <= dt, t/dt < 1, and
// no window with tau >= 1 accepts it either way; the small-input branch is not needed.
template <int MODEL, bool FAST = false>
__device__ __forceinline__ bool traveltime(const TraceGeom &tg, float t0, float t0sq, float m0,
                                           float h0, float apm, const Candidate &k, float *t,
                                           bool *reject) {
    *reject = false;
    if constexpr (MODEL == HSS_CMP) {
        if constexpr (FAST) *t = sqrt_rn_big(t0sq + k.c * tg.h2);
        else *t = sqrtf(t0sq + k.c * tg.h2);
        return true;
    } else if constexpr (MODEL == HSS_ZOCRS) {
        float dm = tg.m - m0;
        float tmp = t0 + k.p1 * dm;
        tmp = tmp * tmp + k.p2 * dm * dm + k.c * tg.h2;
        if (tmp >= 0) {
            if constexpr (FAST) *t = sqrt_rn_big(tmp);
            else *t = sqrtf(tmp);
            return true;
        }
        return false;
    } else {
        float mh;
        int err = oct_displaced_midpoint<FAST>(tg.h, h0, t0, m0, k.c, k.p1, &mh);
        if (err != kNoError || fabsf(tg.m - mh) > apm) {
            *reject = true;
            return false;
        }
        err = oct_time<FAST>(tg.h, h0, t0, m0, tg.m, mh, k.c, k.p1, t);
        return err == kNoError;
    }
}

// Sample index and interpolation weight (COMPUTE_SEMBLANCE, interface.h:27-31):
// tIndex = t/dt, k = (int)tIndex, frac = tIndex - k, trace kept when
// k - tau >= 0 && k + tau + 1 < ns.  A NaN, infinite or out-of-int-range index drops
// the trace (the reference's cast is undefined there); for every other value the
// decision is the reference's.  On success *first = k - tau.  FAST: t/dt by div_rn with
// inv = RN(1/dt) (same bits, see above).
//
// The integer test is done on tIndex itself: with K = ns - tau - 1,
//   tau >= 1:  trunc(x) >= tau  <=>  x >= tau  <=>  x > lo, lo = the float below tau;
//   tau == 0:  trunc(x) >= 0    <=>  x > -1 = lo;
//   K >= 1:    trunc(x) < K     <=>  x < K = hi  (for x > -1);  K <= 0: nothing is kept,
//              hi = -2.
// NaN fails both compares.  index_bounds() computes (lo, hi) on the host; ns < 2^24
// keeps hi exact.
struct IndexBounds {
    float lo, hi;
};

__host__ __device__ inline IndexBounds index_bounds(int ns, int tau) {
    IndexBounds b;
    b.lo = tau == 0 ? -1.0f : __builtin_nextafterf((float)tau, 0.0f);
    const int K = ns - tau - 1;
    b.hi = K >= 1 ? (float)K : -2.0f;
    return b;
}

template <bool FAST = false>
__device__ __forceinline__ bool window_index(float t, float dt, float inv, IndexBounds ib,
                                             int tau, int *first, float *frac) {
    float ti;
    if constexpr (FAST) ti = div_rn(t, dt, inv);
    else ti = t / dt;
    const int k = (int)ti;
    *frac = ti - (float)k;
    *first = k - tau;
    return ti > ib.lo && ti < ib.hi;
}

// The same decision returning the sample index k itself (first = k - tau).  FAST (the
// verified short sequences, used only with tau >= 1, so an accepted t/dt is >= 1): the
// weight is v_fract_f32(t/dt) = t/dt - floor(t/dt), the same bits as t/dt - (float)k for
// every t/dt >= 0 (checked exhaustively on the device with the other short sequences).
template <bool FAST = false>
__device__ __forceinline__ bool window_k(float t, float dt, float inv, IndexBounds ib, int *k_out,
                                         float *frac) {
    float ti;
    if constexpr (FAST) ti = div_rn(t, dt, inv);
    else ti = t / dt;
    const int k = (int)ti;
    if constexpr (FAST) *frac = __builtin_amdgcn_fractf(ti);
    else *frac = ti - (float)k;
    *k_out = k;
    return ti > ib.lo && ti < ib.hi;
}

// ---------------------------------------------------------------------------
// Sample fetch for each storage type.  F16/BF16 values are widened to float and
// multiplied by the trace's scale; F32 is read as is (times scale when one is given).
template <int STORAGE>
struct SampleReader;

template <>
struct SampleReader<HSS_F32> {
    const float *__restrict__ p;
    __device__ __forceinline__ float operator()(size_t i) const { return p[i]; }
};

template <>
struct SampleReader<HSS_F16> {
    const _Float16 *__restrict__ p;
    __device__ __forceinline__ float operator()(size_t i) const { return (float)p[i]; }
};

template <>
struct SampleReader<HSS_BF16> {
    const uint16_t *__restrict__ p;
    __device__ __forceinline__ float operator()(size_t i) const {
        return __uint_as_float((uint32_t)p[i] << 16);
    }
};

// ---------------------------------------------------------------------------
// Workgroup order.  The dispatcher deals workgroups to the chip's 8 XCDs round-robin
// (workgroup b runs on XCD b mod 8) and each XCD has its own L2.  logical_block()
// renumbers them so that every XCD receives one contiguous range of logical ids: the t0
// tiles of one midpoint, which read the same traces, then share one L2 instead of
// fetching those traces into all eight.  A bijection on [0, n).
constexpr int kXcds = 8;

__device__ __forceinline__ int logical_block(int b, int n) {
    const int q = n / kXcds, r = n - (n / kXcds) * kXcds;
    const int x = b % kXcds, i = b / kXcds;
    return x * q + min(x, r) + i;
}

// ---------------------------------------------------------------------------
// Wave-level helpers.
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

}  // namespace hss
