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
// root is picked from the signs of the exact residuals x - s'*s (one FMA each).
// sqrt_rn_big: that sequence alone, for inputs known to be 0 or >= 2^-96 (below, the
// residual products can underflow).
// sqrt_rn: inputs below 2^-96 (negative ones included) are first scaled by 2^64, exactly,
// into that range, and the root by 2^-32 back (exact: the root of a positive float is
// >= 2^-75, a normal number).  The scaling runs only when some active lane of the wave
// has such an input (a wave-uniform branch); it is the identity for the other lanes.
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
    const bool small = x < 0x1.0p-96f;
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(small) == 0, 1)) return sqrt_rn_big(x);
    const float r = sqrt_rn_big(__builtin_amdgcn_ldexpf(x, small ? 64 : 0));
    return __builtin_amdgcn_ldexpf(r, small ? -32 : 0);
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
// ZO-CRS in two parts, the reference's expression in its own operation order: the
// midpoint part (t0 + A*dm)^2 + B*dm*dm, then + C*h^2 and the square root.  The first
// depends on the trace's midpoint only, so a loop over a CMP's traces can keep it.
__device__ __forceinline__ float zocrs_mid_part(float t0, float dm, const Candidate &k) {
    const float tmp = t0 + k.p1 * dm;
    return tmp * tmp + k.p2 * dm * dm;
}

template <bool FAST>
__device__ __forceinline__ bool zocrs_time_from(float mid_part, float c_h2, float *t) {
    const float tmp = mid_part + c_h2;
    if (tmp >= 0) {
        if constexpr (FAST) *t = sqrt_rn_big(tmp);
        else *t = sqrtf(tmp);
        return true;
    }
    return false;
}

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
        return zocrs_time_from<FAST>(zocrs_mid_part(t0, tg.m - m0, k), k.c * tg.h2, t);
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

// The OCT traveltime split by what each term depends on, for loops over (trace,
// candidate) pairs.  OctTrace holds the terms of one trace (its half offset h and the
// run's h0): a wave computes them once per trace instead of once per candidate and lane.
// OctCand holds the terms of one candidate at one t0.  Every term is the expression the
// reference spells (oct/common.cuh:7-94), with its operations in the same order as in
// oct_displaced_midpoint() and oct_time() above, so oct_mid(), oct_time_terms() and
// oct_traveltime_at() give the same bits and the same decisions as traveltime<HSS_OCT>().
struct OctTrace {
    float h, m;
    float gamma_sq;  // gamma*gamma, gamma = 2*sqrt(h*h + h0*h0)
    float p16;       // 16*h*h*h0*h0, left to right
    float hd;        // h*h - h0*h0
    float e1, e2;    // (h+h0)*(h+h0), (h-h0)*(h-h0)
    float f4;        // 4*h*h, left to right
    int branch;      // 1: |h| > |h0|; 2: |h| < |h0|; 0: equal
};

template <bool FAST>
__device__ __forceinline__ OctTrace oct_trace(float h, float m, float h0) {
    OctTrace o;
    o.h = h;
    o.m = m;
    const float gamma = 2.0f * sqrt_of<FAST>(h * h + h0 * h0);
    o.gamma_sq = gamma * gamma;
    o.p16 = 16.0f * h * h * h0 * h0;
    o.hd = h * h - h0 * h0;
    o.e1 = (h + h0) * (h + h0);
    o.e2 = (h - h0) * (h - h0);
    o.f4 = 4.0f * h * h;
    o.branch = fabsf(h) > fabsf(h0) ? 1 : (fabsf(h) < fabsf(h0) ? 2 : 0);
    return o;
}

// Lane jj's value, in a scalar register of the wave.
__device__ __forceinline__ float bcast(float v, int jj) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), jj));
}

// Wave-uniform copies of lane jj's offset-dependent terms: those oct_mid reads and those
// oct_time_terms reads.

__device__ __forceinline__ void bcast_oct_mid_terms(OctTrace &ot, const OctTrace &o, int jj) {
    ot.gamma_sq = bcast(o.gamma_sq, jj);
    ot.p16 = bcast(o.p16, jj);
    ot.hd = bcast(o.hd, jj);
}

__device__ __forceinline__ void bcast_oct_time_terms(OctTrace &ot, const OctTrace &o, int jj) {
    ot.e1 = bcast(o.e1, jj);
    ot.e2 = bcast(o.e2, jj);
    ot.f4 = bcast(o.f4, jj);
    ot.branch = __builtin_amdgcn_readlane(o.branch, jj);
}

struct OctCand {
    float c, slope, t0;
    float theta, theta_sq, two_theta;
    float tn0_sq;    // t0*t0 - c*h0*h0
    float tq_sq;     // tn0_quad*tn0_quad, tn0_quad = tn0_sq*tn0_sq
    float tq_th;     // tn0_quad*theta_sq
    float two_tq;    // 2*tn0_quad
    float q4;        // 4*h0*h0
};

__device__ __forceinline__ OctCand oct_cand(float t0, float h0, const Candidate &k) {
    OctCand o;
    o.c = k.c;
    o.slope = k.p1;
    o.t0 = t0;
    o.theta = t0 * k.p1;
    o.theta_sq = o.theta * o.theta;
    o.two_theta = 2.0f * o.theta;
    o.tn0_sq = t0 * t0 - k.c * h0 * h0;
    const float tn0_quad = o.tn0_sq * o.tn0_sq;
    o.tq_sq = tn0_quad * tn0_quad;
    o.tq_th = tn0_quad * o.theta_sq;
    o.two_tq = 2.0f * tn0_quad;
    o.q4 = 4.0f * h0 * h0;
    return o;
}

// traveltime<HSS_OCT>() from the split terms, in two halves that depend on the trace's
// half offset (through OctTrace) and the candidate only: the displaced midpoint
// (oct_mid) and the time terms (oct_time_terms).  What is left per trace is the aperture
// test |m - mh| > apm and t = th + ah*(m - mh) (oct_traveltime_at), so traces that share
// a half offset can share both halves; the values are the same bits either way
// (oct_traveltime composes them for one trace).  The
// error tests are evaluated as masks, not branches: every lane computes the whole
// expression (a failed square root or division yields a NaN or inf that is then
// discarded).  Radicands of lanes already in error are replaced by 1 so that the square
// root's small-input path stays off.
struct OctMid {
    float mh;
    bool bad;  // negative first radicand or zero second one (the reference's error codes)
};

template <bool FAST>
__device__ __forceinline__ OctMid oct_mid(const OctTrace &tr, const OctCand &k, float m0) {
    // displaced midpoint (oct_displaced_midpoint)
    const float arg1 = k.tq_sq + k.tq_th * tr.gamma_sq + tr.p16 * k.theta_sq * k.theta_sq;
    const bool bad1 = arg1 < 0;
    const float arg2 =
        k.theta_sq * tr.gamma_sq + k.two_tq + 2.0f * sqrt_of<FAST>(bad1 ? 1.0f : arg1);
    OctMid d;
    d.bad = bad1 | (arg2 == 0);
    d.mh = m0 + k.two_theta * tr.hd / sqrt_of<FAST>(d.bad ? 1.0f : arg2);
    return d;
}

struct OctTime {
    float th, ah;  // t = th + ah*(m - mh)
    bool bad;
};

template <bool FAST>
__device__ __forceinline__ OctTime oct_time_terms(const OctTrace &tr, const OctCand &k, float m0,
                                                  const OctMid &d) {
    // time along the trajectory (oct_time)
    const float dmh = (d.mh - m0) * (d.mh - m0);
    const float w1 = tr.e1 - dmh;
    const float w2 = tr.e2 - dmh;
    const bool bad3 = (w1 < 0) | (w2 < 0);
    const float u = sqrt_of<FAST>(bad3 ? 1.0f : w1) + sqrt_of<FAST>(bad3 ? 1.0f : w2);
    const float chh = k.c * tr.h * tr.h;
    float th = k.t0;
    bool bad4 = bad3;
    if (tr.branch == 1) {
        bad4 = bad4 | (u == 0.0f);
        const float a = chh + tr.f4 / (u * u) * k.tn0_sq;
        bad4 = bad4 | (a < 0);
        th = sqrt_of<FAST>(bad4 ? 1.0f : a);
    } else if (tr.branch == 2) {
        const float a = chh + u * u / k.q4 * k.tn0_sq;
        bad4 = bad4 | (a < 0);
        th = sqrt_of<FAST>(bad4 ? 1.0f : a);
    }
    const float tn_sq = th * th - chh;
    OctTime r;
    r.bad = bad4 | (th == 0.0f) | (k.tn0_sq == 0.0f);
    r.th = th;
    r.ah = (k.t0 * tn_sq) / (th * k.tn0_sq) * k.slope;
    return r;
}

// The whole OCT traveltime of one trace: the time terms only where the trace passes the
// aperture test.  *reject is the aperture decision (notUsedCount); the return value says
// whether t is a valid traveltime.
template <bool FAST>
__device__ __forceinline__ bool oct_traveltime(const OctTrace &tr, const OctCand &k, float m0,
                                               float apm, float *t, bool *reject) {
    const OctMid d = oct_mid<FAST>(tr, k, m0);
    const bool rej = d.bad | (fabsf(tr.m - d.mh) > apm);
    *reject = rej;
    if (rej) return false;
    const OctTime tt = oct_time_terms<FAST>(tr, k, m0, d);
    *t = tt.th + tt.ah * (tr.m - d.mh);
    return !tt.bad;
}

// The per-trace remainder from terms computed earlier for the same half offset.
__device__ __forceinline__ bool oct_traveltime_at(const OctMid &d, const OctTime &tt, float m,
                                                  float apm, float *t, bool *reject) {
    const bool rej = d.bad | (fabsf(m - d.mh) > apm);
    *reject = rej;
    *t = tt.th + tt.ah * (m - d.mh);
    return !rej & !tt.bad;
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

__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// A wave-uniform float held in a scalar register: readfirstlane, then an empty asm
// statement with an "s" operand (readfirstlane alone is folded away when the compiler sees
// that the value is uniform, and the value goes back to a vector register).
__device__ __forceinline__ float uniform_f(float x) {
    int b = __builtin_amdgcn_readfirstlane(__float_as_int(x));
    __asm__("" : "+s"(b));
    return __int_as_float(b);
}

// threadIdx.x through an empty asm statement: what is derived from it at a point of use
// (a lane's LDS slots, its output row, its sample) is recomputed there from the one
// register that holds the thread index, instead of being kept in registers of its own (or
// spilled) across the trace loops.
__device__ __forceinline__ int thread_index_here() {
    int t = (int)threadIdx.x;
    __asm__ volatile("" : "+v"(t));
    return t;
}

}  // namespace hss
