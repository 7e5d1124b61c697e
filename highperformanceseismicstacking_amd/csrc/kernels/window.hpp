// Window accumulation of the semblance numerator/denominator and the final reduction.
//
// Restates COMPUTE_SEMBLANCE and REDUCE_SEMBLANCE_STACK
// (common/include/gpu/interface.h:26-55) for one lane = one (t0, candidate) pair.
//   W > 0   : the window size is a compile-time constant (fully unrolled, one basic
//             block, the numerator components A_j live in VGPRs);
//   W == 0  : any w <= 32, guarded by the (wave-uniform) runtime w.
//   HSS_EXACT keeps the reference order: u = (y1 - y0)*frac + y0; A_j += u; lin += u;
//             den += u*u -- no contraction, so the sums equal an IEEE FP32 evaluation
//             of the reference text bit for bit.
//   HSS_FAST  u = fma(y1 - y0, frac, y0); A_j += u; den = fma(u, u, den); the linear sum
//             is recovered at the end as sum_j A_j (same value up to rounding).
#pragma once

#include "device_common.hpp"

namespace hss {

template <int W, int VARIANT>
struct Window {
    static constexpr int N = W > 0 ? W : kMaxWindow;
    float A[N];
    float lin, den;
    unsigned used;

    __device__ __forceinline__ void reset() {
#pragma unroll
        for (int j = 0; j < N; ++j) A[j] = 0.0f;
        lin = 0.0f;
        den = 0.0f;
        used = 0;
    }

    // Adds the w interpolated values starting at sample `base` of the flat sample array.
    template <class Reader>
    __device__ __forceinline__ void add(const Reader &rd, size_t base, float f, float scale,
                                        bool scaled, int w) {
        float y1 = rd(base);
        if (scaled) y1 = y1 * scale;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (W == 0 && j >= w) continue;
            float y0 = y1;
            y1 = rd(base + j + 1);
            if (scaled) y1 = y1 * scale;
            if constexpr (VARIANT == HSS_EXACT) {
                float u = (y1 - y0) * f + y0;
                A[j] += u;
                lin += u;
                den += u * u;
            } else {
                float u = __builtin_fmaf(y1 - y0, f, y0);
                A[j] += u;
                den = __builtin_fmaf(u, u, den);
            }
        }
        used++;
    }

    // Same as add() for a window already decoded (and scaled) into LDS: row[0..w].
    __device__ __forceinline__ void add_lds(const LdsFloat *row, float f, int w) {
        float y1 = row[0];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (W == 0 && j >= w) continue;
            float y0 = y1;
            y1 = row[j + 1];
            if constexpr (VARIANT == HSS_EXACT) {
                float u = (y1 - y0) * f + y0;
                A[j] += u;
                lin += u;
                den += u * u;
            } else {
                float u = __builtin_fmaf(y1 - y0, f, y0);
                A[j] += u;
                den = __builtin_fmaf(u, u, den);
            }
        }
        used++;
    }

    // REDUCE_SEMBLANCE_STACK: S = sum_j A_j^2 / (used * den), stack = lin / (used * w);
    // both stay 0 when no trace was used.
    __device__ __forceinline__ void reduce(int w, float *S, float *stack) const {
        float sem = 0.0f, stk = 0.0f;
        if (used > 0) {
            float num = 0.0f;
            float l = lin;
            if constexpr (VARIANT == HSS_FAST) l = 0.0f;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                if (W == 0 && j >= w) continue;
                if constexpr (VARIANT == HSS_EXACT) {
                    num += A[j] * A[j];
                } else {
                    num = __builtin_fmaf(A[j], A[j], num);
                    l += A[j];
                }
            }
            sem = num / ((float)used * den);
            stk = l / (float)(used * (unsigned)w);
        }
        *S = sem;
        *stack = stk;
    }
};

}  // namespace hss
