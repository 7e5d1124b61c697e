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
//   HSS_KAHAN two-level compensation: the values of up to kKahanChunk consecutive
//             traces are summed as HSS_FAST sums them (A_j += u, den = fma(u, u, den)),
//             then each chunk's sums are Kahan-added into running totals (flush()), and
//             the final sums over j are compensated too.  A chunk is a sum of a few
//             terms, so its rounding stays at a few ulps whatever the number of traces,
//             and the compensation costs a few operations per chunk, not per trace.
#pragma once

#include "device_common.hpp"

namespace hss {

// 16-bit samples read per lane without LDS (compiled window sizes only, all odd): the
// (W+1)/2 + 1 aligned 32-bit words around samples e0 .. e0+W, widened to float (not
// scaled).  FP16: the words shifted by one sample when e0 is odd (v_alignbit), then
// converted.  BF16: one v_perm_b32 per sample moves its two bytes into the high half of a
// zero word (selector byte 0x0c gives 0x00), from either word of the pair, so the shift
// and the widening are one instruction; sample 2i (2i+1) of the window is bytes {1,0}
// ({3,2}) of the pair for even e0 and bytes {3,2} ({5,4}) for odd e0.  Words q[0 .. NW-2] cover samples
// a .. a+W (a = e0 rounded down to even), all inside the window or the sample just before
// it; the last word is needed only for odd e0, where it holds sample e0+W (and the sample
// after it, inside the same aligned word); for even e0 it re-reads q[NW-2].  The window's
// first sample e0 is sample rpar + k of the wave-uniform word row `roww` (the row's first
// sample rounded down to an even sample, i.e. a 4-byte word of the aligned sample array;
// rpar = 1 when it was odd), so the loads take a scalar base and a 32-bit lane offset.
template <int STORAGE, int W>
__device__ __forceinline__ void read16(const uint32_t *roww, unsigned rpar, unsigned k,
                                       float (&y)[W + 1]) {
    static_assert(W & 1, "read16 covers an even number of samples");
    constexpr int NW = (W + 1) / 2 + 1;
    const unsigned e0 = rpar + k;
    const unsigned p = e0 & 1u;
    const uint32_t *q = roww + (e0 >> 1);
    uint32_t D[NW];
#pragma unroll
    for (int i = 0; i < NW - 1; ++i) D[i] = q[i];
    D[NW - 1] = q[NW - 2 + p];
    if constexpr (STORAGE == HSS_F16) {
#pragma unroll
        for (int i = 0; i < NW - 1; ++i) {
            const uint32_t e = __builtin_amdgcn_alignbit(D[i + 1], D[i], 16u * p);
            y[2 * i] = (float)__builtin_bit_cast(_Float16, (uint16_t)(e & 0xffffu));
            y[2 * i + 1] = (float)__builtin_bit_cast(_Float16, (uint16_t)(e >> 16));
        }
    } else {
        const uint32_t sel_even = 0x01000c0cu | (p ? 0x02020000u : 0u);
        const uint32_t sel_odd = sel_even + 0x02020000u;
#pragma unroll
        for (int i = 0; i < NW - 1; ++i) {
            y[2 * i] = __uint_as_float(__builtin_amdgcn_perm(D[i + 1], D[i], sel_even));
            y[2 * i + 1] = __uint_as_float(__builtin_amdgcn_perm(D[i + 1], D[i], sel_odd));
        }
    }
}

// FP16 samples e0 .. e0+W of one window as halves (the words of read16 shifted by one
// sample for odd e0, not converted): arithmetic on (float)h[j] then reads each half
// straight from its word (v_fma_mix_f32) instead of converting it first.
template <int W>
__device__ __forceinline__ void read16_halves(const uint32_t *roww, unsigned rpar, unsigned k,
                                              _Float16 (&h)[W + 1]) {
    static_assert(W & 1, "read16 covers an even number of samples");
    constexpr int NW = (W + 1) / 2 + 1;
    const unsigned e0 = rpar + k;
    const unsigned p = e0 & 1u;
    const uint32_t *q = roww + (e0 >> 1);
    uint32_t D[NW];
#pragma unroll
    for (int i = 0; i < NW - 1; ++i) D[i] = q[i];
    D[NW - 1] = q[NW - 2 + p];
#pragma unroll
    for (int i = 0; i < NW - 1; ++i) {
        const uint32_t e = __builtin_amdgcn_alignbit(D[i + 1], D[i], 16u * p);
        h[2 * i] = __builtin_bit_cast(_Float16, (uint16_t)(e & 0xffffu));
        h[2 * i + 1] = __builtin_bit_cast(_Float16, (uint16_t)(e >> 16));
    }
}

// Kahan step: s + c' = s + c + x with the low part of the sum kept in c (c holds the
// negated error, as in Kahan's original form).
__device__ __forceinline__ void kahan_add(float &s, float &c, float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Traces per compensation chunk of HSS_KAHAN (a power of two that divides the 64 traces of
// a geometry chunk).  16 measured 4% faster than 8 on the C5 line (FP16 grid); a chunk's
// plain sum then rounds at most about 8 ulps, against the ~36 of a plain 72-trace sum.
constexpr int kKahanChunk = 16;
static_assert((kKahanChunk & (kKahanChunk - 1)) == 0 && kKahanChunk <= 64,
              "the chunk must be a power of two dividing 64");

template <int W, int VARIANT>
struct Window {
    static constexpr int N = W > 0 ? W : kMaxWindow;
    static constexpr bool KAHAN = VARIANT == HSS_KAHAN;
    static constexpr int NK = KAHAN ? N : 1;
    float A[N];            // window sums (HSS_KAHAN: of the current chunk)
    float T[NK], Tc[NK];   // HSS_KAHAN: running totals of A_j and their compensation
    float lin, den, tden, denc;
    unsigned used;

    __device__ __forceinline__ void reset() {
#pragma unroll
        for (int j = 0; j < N; ++j) A[j] = 0.0f;
#pragma unroll
        for (int j = 0; j < NK; ++j) {
            T[j] = 0.0f;
            Tc[j] = 0.0f;
        }
        lin = 0.0f;
        den = 0.0f;
        tden = 0.0f;
        denc = 0.0f;
        used = 0;
    }

    // HSS_KAHAN: adds the chunk sums to the compensated totals and starts a new chunk
    // (a no-op for the other variants).  Callers flush every kKahanChunk traces, at
    // wave-uniform points of the trace loop.
    __device__ __forceinline__ void flush(int w) {
        if constexpr (KAHAN) {
#pragma unroll
            for (int j = 0; j < N; ++j) {
                if (W == 0 && j >= w) continue;
                kahan_add(T[j], Tc[j], A[j]);
                A[j] = 0.0f;
            }
            kahan_add(tden, denc, den);
            den = 0.0f;
        }
    }

    // One interpolated sample u = y + frac*d at window position j, d = y_next - y (the
    // same FP32 difference the reference forms, so HSS_EXACT keeps its bits).
    __device__ __forceinline__ void step_d(int j, float y, float d, float f) {
        if constexpr (VARIANT == HSS_EXACT) {
            float u = d * f + y;
            A[j] += u;
            lin += u;
            den += u * u;
        } else {
            // HSS_FAST, and HSS_KAHAN inside a chunk
            float u = __builtin_fmaf(d, f, y);
            A[j] += u;
            den = __builtin_fmaf(u, u, den);
        }
    }

    __device__ __forceinline__ void step(int j, float y0, float y1, float f) {
        step_d(j, y0, y1 - y0, f);
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
            step(j, y0, y1, f);
        }
        used++;
    }

    // FP32 samples of one trace: row[k .. k+w] with `row` wave-uniform (the trace's first
    // sample minus tau) and k the lane's sample index, so the loads take a scalar base and
    // a 32-bit lane offset.
    __device__ __forceinline__ void add_row(const float *row, unsigned k, float f, int w) {
        const float *p = row + k;
        float y1 = p[0];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (W == 0 && j >= w) continue;
            const float y0 = y1;
            y1 = p[j + 1];
            step(j, y0, y1, f);
        }
        used++;
    }

    // 16-bit samples of one window read per lane (read16), then summed.  FACTORED: the
    // non-exact variants factor the trace's scale `sc` out of the window (a line without
    // scales passes sc = 1).  !FACTORED, for a line stored without scales (BF16: FP32's
    // exponent range needs none): the FP32 steps, whose sums need an add per value where
    // the factored scale needs an FMA.  The choice is a template parameter so that a
    // kernel holds one form in its trace loop (both would add registers).
    template <int STORAGE, bool FACTORED = true>
    __device__ __forceinline__ void add16(const uint32_t *roww, unsigned rpar, unsigned k,
                                          float f, float sc) {
        static_assert(W > 0, "add16 needs a compiled window size");
        if constexpr (VARIANT != HSS_EXACT && STORAGE == HSS_F16 && FACTORED) {
            // as below, with y1 - y0 formed as fma(y1, one, -y0) = RN(y1 - y0): the same
            // bits, but an FMA reads both halves straight from the words (v_fma_mix_f32),
            // where a subtraction needs both converted first.  `one` is 1.0 hidden from the
            // optimizer (an empty asm statement), which would otherwise turn the FMA back
            // into a subtraction.
            _Float16 h[W + 1];
            read16_halves<W>(roww, rpar, k, h);
            float one = 1.0f;
            __asm__("" : "+s"(one));
            float dn = 0.0f;
#pragma unroll
            for (int j = 0; j < W; ++j) {
                const float d = __builtin_fmaf((float)h[j + 1], one, -(float)h[j]);
                const float u = __builtin_fmaf(d, f, (float)h[j]);
                A[j] = __builtin_fmaf(sc, u, A[j]);
                dn = __builtin_fmaf(u, u, dn);
            }
            den = __builtin_fmaf(sc * sc, dn, den);
            used++;
            return;
        }
        float y[W + 1];
        read16<STORAGE, W>(roww, rpar, k, y);
        if constexpr (VARIANT != HSS_EXACT) {
            if constexpr (!FACTORED) {
#pragma unroll
                for (int j = 0; j < W; ++j) step(j, y[j], y[j + 1], f);
                used++;
                return;
            }
            // the trace's scale factored out of the window: u = sc * (y0 + frac*(y1 - y0))
            // on the stored values, A_j += sc*u and den += sc^2 * sum_j u^2
            float dn = 0.0f;
#pragma unroll
            for (int j = 0; j < W; ++j) {
                const float u = __builtin_fmaf(y[j + 1] - y[j], f, y[j]);
                A[j] = __builtin_fmaf(sc, u, A[j]);
                dn = __builtin_fmaf(u, u, dn);
            }
            den = __builtin_fmaf(sc * sc, dn, den);
        } else {
            if constexpr (FACTORED) {
#pragma unroll
                for (int j = 0; j <= W; ++j) y[j] = y[j] * sc;
            }
#pragma unroll
            for (int j = 0; j < W; ++j) step(j, y[j], y[j + 1], f);
        }
        used++;
    }

    // Same as add() for a window already decoded (and scaled) into LDS: row[0..w].
    __device__ __forceinline__ void add_lds(const LdsFloat *row, float f, int w) {
        float y1 = row[0];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (W == 0 && j >= w) continue;
            const float y0 = y1;
            y1 = row[j + 1];
            step(j, y0, y1, f);
        }
        used++;
    }

    // REDUCE_SEMBLANCE_STACK: S = sum_j A_j^2 / (used * den), stack = lin / (used * w);
    // both stay 0 when no trace was used.  (HSS_KAHAN: the last chunk is flushed first.)
    __device__ __forceinline__ void reduce(int w, float *S, float *stack) {
        flush(w);
        float sem = 0.0f, stk = 0.0f;
        if (used > 0) {
            float num = 0.0f, numc = 0.0f;
            float l = lin, lc = 0.0f;
            if constexpr (VARIANT != HSS_EXACT) l = 0.0f;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                if (W == 0 && j >= w) continue;
                if constexpr (VARIANT == HSS_EXACT) {
                    num += A[j] * A[j];
                } else if constexpr (VARIANT == HSS_FAST) {
                    num = __builtin_fmaf(A[j], A[j], num);
                    l += A[j];
                } else {
                    // the total carries its compensation: use the corrected value; the
                    // square's rounding goes into the compensation through the FMA
                    const float a = T[j] - Tc[j];
                    const float y = __builtin_fmaf(a, a, -numc);
                    const float t = num + y;
                    numc = (t - num) - y;
                    num = t;
                    kahan_add(l, lc, a);
                }
            }
            float d = den;
            if constexpr (KAHAN) {
                num = num - numc;
                l = l - lc;
                d = tden - denc;
            }
            sem = num / ((float)used * d);
            stk = l / (float)(used * (unsigned)w);
        }
        *S = sem;
        *stack = stk;
    }
};

}  // namespace hss
