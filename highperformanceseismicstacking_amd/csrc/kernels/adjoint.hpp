// Adjoint of the semblance with respect to the traveltime parameters, and the
// Gauss-Newton polish it drives.  No counterpart in the reference (which only searches).
//
// For one (t0, parameter point) lane, with the window sums of REDUCE_SEMBLANCE_STACK
// (common/include/gpu/interface.h:46-55) S = sum_j A_j^2 / (used * den):
//   u_ij = y_ij + f_i d_ij,  d_ij = y_i,k+j+1 - y_i,k+j,  f_i = t_i/dt - k_i,
//   dS/dtheta = 2/(used*den) * ( sum_j A_j G_j - S * used * H ),
//   G_j = sum_i d_ij g_i,  H = sum_i (sum_j u_ij d_ij) g_i,  g_i = (dt_i/dtheta) / dt,
// one pass over the traces (the index k_i and the trace set are locally constant, so
// the gradient is that of the piecewise-smooth S away from the seams where a window
// slides by one sample).  theta is the storage form (c = 4/v^2, A, B or slope); the
// entry points convert dS/dc to dS/dv.  The traveltime derivatives are analytic for CMP
// and ZO-CRS and come from forward-mode dual numbers for OCT.
#pragma once

#include "eval.hpp"
#include "launch_util.hpp"

namespace hss {

template <int MODEL>
constexpr int model_np() {
    return MODEL == HSS_CMP ? 1 : (MODEL == HSS_ZOCRS ? 3 : 2);
}

// ---------------------------------------------------------------------------
// Dual numbers with D derivative slots (OCT traveltime derivatives).
template <int D>
struct Dual {
    float v;
    float d[D];
};

template <int D>
__device__ __forceinline__ Dual<D> dconst(float v) {
    Dual<D> r;
    r.v = v;
#pragma unroll
    for (int m = 0; m < D; ++m) r.d[m] = 0.0f;
    return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator+(Dual<D> a, Dual<D> b) {
    Dual<D> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int m = 0; m < D; ++m) r.d[m] = a.d[m] + b.d[m];
    return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator-(Dual<D> a, Dual<D> b) {
    Dual<D> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int m = 0; m < D; ++m) r.d[m] = a.d[m] - b.d[m];
    return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator*(Dual<D> a, Dual<D> b) {
    Dual<D> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int m = 0; m < D; ++m) r.d[m] = __builtin_fmaf(a.d[m], b.v, a.v * b.d[m]);
    return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator*(Dual<D> a, float s) {
    Dual<D> r;
    r.v = a.v * s;
#pragma unroll
    for (int m = 0; m < D; ++m) r.d[m] = a.d[m] * s;
    return r;
}
template <int D>
__device__ __forceinline__ Dual<D> operator/(Dual<D> a, Dual<D> b) {
    Dual<D> r;
    r.v = a.v / b.v;
    const float rb = 1.0f / b.v;
#pragma unroll
    for (int m = 0; m < D; ++m) r.d[m] = (a.d[m] - r.v * b.d[m]) * rb;
    return r;
}
template <int D>
__device__ __forceinline__ Dual<D> dsqrt(Dual<D> a) {
    Dual<D> r;
    r.v = sqrtf(a.v);
    const float h = r.v > 0.0f ? 0.5f / r.v : 0.0f;
#pragma unroll
    for (int m = 0; m < D; ++m) r.d[m] = a.d[m] * h;
    return r;
}

// OCT displaced midpoint and traveltime (oct/common.cuh:7-94) with c, slope as duals.
__device__ __forceinline__ int oct_mh_dual(float h, float h0, float t0, float m0, Dual<2> c,
                                           Dual<2> slope, Dual<2> *mh) {
    const Dual<2> theta = slope * t0;
    const float gamma = 2.0f * sqrtf(h * h + h0 * h0);
    const Dual<2> tn0_sq = dconst<2>(t0 * t0) - c * (h0 * h0);
    const Dual<2> theta_sq = theta * theta;
    const float gamma_sq = gamma * gamma;
    const Dual<2> tn0_quad = tn0_sq * tn0_sq;
    Dual<2> arg = tn0_quad * tn0_quad + (tn0_quad * theta_sq) * gamma_sq +
                  (theta_sq * theta_sq) * (16.0f * h * h * h0 * h0);
    if (arg.v < 0.0f) return kNegativeSquaredRoot;
    arg = theta_sq * gamma_sq + tn0_quad * 2.0f + dsqrt(arg) * 2.0f;
    if (arg.v == 0.0f) return kDivisionByZero;
    *mh = dconst<2>(m0) + (theta * (2.0f * (h * h - h0 * h0))) / dsqrt(arg);
    return kNoError;
}

// OCT traveltime from the displaced midpoint (oct/common.cuh:7-94) as duals, in its two
// parts: th and the coefficient ah of (m - mh), which depend on the trace's half offset
// only; the traveltime at midpoint m is th + ah * (m - mh) (oct_time_from_terms).
__device__ __forceinline__ int oct_time_terms_dual(float h, float h0, float t0, float m0,
                                                   Dual<2> mh, Dual<2> c, Dual<2> slope,
                                                   Dual<2> *th_out, Dual<2> *ah_out) {
    const Dual<2> tn0_sq = dconst<2>(t0 * t0) - c * (h0 * h0);
    const Dual<2> dmh = mh - dconst<2>(m0);
    const Dual<2> dmh2 = dmh * dmh;
    const Dual<2> w1 = dconst<2>((h + h0) * (h + h0)) - dmh2;
    const Dual<2> w2 = dconst<2>((h - h0) * (h - h0)) - dmh2;
    if (w1.v < 0.0f || w2.v < 0.0f) return kNegativeSquaredRoot;
    const Dual<2> u = dsqrt(w1) + dsqrt(w2);
    Dual<2> th = dconst<2>(t0);
    if (fabsf(h) > fabsf(h0)) {
        if (u.v == 0.0f) return kDivisionByZero;
        const Dual<2> arg = c * (h * h) + (dconst<2>(4.0f * h * h) / (u * u)) * tn0_sq;
        if (arg.v < 0.0f) return kNegativeSquaredRoot;
        th = dsqrt(arg);
    } else if (fabsf(h) < fabsf(h0)) {
        if (h0 == 0.0f) return kDivisionByZero;
        const Dual<2> arg = c * (h * h) + ((u * u) * (1.0f / (4.0f * h0 * h0))) * tn0_sq;
        if (arg.v < 0.0f) return kNegativeSquaredRoot;
        th = dsqrt(arg);
    }
    const Dual<2> tn_sq = th * th - c * (h * h);
    if (th.v == 0.0f || tn0_sq.v == 0.0f) return kDivisionByZero;
    *th_out = th;
    *ah_out = ((tn_sq * t0) / (th * tn0_sq)) * slope;
    return kNoError;
}

// The dual terms of one half offset for one parameter point: displaced midpoint, th and
// ah (all independent of the trace's midpoint), or the error that stopped them.
struct OctDualTerms {
    Dual<2> mh, th, ah;
    int err;
};

__device__ __forceinline__ OctDualTerms oct_dual_terms(float h, float h0, float t0, float m0,
                                                       Dual<2> c, Dual<2> slope) {
    OctDualTerms r;
    r.err = oct_mh_dual(h, h0, t0, m0, c, slope, &r.mh);
    if (r.err == kNoError) r.err = oct_time_terms_dual(h, h0, t0, m0, r.mh, c, slope, &r.th, &r.ah);
    return r;
}

// The traveltime at midpoint m from the offset's terms.
__device__ __forceinline__ Dual<2> oct_time_from_terms(const OctDualTerms &q, float m) {
    return q.th + q.ah * (dconst<2>(m) - q.mh);
}

// The point's (c, slope) as duals seeded for d/dc and d/dslope.
__device__ __forceinline__ void oct_seed_duals(const Candidate &k, Dual<2> *c, Dual<2> *sl) {
    *c = dconst<2>(k.c);
    *sl = dconst<2>(k.p1);
    c->d[0] = 1.0f;
    sl->d[1] = 1.0f;
}

// Traveltime and dt/dtheta (theta in storage form).  The value and the acceptance are
// traveltime()'s, bit for bit, so a lane keeps the same windows as the search kernels.
// OCT: with use_kept, `kept` holds the dual terms of the trace's half offset for this
// point (the same bits oct_dual_terms gives for it).
template <int MODEL, bool FAST>
__device__ __forceinline__ bool traveltime_grad(const TraceGeom &tg, float t0, float t0sq, float m0,
                                                float h0, float apm, const Candidate &k, float *t,
                                                float *dtd, bool use_kept, const OctDualTerms &kept) {
    if constexpr (MODEL == HSS_CMP) {
        bool r;
        traveltime<MODEL, FAST>(tg, t0, t0sq, m0, h0, apm, k, t, &r);
        dtd[0] = *t > 0.0f ? 0.5f * tg.h2 / *t : 0.0f;
        return true;
    } else if constexpr (MODEL == HSS_ZOCRS) {
        const float dm = tg.m - m0;
        const float a = t0 + k.p1 * dm;
        bool rj;
        if (!traveltime<MODEL, FAST>(tg, t0, t0sq, m0, h0, apm, k, t, &rj)) return false;
        const float r = *t > 0.0f ? 1.0f / *t : 0.0f;
        dtd[0] = 0.5f * tg.h2 * r;
        dtd[1] = a * dm * r;
        dtd[2] = 0.5f * dm * dm * r;
        return true;
    } else {
        // value and decisions from the forward traveltime (the same bits as the search
        // kernels), derivatives from the dual-number twin of the same formulas
        bool r;
        if (!traveltime<MODEL, FAST>(tg, t0, t0sq, m0, h0, apm, k, t, &r)) return false;
        OctDualTerms q;
        if (use_kept) {
            q = kept;
        } else {
            Dual<2> c, sl;
            oct_seed_duals(k, &c, &sl);
            q = oct_dual_terms(tg.h, h0, t0, m0, c, sl);
        }
        if (q.err == kNoError) {
            const Dual<2> tt = oct_time_from_terms(q, tg.m);
            dtd[0] = tt.d[0];
            dtd[1] = tt.d[1];
        } else {
            dtd[0] = dtd[1] = 0.0f;
        }
        return true;
    }
}

// S, stack and dS/dtheta for one lane's own parameter point.  FMA arithmetic (as
// HSS_FAST); per-lane windows are read with per-lane loads.  Every lane of the wave must
// call it (the trace geometry is broadcast across the wave).
template <int MODEL, int W, int STORAGE>
__device__ __forceinline__ void evaluate_grad(const EvalCtx &e, float t0, float t0sq,
                                              const Candidate &k, bool active, float *S,
                                              float *stack, float *grad) {
    constexpr int D = model_np<MODEL>();
    constexpr int N = W > 0 ? W : kMaxWindow;
    constexpr bool FASTIDX = W != 0;
    using Reader = SampleReader<STORAGE>;
    const bool scaled = (STORAGE != HSS_F32) && e.scale != nullptr;
    const Reader rd{(decltype(Reader::p))e.samples};
    float A[N], G[N][D], H[D];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        A[j] = 0.0f;
#pragma unroll
        for (int m = 0; m < D; ++m) G[j][m] = 0.0f;
    }
#pragma unroll
    for (int m = 0; m < D; ++m) H[m] = 0.0f;
    float den = 0.0f;
    unsigned used = 0;
    const int lane = threadIdx.x & (kWave - 1);
    // OCT: the point's dual terms of the current half offset, kept while consecutive
    // traces share the offset's bit pattern (a wave-uniform test) when the list comes in
    // runs of one offset (grouped by offset); in the reference's order each trace computes
    // its own, only when the aperture keeps it
    bool keep = false;
    if constexpr (MODEL == HSS_OCT) keep = oct_offset_runs_long(e);
    OctDualTerms kept{};
    uint32_t kept_key = 0;
    bool kept_valid = false;
    for (int j0 = 0; j0 < e.cnt; j0 += kWave) {
        const int mine = j0 + lane;
        const bool have = mine < e.cnt;
        const int ti_l = have ? (e.midx ? e.midx[e.beg + mine] : e.beg + mine) : 0;
        const TraceGeom g_l = have ? load_geom<MODEL>(e.mid, e.hoff, e.hoff2, ti_l)
                                   : TraceGeom{0.0f, 0.0f, 0.0f};
        const float sc_l = (scaled && have) ? e.scale[ti_l] : 1.0f;
        const int nchunk = min(kWave, e.cnt - j0);
        for (int jj = 0; jj < nchunk; ++jj) {
            const int ti = __builtin_amdgcn_readlane(ti_l, jj);
            TraceGeom tg;
            tg.m = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g_l.m), jj));
            tg.h = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g_l.h), jj));
            tg.h2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(g_l.h2), jj));
            const float sc =
                scaled ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sc_l), jj)) : 1.0f;
            float t, dtd[D > 0 ? D : 1];
            if constexpr (MODEL == HSS_OCT) {
                if (keep && (!kept_valid || __float_as_uint(tg.h) != kept_key)) {
                    Dual<2> c, sl;
                    oct_seed_duals(k, &c, &sl);
                    kept = oct_dual_terms(tg.h, e.h0, t0, e.m0, c, sl);
                    kept_key = __float_as_uint(tg.h);
                    kept_valid = true;
                }
            }
            if (!traveltime_grad<MODEL, FASTIDX>(tg, t0, t0sq, e.m0, e.h0, e.apm, k, &t, dtd, keep,
                                                 kept))
                continue;
            int first;
            float f;
            if (!window_index<FASTIDX>(t, e.dt, e.inv_dt, e.ib, e.tau, &first, &f)) continue;
            if (!active) continue;
            float gi[D];
#pragma unroll
            for (int m = 0; m < D; ++m) gi[m] = dtd[m] * e.inv_dt;
            // the window's samples y[0..w]: compiled window sizes read them from the trace's
            // row (a wave-uniform base and a 32-bit lane offset; 16-bit storage as aligned
            // words, read16), the runtime-w instantiation sample by sample
            const size_t trace0 = (size_t)ti * (size_t)e.ns;
            float y[N + 1];
            if constexpr (W > 0 && STORAGE == HSS_F32) {
                const float *p = (const float *)e.samples + trace0 + (unsigned)first;
#pragma unroll
                for (int j = 0; j <= N; ++j) y[j] = p[j];
            } else if constexpr (W > 0) {
                const size_t row = trace0 - (size_t)e.tau;
                const unsigned rpar = (unsigned)row & 1u;
                const uint32_t *roww =
                    (const uint32_t *)((const uint16_t *)e.samples + (row - rpar));
                read16<STORAGE, W>(roww, rpar, (unsigned)(first + e.tau), y);
#pragma unroll
                for (int j = 0; j <= N; ++j) y[j] = y[j] * sc;
            } else {
#pragma unroll
                for (int j = 0; j <= N; ++j)
                    y[j] = (j <= e.w) ? rd(trace0 + (size_t)first + j) * sc : 0.0f;
            }
            float eu = 0.0f;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                if (W == 0 && j >= e.w) continue;
                const float y0 = y[j];
                const float y1 = y[j + 1];
                const float d = y1 - y0;
                const float u = __builtin_fmaf(d, f, y0);
                A[j] += u;
                den = __builtin_fmaf(u, u, den);
                eu = __builtin_fmaf(u, d, eu);
#pragma unroll
                for (int m = 0; m < D; ++m) G[j][m] = __builtin_fmaf(d, gi[m], G[j][m]);
            }
#pragma unroll
            for (int m = 0; m < D; ++m) H[m] = __builtin_fmaf(eu, gi[m], H[m]);
            used++;
        }
    }
    float sem = 0.0f, stk = 0.0f;
    float gr[D];
#pragma unroll
    for (int m = 0; m < D; ++m) gr[m] = 0.0f;
    if (used > 0 && den > 0.0f) {
        float num = 0.0f, lin = 0.0f;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (W == 0 && j >= e.w) continue;
            num = __builtin_fmaf(A[j], A[j], num);
            lin += A[j];
        }
        const float uden = (float)used * den;
        sem = num / uden;
        stk = lin / (float)(used * (unsigned)e.w);
        const float scl = 2.0f / uden;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            float a = 0.0f;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                if (W == 0 && j >= e.w) continue;
                a = __builtin_fmaf(A[j], G[j][m], a);
            }
            gr[m] = scl * (a - sem * (float)used * H[m]);
        }
    }
    *S = sem;
    *stack = stk;
#pragma unroll
    for (int m = 0; m < D; ++m) grad[m] = gr[m];
}

// ---------------------------------------------------------------------------
// Kernels: one wave per (midpoint, 64 t0 samples); lane = t0 sample with its own point.
struct AdjointArgs {
    EvalCtx e;
    const int *mbeg;
    const int *mcnt;
    const float *mm0;
    const int *morder;  // launch order (launch_midpoint) or NULL
    int nmid;
    int ntiles;
    const float *points;  // [n][(2+np)][ns]; rows 2.. are the parameters (v space)
    float *out;           // [n][(2+np)][ns]
    float lo[3], hi[3];   // polish bounds (v space)
    int steps;            // polish steps
};

constexpr int kAdjThreads = kWave;

template <int MODEL>
__device__ __forceinline__ Candidate candidate_from_v(const float *p) {
    Candidate k{0.0f, 0.0f, 0.0f};
    k.c = 4.0f / (p[0] * p[0]);
    if constexpr (model_np<MODEL>() >= 2) k.p1 = p[1];
    if constexpr (model_np<MODEL>() >= 3) k.p2 = p[2];
    return k;
}

// gradient in v space: dS/dv = dS/dc * dc/dv, c = 4/v^2
template <int MODEL>
__device__ __forceinline__ void grad_to_v(const float *p, float *g) {
    g[0] = g[0] * (-8.0f / (p[0] * p[0] * p[0]));
}

template <int MODEL, int W, int STORAGE>
__global__ __launch_bounds__(kAdjThreads) void gradient_kernel(AdjointArgs a) {
    constexpr int D = model_np<MODEL>();
    const int lane = threadIdx.x;
    const int blk = logical_block(blockIdx.x, gridDim.x);
    const int slot = blk / a.ntiles;
    const int tile = blk - slot * a.ntiles;
    const int g = launch_midpoint(a.morder, a.nmid, slot);
    if (g < 0) return;
    const int s = tile * kWave + lane;
    const int ns = a.e.ns;
    const bool active = s < ns;
    const int sc = active ? s : 0;
    const EvalCtx e = midpoint_ctx(a.e, a.mbeg, a.mcnt, a.mm0, g);
    const float *row = a.points + (size_t)g * (2 + D) * ns;
    float p[D];
#pragma unroll
    for (int m = 0; m < D; ++m) p[m] = row[(2 + m) * ns + sc];
    const bool ok = active && p[0] > 0.0f;
    if (!(p[0] > 0.0f)) p[0] = 1.0f;
    const float t0 = (float)sc * e.dt;
    float S, stk, gr[D];
    evaluate_grad<MODEL, W, STORAGE>(e, t0, t0 * t0, candidate_from_v<MODEL>(p), ok, &S, &stk, gr);
    grad_to_v<MODEL>(p, gr);
    if (!active) return;
    float *out = a.out + (size_t)g * (2 + D) * ns;
    out[s] = ok ? S : 0.0f;
    out[ns + s] = ok ? stk : 0.0f;
#pragma unroll
    for (int m = 0; m < D; ++m) out[(2 + m) * ns + s] = ok ? gr[m] : 0.0f;
}

// Gauss-Newton polish of each lane's point: the residual r = 1 - S (S <= 1) has the
// Jacobian -dS/dx in the box-normalised coordinates x = (p - lo)/(hi - lo); a
// Levenberg-Marquardt damped Gauss-Newton step for one residual is
//   dx = r * dS/dx / (|dS/dx|^2 + mu),
// kept when S increases (mu /= 3), rejected otherwise (mu *= 4); mu starts at |dS/dx|^2.
// S never decreases; the point stays in the box.
template <int MODEL, int W, int STORAGE>
__global__ __launch_bounds__(kAdjThreads) void polish_kernel(AdjointArgs a) {
    constexpr int D = model_np<MODEL>();
    const int lane = threadIdx.x;
    const int blk = logical_block(blockIdx.x, gridDim.x);
    const int slot = blk / a.ntiles;
    const int tile = blk - slot * a.ntiles;
    const int g = launch_midpoint(a.morder, a.nmid, slot);
    if (g < 0) return;
    const int s = tile * kWave + lane;
    const int ns = a.e.ns;
    const bool active = s < ns;
    const int sc = active ? s : 0;
    const EvalCtx e = midpoint_ctx(a.e, a.mbeg, a.mcnt, a.mm0, g);
    const float *row = a.points + (size_t)g * (2 + D) * ns;
    float p[D], span[D];
#pragma unroll
    for (int m = 0; m < D; ++m) {
        span[m] = a.hi[m] - a.lo[m];
        p[m] = fminf(fmaxf(row[(2 + m) * ns + sc], a.lo[m]), a.hi[m]);
    }
    const float t0 = (float)sc * e.dt;
    const float t0sq = t0 * t0;
    float S, stk, gr[D];
    evaluate_grad<MODEL, W, STORAGE>(e, t0, t0sq, candidate_from_v<MODEL>(p), active, &S, &stk, gr);
    grad_to_v<MODEL>(p, gr);
    float mu = -1.0f;
    for (int it = 0; it < a.steps; ++it) {
        float gx[D], g2 = 0.0f;
#pragma unroll
        for (int m = 0; m < D; ++m) {
            gx[m] = gr[m] * span[m];
            g2 = __builtin_fmaf(gx[m], gx[m], g2);
        }
        const bool live = active && g2 > 0.0f && S > 0.0f && S < 1.0f;
        if (mu < 0.0f) mu = g2;
        const float r = 1.0f - S;
        float q[D];
#pragma unroll
        for (int m = 0; m < D; ++m) {
            const float dx = live ? r * gx[m] / (g2 + mu) : 0.0f;
            q[m] = fminf(fmaxf(p[m] + dx * span[m], a.lo[m]), a.hi[m]);
        }
        float S2, stk2, gr2[D];
        evaluate_grad<MODEL, W, STORAGE>(e, t0, t0sq, candidate_from_v<MODEL>(q), live, &S2, &stk2,
                                         gr2);
        grad_to_v<MODEL>(q, gr2);
        if (live && S2 > S) {
            S = S2;
            stk = stk2;
#pragma unroll
            for (int m = 0; m < D; ++m) {
                p[m] = q[m];
                gr[m] = gr2[m];
            }
            mu *= (1.0f / 3.0f);
        } else {
            mu *= 4.0f;
        }
    }
    if (!active) return;
    float *out = a.out + (size_t)g * (2 + D) * ns;
    out[s] = S;
    out[ns + s] = stk;
#pragma unroll
    for (int m = 0; m < D; ++m) out[(2 + m) * ns + s] = p[m];
}

template <int MODEL>
int launch_adjoint(const AdjointArgs &a, int n_mid, int storage, bool polish, hipStream_t stream) {
    const unsigned blocks = (unsigned)n_mid * (unsigned)a.ntiles;
    if (blocks == 0) return HSS_OK;
    int rc = dispatch_storage(storage, [&](auto S) {
        dispatch_window(a.e.w, fast_index_verified(a.e.dt, a.e.tau, stream), [&](auto Wc) {
            constexpr int WW = decltype(Wc)::value, SS = decltype(S)::value;
            if (polish)
                hipLaunchKernelGGL((polish_kernel<MODEL, WW, SS>), dim3(blocks), dim3(kAdjThreads), 0,
                                   stream, a);
            else
                hipLaunchKernelGGL((gradient_kernel<MODEL, WW, SS>), dim3(blocks), dim3(kAdjThreads),
                                   0, stream, a);
        });
    });
    if (rc != HSS_OK) return rc;
    return hip_status(hipGetLastError());
}

int launch_adjoint_cmp(const AdjointArgs &a, int n_mid, int storage, bool polish, hipStream_t st);
int launch_adjoint_zocrs(const AdjointArgs &a, int n_mid, int storage, bool polish, hipStream_t st);
int launch_adjoint_oct(const AdjointArgs &a, int n_mid, int storage, bool polish, hipStream_t st);

}  // namespace hss
