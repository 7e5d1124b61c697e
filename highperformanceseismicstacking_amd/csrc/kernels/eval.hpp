// One (t0, candidate) semblance evaluation by one lane, shared by the grid, differential
// evolution and stretch-free kernels.
//
// The lane walks the midpoint's traces in order -- the order of the reference's
// per-thread loop (cmp/linear_search.cu:55-76 and twins) -- and computes the traveltime,
// t/dt, the sample index and the interpolation weight with the reference's IEEE
// operations in both variants: one ulp of t/dt is up to 6e-5 of interpolation weight
// at t/dt ~ 1000, so the index chain must round as the reference's does for the results
// to agree within 1e-5.  HSS_EXACT then reproduces an IEEE FP32 evaluation of the
// reference text bit for bit; HSS_FAST contracts the window sums (window.hpp).
#pragma once

#include "device_common.hpp"
#include "window.hpp"

namespace hss {

struct EvalCtx {
    const void *samples;
    const float *scale;
    const float *mid;
    const float *hoff;
    const float *hoff2;
    const int *midx;  // NULL: traces beg .. beg+cnt-1
    int beg, cnt;
    int ns, tau, w;
    float dt, inv_dt;  // inv_dt = RN(1/dt)
    IndexBounds ib;    // index_bounds(ns, tau)
    float m0, h0, apm;
};

template <int MODEL>
__device__ __forceinline__ TraceGeom load_geom(const float *__restrict__ mid,
                                               const float *__restrict__ hoff,
                                               const float *__restrict__ hoff2, int ti) {
    TraceGeom g{0.0f, 0.0f, 0.0f};
    if constexpr (MODEL == HSS_CMP) {
        g.h2 = hoff2[ti];
    } else if constexpr (MODEL == HSS_ZOCRS) {
        g.m = mid[ti];
        g.h2 = hoff2[ti];
    } else {
        g.m = mid[ti];
        g.h = hoff[ti];
    }
    return g;
}

// SHIFT: stretch-free evaluation, t = t_n + n*dt (cmp/stretch_free.cu:54-56); `shift` is
// that n*dt.  Returns S and stack through the pointers and adds the number of model
// rejections (OCT notUsedCount) to *rejected.
// The compiled window sizes (W > 0) compute sqrt and t/dt with the verified short
// sequences (sqrt_rn, div_rn); the runtime-w instantiation (W == 0) uses the library
// operations and is also what the launchers pick when the check failed for this dt.
template <int MODEL, int W, int VARIANT, int STORAGE, bool SHIFT>
__device__ __forceinline__ void evaluate(const EvalCtx &e, float t0, float t0sq,
                                         const Candidate &k, float shift, bool active,
                                         float *S, float *stack, unsigned *rejected) {
    using Reader = SampleReader<STORAGE>;
    const bool scaled = (STORAGE != HSS_F32) && e.scale != nullptr;
    Reader rd{(decltype(Reader::p))e.samples};
    Window<W, VARIANT> acc;
    acc.reset();
    unsigned rej = 0;
    const int lane = threadIdx.x & (kWave - 1);
    // Traces go by in chunks of 64: lane i loads trace j0+i's index, geometry and scale
    // with one coalesced access, and the trace loop broadcasts them with v_readlane, so
    // no memory latency sits inside the per-trace loop.
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
            int first;
            float f;
            float t;
            bool r;
            constexpr bool FASTIDX = W != 0;
            bool ok = traveltime<MODEL, FASTIDX>(tg, t0, t0sq, e.m0, e.h0, e.apm, k, &t, &r);
            rej += r ? 1u : 0u;
            if constexpr (SHIFT) t = t + shift;
            ok = ok && window_index<FASTIDX>(t, e.dt, e.inv_dt, e.ib, e.tau, &first, &f);
            if (ok && active) {
                const float sc = scaled ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sc_l), jj)) : 1.0f;
                acc.add(rd, (size_t)ti * (size_t)e.ns + (size_t)first, f, sc, scaled, e.w);
            }
        }
    }
    acc.reduce(e.w, S, stack);
    *rejected += rej;
}

// Context for midpoint g of a batch (fields that do not depend on g come from `base`).
__device__ __forceinline__ EvalCtx midpoint_ctx(const EvalCtx &base, const int *mbeg,
                                                const int *mcnt, const float *mm0, int g) {
    EvalCtx e = base;
    e.beg = mbeg[g];
    e.cnt = mcnt[g];
    e.m0 = mm0[g];
    return e;
}

}  // namespace hss
