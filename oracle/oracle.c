/*
 * CPU oracle for the semblance hot path -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Build: -O2 -ffp-contract=off, no fast-math (oracle/Makefile).  Every float
 * expression below is written in the order of the reference source text so that
 * with contraction off each operation is one IEEE-rounded step, as the reference
 * source spells it.  (The reference itself was built with fast-math:
 * cuda/Makefile:26, OpenCLComputeAlgorithm.cpp:45; its results can therefore differ
 * from these by rounding.)
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAX_WINDOW 32  /* MAX_WINDOW_SIZE, common/include/gpu/interface.h:10 */
#define ORC_F_FAC 0.85f    /* common/include/gpu/interface.h:5 */
#define ORC_CR 0.5f        /* common/include/gpu/interface.h:6 */

/* ------------------------------------------------------------------------- */
/* OCT displaced midpoint: cuda/include/semblance/kernel/oct/common.cuh:7-35   */
int orc_oct_displaced_midpoint(float h, float h0, float t0, float m0, float c,
                               float slope, float *mh)
{
    float theta, theta_sq, gamma, gamma_sq, tn0_sq, tn0_quad, sqrt_arg;

    theta = t0 * slope;
    gamma = 2.0f * sqrtf(h * h + h0 * h0);
    tn0_sq = t0 * t0 - c * h0 * h0;

    theta_sq = theta * theta;
    gamma_sq = gamma * gamma;
    tn0_quad = tn0_sq * tn0_sq;

    sqrt_arg = tn0_quad * tn0_quad + tn0_quad * theta_sq * gamma_sq
             + 16.0f * h * h * h0 * h0 * theta_sq * theta_sq;
    if (sqrt_arg < 0) {
        return ORC_NEGATIVE_SQUARED_ROOT;
    }
    sqrt_arg = theta_sq * gamma_sq + 2.0f * tn0_quad + 2.0f * sqrtf(sqrt_arg);
    if (sqrt_arg == 0) {
        return ORC_DIVISION_BY_ZERO;
    }
    *mh = m0 + 2.0f * theta * (h * h - h0 * h0) / sqrtf(sqrt_arg);
    return ORC_NO_ERROR;
}

/* OCT traveltime: cuda/include/semblance/kernel/oct/common.cuh:37-94 */
int orc_oct_time(float h, float h0, float t0, float m0, float m, float mh,
                 float c, float slope, float *out)
{
    float tn0_sq, tn_sq, w_sqrt_1, w_sqrt_2, u, sqrt_arg, ah, th;

    tn0_sq = t0 * t0 - c * h0 * h0;
    w_sqrt_1 = (h + h0) * (h + h0) - (mh - m0) * (mh - m0);
    w_sqrt_2 = (h - h0) * (h - h0) - (mh - m0) * (mh - m0);
    if (w_sqrt_1 < 0 || w_sqrt_2 < 0) {
        return ORC_NEGATIVE_SQUARED_ROOT;
    }
    u = sqrtf(w_sqrt_1) + sqrtf(w_sqrt_2);

    th = t0;
    if (fabsf(h) > fabsf(h0)) {
        if (!u) {
            return ORC_DIVISION_BY_ZERO;
        }
        sqrt_arg = c * h * h + 4.0f * h * h / (u * u) * tn0_sq;
        if (sqrt_arg < 0) {
            return ORC_NEGATIVE_SQUARED_ROOT;
        }
        th = sqrtf(sqrt_arg);
    } else if (fabsf(h) < fabsf(h0)) {
        if (!h0) {
            return ORC_DIVISION_BY_ZERO;
        }
        sqrt_arg = c * h * h + u * u / (4.0f * h0 * h0) * tn0_sq;
        if (sqrt_arg < 0) {
            return ORC_NEGATIVE_SQUARED_ROOT;
        }
        th = sqrtf(sqrt_arg);
    }

    tn_sq = th * th - c * h * h;
    if (!th || !tn0_sq) {
        return ORC_DIVISION_BY_ZERO;
    }
    ah = (t0 * tn_sq) / (th * tn0_sq) * slope;
    *out = th + ah * (m - mh);
    return ORC_NO_ERROR;
}

/* ------------------------------------------------------------------------- */
/* Philox4x32-10, the counter-based generator of Salmon et al., "Parallel random
 * numbers: as easy as 1, 2, 3" (SC'11): ten rounds of two 32x32->64 multiplies
 * with the published multipliers and Weyl key increments. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 24-bit uniform in (0, 1], the half-open side curand_uniform also uses. */
static inline float orc_u01(uint32_t x)
{
    return ((float)(x >> 8) + 1.0f) * 5.9604644775390625e-08f; /* 2^-24 */
}

/* ------------------------------------------------------------------------- */
/* Sample index and interpolation weight: the first three lines and the guard of
 * COMPUTE_SEMBLANCE (common/include/gpu/interface.h:26-31).  A NaN, infinite or
 * out-of-int-range index excludes the trace; in the reference that cast is
 * undefined behaviour.  For every finite in-range index the decision is the
 * reference's. */
static inline int orc_window_index(float t, float dt, int ns, int tau, int *k_out, float *f_out)
{
    float ti = t / dt;
    if (!(ti > -1.0f && ti < 2.0e9f)) {
        return 0;
    }
    int k = (int)ti;
    float f = ti - (float)k;
    if ((k - tau >= 0) && (k + tau + 1 < ns)) {
        *k_out = k;
        *f_out = f;
        return 1;
    }
    return 0;
}

/* Traveltime of one trace for one candidate.  Returns 1 with *t set when the trace
 * enters the window sum, 0 when the model skips it; *nu is incremented where the
 * reference increments notUsedCount.
 *   CMP    cmp/linear_search.cu:60-62
 *   ZO-CRS zocrs/linear_search.cu:83-90
 *   OCT    oct/linear_search.cu:84-98 */
static inline int orc_traveltime(int model, const orc_gather *g, int i, float t0,
                                 const float *prm, float *t, unsigned *nu)
{
    switch (model) {
    case ORC_CMP: {
        float c = prm[0];
        float h_sq = g->hoff2[i];
        *t = sqrtf(t0 * t0 + c * h_sq);
        return 1;
    }
    case ORC_ZOCRS: {
        float c = prm[0], a = prm[1], b = prm[2];
        float m = g->mid[i];
        float h_sq = g->hoff2[i];
        float dm = m - g->m0;
        float tmp = t0 + a * dm;
        tmp = tmp * tmp + b * dm * dm + c * h_sq;
        if (tmp >= 0) {
            *t = sqrtf(tmp);
            return 1;
        }
        return 0;
    }
    case ORC_OCT: {
        float c = prm[0], slope = prm[1];
        float m = g->mid[i], h = g->hoff[i], mh;
        int err = orc_oct_displaced_midpoint(h, g->h0, t0, g->m0, c, slope, &mh);
        if (err != ORC_NO_ERROR || fabsf(m - mh) > g->apm) {
            (*nu)++;
            return 0;
        }
        err = orc_oct_time(h, g->h0, t0, g->m0, m, mh, c, slope, t);
        return err == ORC_NO_ERROR;
    }
    }
    return 0;
}

/* Exported forms of the two helpers above for the gradient oracle (oracle_grad.c). */
int orc_traveltime_f32(int model, const orc_gather *g, int i, float t0, const float *prm,
                       float *t)
{
    unsigned nu = 0;
    return orc_traveltime(model, g, i, t0, prm, t, &nu);
}

int orc_window_index_f32(float t, float dt, int ns, int tau, int *k_out, float *f_out)
{
    return orc_window_index(t, dt, ns, tau, k_out, f_out);
}

/* One (t0, candidate) semblance in FP32: the body of computeSemblancesFor*
 * (cmp/linear_search.cu:37-80) with COMPUTE_SEMBLANCE (interface.h:26-44) and
 * REDUCE_SEMBLANCE_STACK (interface.h:46-55).  shift_n != 0 adds the stretch-free
 * shift t = t_n + n*dt (cmp/stretch_free.cu:56). */
void orc_eval_f32(int model, const orc_gather *g, int s, int shift_n,
                  const float *prm, float *S, float *stack,
                  unsigned *used_out, unsigned *not_used_out)
{
    float A[ORC_MAX_WINDOW];
    float den = 0, lin = 0;
    unsigned used = 0, nu = 0;
    const int w = g->w;
    for (int j = 0; j < ORC_MAX_WINDOW; j++) {
        A[j] = 0;
    }
    float t0 = (float)s * g->dt;
    for (int i = 0; i < g->ntr; i++) {
        float t;
        if (!orc_traveltime(model, g, i, t0, prm, &t, &nu)) {
            continue;
        }
        if (shift_n != 0) {
            t = t + (float)shift_n * g->dt;
        }
        int k;
        float f;
        if (!orc_window_index(t, g->dt, g->ns, g->tau, &k, &f)) {
            continue;
        }
        const float *tr = g->samples + (size_t)i * g->ns;
        int kk = k - g->tau;
        float y0, y1 = tr[kk], u;
        for (int j = 0; j < w; j++, kk++) {
            y0 = y1;
            y1 = tr[kk + 1];
            u = (y1 - y0) * f + y0;
            A[j] += u;
            lin += u;
            den += u * u;
        }
        used++;
    }
    float sem = 0, stk = 0;
    if (used > 0) {
        float num = 0;
        for (int j = 0; j < w; j++) {
            num += A[j] * A[j];
        }
        sem = num / ((float)used * den);
        stk = lin / (float)(used * (unsigned)w);
    }
    *S = sem;
    *stack = stk;
    if (used_out) *used_out = used;
    if (not_used_out) *not_used_out = nu;
}

/* FP64-accumulating twin of orc_eval_f32: same traveltime, index and inclusion
 * decisions (FP32), interpolation and sums in FP64. */
void orc_eval_f64(int model, const orc_gather *g, int s, int shift_n,
                  const float *prm, double *S, double *stack,
                  unsigned *used_out, unsigned *not_used_out)
{
    double A[ORC_MAX_WINDOW];
    double den = 0, lin = 0;
    unsigned used = 0, nu = 0;
    const int w = g->w;
    for (int j = 0; j < ORC_MAX_WINDOW; j++) {
        A[j] = 0;
    }
    float t0 = (float)s * g->dt;
    for (int i = 0; i < g->ntr; i++) {
        float t;
        if (!orc_traveltime(model, g, i, t0, prm, &t, &nu)) {
            continue;
        }
        if (shift_n != 0) {
            t = t + (float)shift_n * g->dt;
        }
        int k;
        float f;
        if (!orc_window_index(t, g->dt, g->ns, g->tau, &k, &f)) {
            continue;
        }
        const float *tr = g->samples + (size_t)i * g->ns;
        int kk = k - g->tau;
        double fd = (double)f;
        for (int j = 0; j < w; j++, kk++) {
            double y0 = tr[kk], y1 = tr[kk + 1];
            double u = (y1 - y0) * fd + y0;
            A[j] += u;
            lin += u;
            den += u * u;
        }
        used++;
    }
    double sem = 0, stk = 0;
    if (used > 0) {
        double num = 0;
        for (int j = 0; j < w; j++) {
            num += A[j] * A[j];
        }
        sem = num / ((double)used * den);
        stk = lin / ((double)used * (double)w);
    }
    *S = sem;
    *stack = stk;
    if (used_out) *used_out = used;
    if (not_used_out) *not_used_out = nu;
}

static int orc_model_np(int model)
{
    return model == ORC_CMP ? 1 : (model == ORC_ZOCRS ? 3 : (model == ORC_OCT ? 2 : -1));
}

/* ------------------------------------------------------------------------- */
/* buildParameterArrayFor*: v-major index decode, storage c = 4/v^2.
 * CMP    cmp/linear_search.cu:4-17
 * ZO-CRS zocrs/linear_search.cu:4-34  (iv = (k/(nA*nB))%nV, iA = (k/nB)%nA, iB = k%nB)
 * OCT    oct/linear_search.cu:4-28    (iv = (k/nS)%nV,     iS = k%nS)              */
int orc_build_parameter_array(int model, const float *lower, const float *step,
                              const int *count, float *out)
{
    int np = orc_model_np(model);
    if (np < 0) return -ORC_INVALID_MODEL;
    unsigned total = 1;
    for (int d = 0; d < np; d++) total *= (unsigned)count[d];
    for (unsigned k = 0; k < total; k++) {
        if (model == ORC_CMP) {
            float v = lower[0] + (float)k * step[0];
            out[k] = 4.0f / (v * v);
        } else if (model == ORC_ZOCRS) {
            unsigned nV = (unsigned)count[0], nA = (unsigned)count[1], nB = (unsigned)count[2];
            unsigned iv = (k / (nA * nB)) % nV, ia = (k / nB) % nA, ib = k % nB;
            float v = lower[0] + (float)iv * step[0];
            float a = lower[1] + (float)ia * step[1];
            float b = lower[2] + (float)ib * step[2];
            out[3 * k] = 4.0f / (v * v);
            out[3 * k + 1] = a;
            out[3 * k + 2] = b;
        } else {
            unsigned nV = (unsigned)count[0], nS = (unsigned)count[1];
            unsigned iv = (k / nS) % nV, is = k % nS;
            float v = lower[0] + (float)iv * step[0];
            float sl = lower[1] + (float)is * step[1];
            out[2 * k] = 4.0f / (v * v);
            out[2 * k + 1] = sl;
        }
    }
    return (int)total;
}

/* ------------------------------------------------------------------------- */
/* Grid search: computeSemblancesFor* then selectBestSemblancesFor*
 * (cmp/linear_search.cu:83-115, zocrs/linear_search.cu:111-151,
 *  oct/linear_search.cu:116-153): best starts at -1 and moves on strict '>', so the
 * lowest index among equal maxima wins and NaN never wins.  When nothing wins the
 * reference writes uninitialised stack/parameter values; the oracle writes 0. */
void orc_grid_f32(int model, const orc_gather *g, const float *params, int P,
                  float *sembl, float *stack, float *result, int *best_idx,
                  double *not_used_total, int nthreads)
{
    const int np = orc_model_np(model);
    const int ns = g->ns;
    double nu_total = 0;
    if (nthreads <= 0) nthreads = 1;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads) reduction(+:nu_total)
    for (int s = 0; s < ns; s++) {
        float best = -1, bstack = 0;
        int bi = -1;
        for (int p = 0; p < P; p++) {
            float S, stk;
            unsigned nu;
            orc_eval_f32(model, g, s, 0, params + (size_t)p * np, &S, &stk, NULL, &nu);
            nu_total += nu;
            if (sembl) sembl[(size_t)s * P + p] = S;
            if (stack) stack[(size_t)s * P + p] = stk;
            if (S > best) {
                best = S;
                bstack = stk;
                bi = p;
            }
        }
        result[s] = best;
        result[ns + s] = bi >= 0 ? bstack : 0.0f;
        if (bi >= 0) {
            const float *pr = params + (size_t)bi * np;
            result[2 * ns + s] = 2 / sqrtf(pr[0]);
            for (int d = 1; d < np; d++) result[(2 + d) * ns + s] = pr[d];
        } else {
            for (int d = 0; d < np; d++) result[(2 + d) * ns + s] = 0.0f;
        }
        if (best_idx) best_idx[s] = bi;
    }
    if (not_used_total) *not_used_total = nu_total;
}

void orc_grid_f64(int model, const orc_gather *g, const float *params, int P,
                  double *sembl, double *stack, double *result, int *best_idx,
                  int nthreads)
{
    const int np = orc_model_np(model);
    const int ns = g->ns;
    if (nthreads <= 0) nthreads = 1;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int s = 0; s < ns; s++) {
        double best = -1, bstack = 0;
        int bi = -1;
        for (int p = 0; p < P; p++) {
            double S, stk;
            orc_eval_f64(model, g, s, 0, params + (size_t)p * np, &S, &stk, NULL, NULL);
            if (sembl) sembl[(size_t)s * P + p] = S;
            if (stack) stack[(size_t)s * P + p] = stk;
            if (S > best) {
                best = S;
                bstack = stk;
                bi = p;
            }
        }
        result[s] = best;
        result[ns + s] = bi >= 0 ? bstack : 0.0;
        if (bi >= 0) {
            const float *pr = params + (size_t)bi * np;
            result[2 * ns + s] = 2.0 / sqrt((double)pr[0]);
            for (int d = 1; d < np; d++) result[(2 + d) * ns + s] = pr[d];
        } else {
            for (int d = 0; d < np; d++) result[(2 + d) * ns + s] = 0.0;
        }
        if (best_idx) best_idx[s] = bi;
    }
}

/* ------------------------------------------------------------------------- */
/* Stretch-free: for each (s, n) the candidate is the parameter picked at sample
 * s_n = s - n, c_n = 4/v_n^2, and the traveltime is shifted by n*dt
 * (cmp/stretch_free.cu:27-58; zocrs/stretch_free.cu:34-70; oct/stretch_free.cu:36-84);
 * argmax over n by common/stretch_free.cu:3-35. */
static void orc_sf_candidate(int model, const float *vparams, int ns, int sn, float *prm)
{
    float v = vparams[sn];
    prm[0] = 4.0f / (v * v);
    if (model == ORC_ZOCRS) {
        prm[1] = vparams[ns + sn];
        prm[2] = vparams[2 * ns + sn];
    } else if (model == ORC_OCT) {
        prm[1] = vparams[ns + sn];
    }
}

void orc_stretch_free_f32(int model, const orc_gather *g, const float *vparams,
                          const float *nvals, int nN, float *sembl, float *stack,
                          float *result, int nthreads)
{
    const int ns = g->ns;
    if (nthreads <= 0) nthreads = 1;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int s = 0; s < ns; s++) {
        float best = -1, bstack = 0, bn = 0;
        int found = 0;
        for (int q = 0; q < nN; q++) {
            int n = (int)nvals[q];
            int sn = s - n;
            float S = 0, stk = 0;
            if (sn >= 0 && sn < ns) {
                float prm[3];
                orc_sf_candidate(model, vparams, ns, sn, prm);
                orc_eval_f32(model, g, sn, n, prm, &S, &stk, NULL, NULL);
            }
            if (sembl) sembl[(size_t)s * nN + q] = S;
            if (stack) stack[(size_t)s * nN + q] = stk;
            if (S > best) {
                best = S;
                bstack = stk;
                bn = nvals[q];
                found = 1;
            }
        }
        result[s] = best;
        result[ns + s] = found ? bstack : 0.0f;
        result[2 * ns + s] = found ? bn : 0.0f;
    }
}

void orc_stretch_free_f64(int model, const orc_gather *g, const float *vparams,
                          const float *nvals, int nN, double *result, int nthreads)
{
    const int ns = g->ns;
    if (nthreads <= 0) nthreads = 1;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int s = 0; s < ns; s++) {
        double best = -1, bstack = 0, bn = 0;
        int found = 0;
        for (int q = 0; q < nN; q++) {
            int n = (int)nvals[q];
            int sn = s - n;
            double S = 0, stk = 0;
            if (sn >= 0 && sn < ns) {
                float prm[3];
                orc_sf_candidate(model, vparams, ns, sn, prm);
                orc_eval_f64(model, g, sn, n, prm, &S, &stk, NULL, NULL);
            }
            if (S > best) {
                best = S;
                bstack = stk;
                bn = nvals[q];
                found = 1;
            }
        }
        result[s] = best;
        result[ns + s] = found ? bstack : 0.0;
        result[2 * ns + s] = found ? bn : 0.0;
    }
}

/* ------------------------------------------------------------------------- */
/* Differential evolution for one midpoint.  Operator semantics follow
 * common/differential_evolution.cu:19-155 (synchronous generations; donors drawn
 * per parameter as floor(U*(NP-1)); crossover index l = floor(U*(D-1));
 * u = (r > CR && l != d) ? x : v; selection on strict '>' of S) and the best-of-
 * population pick of cmp|zocrs|oct/differential_evolution.cu selectBestIndividuals*.
 * The uniforms come from Philox4x32-10 instead of curand: for sample s, individual
 * i, generation gen, uniform q lives in word q%4 of
 *   philox(ctr = {s, stream_id, i, (gen << 4) | (q / 4)}, key = {seed_lo, seed_hi}).
 * Generation 0 draws q = 0..D-1 for the start population; generation >= 1 draws
 * q = 3d+{0,1,2} for the donors of parameter d, q = 3D for l and q = 3D+1+d for the
 * crossover coin of parameter d. */
static void orc_de_uniforms(uint64_t seed, uint32_t stream_id, int s, int i, int gen,
                            int nq, float *u)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int b = 0; b * 4 < nq; b++) {
        uint32_t ctr[4] = {(uint32_t)s, stream_id, (uint32_t)i, ((uint32_t)gen << 4) | (uint32_t)b};
        uint32_t r[4];
        orc_philox4x32_10(ctr, key, r);
        for (int q = 0; q < 4 && b * 4 + q < nq; q++) u[b * 4 + q] = orc_u01(r[q]);
    }
}

static void orc_de_eval(int model, const orc_gather *g, int s, const float *xv,
                        float *S, float *stk, uint64_t *not_used)
{
    float prm[3];
    float v = xv[0];
    prm[0] = 4.0f / (v * v);
    if (model == ORC_ZOCRS) {
        prm[1] = xv[1];
        prm[2] = xv[2];
    } else if (model == ORC_OCT) {
        prm[1] = xv[1];
    }
    unsigned nu = 0;
    orc_eval_f32(model, g, s, 0, prm, S, stk, NULL, &nu);
    *not_used += nu;
}

/* One start individual from D uniforms: x_d = lower_d + U_d * (upper_d - lower_d)
 * (common/differential_evolution.cu:19-40). */
void orc_de_start_individual(int D, const float *lower, const float *upper, const float *u,
                             float *xi)
{
    for (int d = 0; d < D; d++) {
        xi[d] = lower[d] + u[d] * (upper[d] - lower[d]);
    }
}

/* Mutant and trial of individual i of population x [NP][D] from explicit uniforms:
 * u[3d + 0..2] are the donor draws of parameter d (mutatePopulations,
 * common/differential_evolution.cu:43-86: r = floor(U * (NP - 1)), v = x_r1 + F (x_r2 - x_r3),
 * fminf with the upper bound, then fmaxf with the lower), u[3D] the forced index
 * l = floor(U * (D - 1)) and u[3D + 1 + d] the coin of parameter d (crossoverPopulations,
 * :88-117: the trial keeps x_d when coin > CR and d != l, else takes v_d). */
void orc_de_trial(int D, int NP, const float *x, int i, const float *lower, const float *upper,
                  const float *u, float *mutant, float *trial)
{
    for (int d = 0; d < D; d++) {
        unsigned r1 = (unsigned)(u[3 * d] * (float)(NP - 1));
        unsigned r2 = (unsigned)(u[3 * d + 1] * (float)(NP - 1));
        unsigned r3 = (unsigned)(u[3 * d + 2] * (float)(NP - 1));
        float nv = x[r1 * D + d] + ORC_F_FAC * (x[r2 * D + d] - x[r3 * D + d]);
        nv = fminf(nv, upper[d]);
        nv = fmaxf(nv, lower[d]);
        mutant[d] = nv;
    }
    unsigned l = (unsigned)(u[3 * D] * (float)(D - 1));
    for (int d = 0; d < D; d++) {
        float r = u[3 * D + 1 + d];
        trial[d] = (r > ORC_CR && l != (unsigned)d) ? x[i * D + d] : mutant[d];
    }
}

void orc_de_f32(int model, const orc_gather *g, const float *lower, const float *upper,
                int NP, int G, uint64_t seed, uint32_t stream_id, float *result,
                float *x_out, float *fx_out, uint64_t *not_used_total, int nthreads)
{
    const int D = orc_model_np(model);
    const int ns = g->ns;
    if (nthreads <= 0) nthreads = 1;
    (void)nthreads;
    uint64_t nu_all = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : nu_all)
    {
        float *x = (float *)malloc(sizeof(float) * NP * D);
        float *vv = (float *)malloc(sizeof(float) * NP * D);
        float *uu = (float *)malloc(sizeof(float) * NP * D);
        float *fx = (float *)malloc(sizeof(float) * NP * 2);
        float *fu = (float *)malloc(sizeof(float) * NP * 2);
#pragma omp for schedule(dynamic, 4)
        for (int s = 0; s < ns; s++) {
            float un[16];
            for (int i = 0; i < NP; i++) {
                orc_de_uniforms(seed, stream_id, s, i, 0, D, un);
                orc_de_start_individual(D, lower, upper, un, x + i * D);
            }
            for (int i = 0; i < NP; i++) {
                orc_de_eval(model, g, s, x + i * D, &fx[2 * i], &fx[2 * i + 1], &nu_all);
            }
            for (int gen = 1; gen < G; gen++) {
                /* synchronous generation: every trial is formed from the population of
                 * the previous generation before any individual is replaced */
                for (int i = 0; i < NP; i++) {
                    orc_de_uniforms(seed, stream_id, s, i, gen, 4 * D + 1, un);
                    orc_de_trial(D, NP, x, i, lower, upper, un, vv + i * D, uu + i * D);
                }
                for (int i = 0; i < NP; i++) {
                    orc_de_eval(model, g, s, uu + i * D, &fu[2 * i], &fu[2 * i + 1], &nu_all);
                }
                for (int i = 0; i < NP; i++) {
                    if (fu[2 * i] > fx[2 * i]) {
                        for (int d = 0; d < D; d++) x[i * D + d] = uu[i * D + d];
                        fx[2 * i] = fu[2 * i];
                        fx[2 * i + 1] = fu[2 * i + 1];
                    }
                }
            }
            float best = -1, bstack = 0;
            int bi = -1;
            for (int i = 0; i < NP; i++) {
                if (fx[2 * i] > best) {
                    best = fx[2 * i];
                    bstack = fx[2 * i + 1];
                    bi = i;
                }
            }
            result[s] = best;
            result[ns + s] = bi >= 0 ? bstack : 0.0f;
            for (int d = 0; d < D; d++) result[(2 + d) * ns + s] = bi >= 0 ? x[bi * D + d] : 0.0f;
            if (x_out) memcpy(x_out + (size_t)s * NP * D, x, sizeof(float) * NP * D);
            if (fx_out) memcpy(fx_out + (size_t)s * NP * 2, fx, sizeof(float) * NP * 2);
        }
        free(x); free(vv); free(uu); free(fx); free(fu);
    }
    if (not_used_total) *not_used_total = nu_all;
}

/* ------------------------------------------------------------------------- */
/* Trace selection masks. */
void orc_select_traces(int model, const float *mid, const float *hoff, int ntot,
                       float m0, float apm, float h0, int ns, float dt,
                       const float *pairs, int npairs, int per_sample,
                       unsigned char *mask)
{
    for (int i = 0; i < ntot; i++) {
        if (model == ORC_CMP) {
            mask[i] = (m0 == mid[i]);           /* cmp/common.cu:13 */
        } else if (model == ORC_ZOCRS) {
            mask[i] = fabsf(m0 - mid[i]) <= apm; /* zocrs/common.cu:14 */
        } else {
            /* oct/linear_search.cu:170-190: OR over (s, pair) with early exit */
            float m = mid[i], h = hoff[i], mh;
            unsigned char used = 0;
            for (int s = 0; !used && s < ns; s++) {
                float t0 = (float)s * dt;
                const float *pp = pairs + (per_sample ? (size_t)s * npairs * 2 : 0);
                for (int p = 0; !used && p < npairs; p++) {
                    float c = pp[2 * p], slope = pp[2 * p + 1];
                    int err = orc_oct_displaced_midpoint(h, h0, t0, m0, c, slope, &mh);
                    used = (err == ORC_NO_ERROR && fabsf(m - mh) <= apm);
                }
            }
            mask[i] = used;
        }
    }
}
