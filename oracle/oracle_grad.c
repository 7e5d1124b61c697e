/*
 * FP64 semblance and its gradient with respect to the traveltime parameters --
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Everything is carried in double: the traveltime, t/dt, the sample index and weight,
 * the interpolation and the sums.  S is then a continuous, piecewise smooth function
 * of the parameters (the window slides by one sample exactly where the weight wraps
 * from 1 to 0), and its gradient is exact away from those seams.  The parameter
 * derivatives of the traveltime come from forward-mode dual numbers pushed through the
 * same formulas the reference spells (cmp/linear_search.cu:60-62,
 * zocrs/linear_search.cu:83-90, oct/common.cuh:7-94); the semblance part is the
 * derivative of REDUCE_SEMBLANCE_STACK (common/include/gpu/interface.h:46-55):
 *   dS/dtheta = 2/(used*den) * (sum_j A_j dA_j - S * used * sum_ij u_ij du_ij),
 *   du_ij = (y_{k+j+1} - y_{k+j}) * d(t/dt)/dtheta.
 * The reference has no gradient; this is the oracle for the adjoint kernel.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

#define G_MAXW 32

typedef struct {
    double v, d[3];
} dn;

static dn dn_c(double v) { dn r = {v, {0, 0, 0}}; return r; }
static dn dn_add(dn a, dn b) { dn r = {a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}}; return r; }
static dn dn_sub(dn a, dn b) { dn r = {a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}}; return r; }
static dn dn_mul(dn a, dn b)
{
    dn r;
    r.v = a.v * b.v;
    for (int m = 0; m < 3; m++) r.d[m] = a.d[m] * b.v + a.v * b.d[m];
    return r;
}
static dn dn_scl(dn a, double s) { dn r = {a.v * s, {a.d[0] * s, a.d[1] * s, a.d[2] * s}}; return r; }
static dn dn_div(dn a, dn b)
{
    dn r;
    r.v = a.v / b.v;
    for (int m = 0; m < 3; m++) r.d[m] = (a.d[m] * b.v - a.v * b.d[m]) / (b.v * b.v);
    return r;
}
static dn dn_sqrt(dn a)
{
    dn r;
    r.v = sqrt(a.v);
    for (int m = 0; m < 3; m++) r.d[m] = r.v > 0 ? a.d[m] / (2.0 * r.v) : 0.0;
    return r;
}

/* OCT displaced midpoint in dual arithmetic (oct/common.cuh:7-35); c and slope carry
 * the derivatives. */
static int g_oct_mh(double h, double h0, double t0, double m0, dn c, dn slope, dn *mh)
{
    dn theta = dn_scl(slope, t0);
    double gamma = 2.0 * sqrt(h * h + h0 * h0);
    dn tn0_sq = dn_sub(dn_c(t0 * t0), dn_scl(c, h0 * h0));
    dn theta_sq = dn_mul(theta, theta);
    double gamma_sq = gamma * gamma;
    dn tn0_quad = dn_mul(tn0_sq, tn0_sq);
    dn arg = dn_add(dn_add(dn_mul(tn0_quad, tn0_quad), dn_scl(dn_mul(tn0_quad, theta_sq), gamma_sq)),
                    dn_scl(dn_mul(theta_sq, theta_sq), 16.0 * h * h * h0 * h0));
    if (arg.v < 0) return ORC_NEGATIVE_SQUARED_ROOT;
    arg = dn_add(dn_add(dn_scl(theta_sq, gamma_sq), dn_scl(tn0_quad, 2.0)), dn_scl(dn_sqrt(arg), 2.0));
    if (arg.v == 0) return ORC_DIVISION_BY_ZERO;
    *mh = dn_add(dn_c(m0), dn_div(dn_scl(theta, 2.0 * (h * h - h0 * h0)), dn_sqrt(arg)));
    return ORC_NO_ERROR;
}

/* OCT traveltime in dual arithmetic (oct/common.cuh:37-94). */
static int g_oct_time(double h, double h0, double t0, double m0, double m, dn mh, dn c, dn slope,
                      dn *out)
{
    dn tn0_sq = dn_sub(dn_c(t0 * t0), dn_scl(c, h0 * h0));
    dn dmh = dn_sub(mh, dn_c(m0));
    dn dmh2 = dn_mul(dmh, dmh);
    dn w1 = dn_sub(dn_c((h + h0) * (h + h0)), dmh2);
    dn w2 = dn_sub(dn_c((h - h0) * (h - h0)), dmh2);
    if (w1.v < 0 || w2.v < 0) return ORC_NEGATIVE_SQUARED_ROOT;
    dn u = dn_add(dn_sqrt(w1), dn_sqrt(w2));
    dn th = dn_c(t0);
    if (fabs(h) > fabs(h0)) {
        if (u.v == 0) return ORC_DIVISION_BY_ZERO;
        dn arg = dn_add(dn_scl(c, h * h), dn_mul(dn_div(dn_c(4.0 * h * h), dn_mul(u, u)), tn0_sq));
        if (arg.v < 0) return ORC_NEGATIVE_SQUARED_ROOT;
        th = dn_sqrt(arg);
    } else if (fabs(h) < fabs(h0)) {
        if (h0 == 0) return ORC_DIVISION_BY_ZERO;
        dn arg = dn_add(dn_scl(c, h * h), dn_mul(dn_scl(dn_mul(u, u), 1.0 / (4.0 * h0 * h0)), tn0_sq));
        if (arg.v < 0) return ORC_NEGATIVE_SQUARED_ROOT;
        th = dn_sqrt(arg);
    }
    dn tn_sq = dn_sub(dn_mul(th, th), dn_scl(c, h * h));
    if (th.v == 0 || tn0_sq.v == 0) return ORC_DIVISION_BY_ZERO;
    dn ah = dn_mul(dn_div(dn_scl(tn_sq, t0), dn_mul(th, tn0_sq)), slope);
    *out = dn_add(th, dn_mul(ah, dn_sub(dn_c(m), mh)));
    return ORC_NO_ERROR;
}

/* Traveltime of trace i in FP64 with d/d(c, p1, p2); 0 when the trace is dropped. */
static int g_traveltime(int model, const orc_gather *g, int i, double t0, dn c, dn p1, dn p2,
                        dn *t)
{
    const double h = g->hoff[i], m = g->mid[i];
    const double h2 = h * h;
    if (model == ORC_CMP) {
        *t = dn_sqrt(dn_add(dn_c(t0 * t0), dn_scl(c, h2)));
        return 1;
    }
    if (model == ORC_ZOCRS) {
        const double dm = m - g->m0;
        dn a = dn_add(dn_c(t0), dn_scl(p1, dm));
        dn tmp = dn_add(dn_add(dn_mul(a, a), dn_scl(p2, dm * dm)), dn_scl(c, h2));
        if (tmp.v < 0) return 0;
        *t = dn_sqrt(tmp);
        return 1;
    }
    dn mh;
    if (g_oct_mh(h, g->h0, t0, g->m0, c, p1, &mh) != ORC_NO_ERROR) return 0;
    if (fabs(m - mh.v) > g->apm) return 0;
    return g_oct_time(h, g->h0, t0, g->m0, m, mh, c, p1, t) == ORC_NO_ERROR;
}

void orc_eval_grad_f64(int model, const orc_gather *g, int s, const double *pv, int f32_index,
                       double *S, double *stack, double *grad)
{
    float prm32[3] = {0, 0, 0};
    const float v32 = (float)pv[0];
    prm32[0] = 4.0f / (v32 * v32);
    if (model != ORC_CMP) prm32[1] = (float)pv[1];
    if (model == ORC_ZOCRS) prm32[2] = (float)pv[2];
    const float t0f = (float)s * g->dt;
    const int np = model == ORC_CMP ? 1 : (model == ORC_ZOCRS ? 3 : 2);
    const double v = pv[0];
    dn c = dn_c(4.0 / (v * v));
    c.d[0] = 1.0;
    dn p1 = dn_c(np > 1 ? pv[1] : 0.0);
    p1.d[1] = 1.0;
    dn p2 = dn_c(np > 2 ? pv[2] : 0.0);
    p2.d[2] = 1.0;
    const double dt = g->dt, t0 = (double)s * dt;
    const int w = g->w, tau = g->tau;
    double A[G_MAXW] = {0}, G[G_MAXW][3], H[3] = {0, 0, 0};
    memset(G, 0, sizeof(G));
    double den = 0, lin = 0;
    unsigned used = 0;
    for (int i = 0; i < g->ntr; i++) {
        dn t;
        int k;
        double f;
        if (f32_index) {
            /* decisions, index and weight as the FP32 evaluation takes them; the
             * derivative of the FP64 traveltime at the same point */
            float t32, f32;
            if (!orc_traveltime_f32(model, g, i, t0f, prm32, &t32)) continue;
            if (!orc_window_index_f32(t32, g->dt, g->ns, tau, &k, &f32)) continue;
            f = (double)f32;
            if (!g_traveltime(model, g, i, t0, c, p1, p2, &t)) t = dn_c(0.0);
        } else {
            if (!g_traveltime(model, g, i, t0, c, p1, p2, &t)) continue;
            const double ti = t.v / dt;
            if (!(ti > -1.0 && ti < 2.0e9)) continue;
            k = (int)ti;
            if (k - tau < 0 || k + tau + 1 >= g->ns) continue;
            f = ti - (double)k;
        }
        double gi[3];
        for (int m = 0; m < 3; m++) gi[m] = t.d[m] / dt;
        const float *tr = g->samples + (size_t)i * g->ns + (k - tau);
        double e = 0;
        for (int j = 0; j < w; j++) {
            const double y0 = tr[j], d = (double)tr[j + 1] - y0;
            const double u = y0 + f * d;
            A[j] += u;
            lin += u;
            den += u * u;
            e += u * d;
            for (int m = 0; m < 3; m++) G[j][m] += d * gi[m];
        }
        for (int m = 0; m < 3; m++) H[m] += e * gi[m];
        used++;
    }
    double sem = 0, stk = 0, gt[3] = {0, 0, 0};
    if (used > 0 && den > 0) {
        double num = 0;
        for (int j = 0; j < w; j++) num += A[j] * A[j];
        sem = num / ((double)used * den);
        stk = lin / ((double)used * (double)w);
        const double sc = 2.0 / ((double)used * den);
        for (int m = 0; m < 3; m++) {
            double a = 0;
            for (int j = 0; j < w; j++) a += A[j] * G[j][m];
            gt[m] = sc * (a - sem * (double)used * H[m]);
        }
    }
    *S = sem;
    *stack = stk;
    if (grad) {
        grad[0] = gt[0] * (-8.0 / (v * v * v));  /* c = 4/v^2 */
        for (int m = 1; m < np; m++) grad[m] = gt[m];
    }
}
