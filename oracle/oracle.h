/*
 * CPU oracle for the semblance hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker.  The product path never calls it.
 *
 * A plain-C, IEEE-754 restatement of the reference arithmetic
 * (lmcad-unicamp/HighPerformanceSeismicStacking).  It is compiled with
 * -ffp-contract=off and without fast-math so every float operation is one
 * correctly rounded IEEE operation, performed in the order the reference source
 * text spells it.  Each function names the reference file:line it restates.
 *
 * Two arithmetic flavours:
 *   *_f32   : the reference's FP32 arithmetic, op for op.
 *   *_f64   : discrete decisions (sample index k, trace inclusion, OCT error
 *             branches) are taken exactly as in *_f32; the interpolation and the
 *             window sums are then carried in FP64.  This measures the rounding
 *             error of a kernel's sums without mixing in the (discontinuous)
 *             effect of a flipped index decision.
 */
#ifndef HSS_ORACLE_H
#define HSS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* traveltime model ids (common/include/capability.h:3-7) */
enum { ORC_CMP = 0, ORC_ZOCRS = 1, ORC_OCT = 2 };

/* error codes (common/include/gpu/interface.h:12-18) */
enum {
    ORC_NO_ERROR = 0,
    ORC_DIVISION_BY_ZERO = 1,
    ORC_NEGATIVE_SQUARED_ROOT = 2,
    ORC_INVALID_RANGE = 3,
    ORC_INVALID_MODEL = 4
};

/* The traces selected for one midpoint, in the layout the reference's
 * FILT_* device arrays hold (ComputeAlgorithm.cpp:107-153). */
typedef struct {
    const float *samples; /* [ntr][ns] trace-major                       */
    const float *mid;     /* [ntr] midpoint m                            */
    const float *hoff;    /* [ntr] signed half offset h                  */
    const float *hoff2;   /* [ntr] h*h as the host computes it           */
    int ntr;
    int ns;
    float dt;             /* sample period in seconds                    */
    int tau;              /* tauIndexDisplacement                        */
    int w;                /* window size 2*tau+1                         */
    float m0;
    float h0;
    float apm;
} orc_gather;

/* ---- OCT device functions (cuda/include/semblance/kernel/oct/common.cuh:7-94) */
int orc_oct_displaced_midpoint(float h, float h0, float t0, float m0, float c,
                               float slope, float *mh);
int orc_oct_time(float h, float h0, float t0, float m0, float m, float mh,
                 float c, float slope, float *out);

/* ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- one (t0, candidate) evaluation.
 * prm holds the candidate in storage form: CMP {c}; ZO-CRS {c, A, B};
 * OCT {c, slope} with c = 4/v^2.  shift_n is the stretch-free sample shift
 * (0 for grid and DE).  Outputs S, stack, used and notUsed counts. */
void orc_eval_f32(int model, const orc_gather *g, int s, int shift_n,
                  const float *prm, float *S, float *stack,
                  unsigned *used, unsigned *not_used);
void orc_eval_f64(int model, const orc_gather *g, int s, int shift_n,
                  const float *prm, double *S, double *stack,
                  unsigned *used, unsigned *not_used);

/* ---- grid parameter arrays
 * (cmp/linear_search.cu:4-17, zocrs/linear_search.cu:4-34, oct/linear_search.cu:4-28) */
int orc_build_parameter_array(int model, const float *lower, const float *step,
                              const int *count, float *out);

/* ---- grid search over one midpoint: computeSemblancesFor* + selectBest*
 * (cmp/linear_search.cu:19-115, zocrs/linear_search.cu:36-151,
 *  oct/linear_search.cu:30-153).
 * params [P][np] storage form.  sembl/stack [ns][P] (nullable).
 * result [(2+np)][ns].  best_idx [ns] (nullable; -1 when no candidate won).
 * not_used_total: sum of notUsed over all (s, p) (nullable).
 * nthreads: OpenMP threads over t0 (<=0: 1). */
void orc_grid_f32(int model, const orc_gather *g, const float *params, int P,
                  float *sembl, float *stack, float *result, int *best_idx,
                  double *not_used_total, int nthreads);
void orc_grid_f64(int model, const orc_gather *g, const float *params, int P,
                  double *sembl, double *stack, double *result, int *best_idx,
                  int nthreads);

/* ---- stretch-free re-stack (cmp/stretch_free.cu:4-66, zocrs/stretch_free.cu:4-85,
 * oct/stretch_free.cu:5-97, common/stretch_free.cu:3-35).
 * vparams [np][ns] in v space (velocity, then A,B or slope); nvals [nN].
 * result [3][ns] = S, stack, n. */
void orc_stretch_free_f32(int model, const orc_gather *g, const float *vparams,
                          const float *nvals, int nN, float *sembl, float *stack,
                          float *result, int nthreads);
void orc_stretch_free_f64(int model, const orc_gather *g, const float *vparams,
                          const float *nvals, int nN, double *result, int nthreads);

/* ---- differential evolution over one midpoint with a Philox stream
 * (common/differential_evolution.cu:19-155, cmp|zocrs|oct/differential_evolution.cu).
 * lower/upper [np] in v space.  result [(2+np)][ns].
 * x_out [ns][NP][np], fx_out [ns][NP][2] (nullable): final population.
 * not_used_total (nullable): the traces the model rejected, summed over every evaluation
 * of the run (the start population and each generation's trials) and every sample. */
void orc_de_f32(int model, const orc_gather *g, const float *lower, const float *upper,
                int NP, int G, uint64_t seed, uint32_t stream_id, float *result,
                float *x_out, float *fx_out, uint64_t *not_used_total, int nthreads);

/* The DE operators on explicit uniforms (the pieces orc_de_f32 is made of):
 * orc_de_start_individual: x_d = lower_d + u[d] * (upper_d - lower_d);
 * orc_de_trial: the mutant and trial of individual i of x [NP][D], with u[3d+0..2] the
 * donor draws of parameter d, u[3D] the forced crossover index draw and u[3D+1+d] the
 * crossover coin of parameter d. */
void orc_de_start_individual(int D, const float *lower, const float *upper, const float *u,
                             float *xi);
void orc_de_trial(int D, int NP, const float *x, int i, const float *lower, const float *upper,
                  const float *u, float *mutant, float *trial);

/* FP64 semblance, stack and dS/d(parameters) at one sample s for one parameter point
 * pv in v space (v; v,A,B; v,slope) (oracle_grad.c).  grad[np] may be NULL.
 *   f32_index = 0: everything in double, including the traveltime, t/dt and the sample
 *                  index -- S is then the piecewise-smooth function the gradient is
 *                  exact for;
 *   f32_index = 1: the traveltime, t/dt, sample index, weight and trace decisions are
 *                  taken in FP32 exactly as orc_eval_f32 takes them (c = 4/v^2 in
 *                  float), the traveltime derivatives and all sums in double -- the
 *                  counterpart of a FP32 kernel, free of index disagreements at seams. */
void orc_eval_grad_f64(int model, const orc_gather *g, int s, const double *pv, int f32_index,
                       double *S, double *stack, double *grad);
int orc_traveltime_f32(int model, const orc_gather *g, int i, float t0, const float *prm,
                       float *t);
int orc_window_index_f32(float t, float dt, int ns, int tau, int *k_out, float *f_out);

/* ---- trace selection masks over a whole line
 * (cmp/common.cu:3-15, zocrs/common.cu:3-16, oct/linear_search.cu:155-192,
 *  oct/differential_evolution.cu:130-168, oct/stretch_free.cu:99-138).
 * OCT: pairs [npairs_per_s * (per_sample ? ns : 1)][2] = {c, slope} storage form;
 * per_sample=0: the same npairs for every t0 (grid); per_sample=1: pairs for sample s
 * start at s*npairs (DE populations, stretch-free per-sample parameters). */
void orc_select_traces(int model, const float *mid, const float *hoff, int ntot,
                       float m0, float apm, float h0, int ns, float dt,
                       const float *pairs, int npairs, int per_sample,
                       unsigned char *mask);

/* default_random_engine + uniform_real_distribution<float> draw of the OCT grid
 * selection samples (CudaLinearSearchAlgorithm.cu:268-308), written as
 * [count][2] = {4/v^2, slope}.  Implemented over libstdc++ <random>. */
void orc_oct_selection_samples(float vmin, float vmax, float smin, float smax,
                               int count, float *out);

#ifdef __cplusplus
}
#endif
#endif
