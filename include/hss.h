/*
 * hss.h -- C ABI of the MI355X (gfx950) semblance backend.
 *
 * This is the layer beneath the reference's C++ plugin interface
 * (lmcad-unicamp/HighPerformanceSeismicStacking, common/include/semblance/algorithm/
 * ComputeAlgorithm.hpp:29-100 and its Cuda* subclasses).  Each entry point replaces
 * one reference kernel family; the citation on each declaration names the CUDA
 * launch site it stands in for.  The host classes in
 * highperformanceseismicstacking_amd/csrc/host (HipLinearSearchAlgorithm,
 * HipDifferentialEvolutionAlgorithm, HipStretchFreeAlgorithm) call exactly these.
 *
 * Conventions
 *   - plain C: raw device pointers, element counts, an opaque HIP stream
 *     (hipStream_t passed as void*; NULL = the null stream);
 *   - every call is asynchronous on `stream` and returns an int status:
 *     0 = success, >0 = a hipError_t value, <0 = an hss_status argument error;
 *   - the caller owns all memory; no entry point allocates device memory, except
 *     where a declaration says "allocates" (none in this version);
 *   - the first semblance call on a device for a given dt runs a one-time check of the
 *     short square-root and division sequences the kernels use (hss_verify_index_ops)
 *     and waits for it (outside stream capture: hss_index_ops_state); later calls are
 *     fully asynchronous;
 *   - floating point in/out is float32; trace samples may be stored as float32,
 *     float16 or bfloat16 (hss_traces.storage) with an optional per-trace float32 scale
 *     (float16 needs one for range; bfloat16, with float32's exponent range, is best
 *     stored without: its windows are then summed as float32 windows).
 *
 * Model ids follow traveltime_t (common/include/capability.h:3-7).
 */
#ifndef HSS_H
#define HSS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HSS_ABI_VERSION 2

/* traveltime models (common/include/capability.h:3-7) */
enum hss_model { HSS_CMP = 0, HSS_ZOCRS = 1, HSS_OCT = 2 };

/* trace sample storage */
enum hss_storage { HSS_F32 = 0, HSS_F16 = 1, HSS_BF16 = 2 };

/* arithmetic variant of the window sums
 *   HSS_EXACT: the reference op order, no FMA contraction -- results are bit-identical
 *              to an IEEE FP32 evaluation of common/include/gpu/interface.h:26-55;
 *   HSS_FAST : FMA-contracted interpolation and sums, stack sum derived from the
 *              window components; within 1e-5 relative of HSS_EXACT;
 *   HSS_KAHAN: as HSS_FAST with Kahan-compensated window sums (A_j, sum of squares,
 *              and the sums over the window); no reference counterpart. */
enum hss_variant { HSS_EXACT = 0, HSS_FAST = 1, HSS_KAHAN = 2 };

/* argument errors (negative statuses) */
enum hss_status {
    HSS_OK = 0,
    HSS_ERR_INVALID_MODEL = -1,
    HSS_ERR_INVALID_WINDOW = -2,   /* w != 2*tau+1, or w > 32 (MAX_WINDOW_SIZE) */
    HSS_ERR_INVALID_SIZE = -3,     /* a count is negative, zero where it must not be, or too large */
    HSS_ERR_NULL_POINTER = -4,
    HSS_ERR_INVALID_STORAGE = -5,  /* unknown storage, or samples not 4-byte aligned */
    HSS_ERR_INVALID_VARIANT = -6,
    HSS_ERR_NO_DEVICE = -7,
    HSS_ERR_SCRATCH_LIMIT = -8     /* the call needs more than the static scratch holds */
};

/* All traces of a line (or of one midpoint), resident on the device, in the order the
 * reference's Gather sorts them (by midpoint, then half offset; Trace.cpp:63-74).
 * The float arrays are the GatherData MDPNT / HLFOFFST / HLFOFFST_SQ columns
 * (ComputeAlgorithm.hpp:18-27, ComputeAlgorithm.cpp:31-61). */
typedef struct hss_traces {
    const void  *samples;   /* [ntr][ns] in `storage`, 4-byte aligned              */
    const float *scale;     /* [ntr] per-trace scale; NULL => 1                     */
    const float *mid;       /* [ntr] midpoint m                                    */
    const float *hoff;      /* [ntr] signed half offset h                          */
    const float *hoff2;     /* [ntr] h*h                                           */
    int32_t ntr;
    int32_t ns;             /* samples per trace                                   */
    float   dt;             /* sample period in seconds (Gather.cpp:73)            */
    int32_t tau;            /* tauIndexDisplacement = (int)(tau_s / dt) (Gather.cpp:75) */
    int32_t w;              /* window size 2*tau+1 (Gather.cpp:76), <= 32          */
    int32_t storage;        /* enum hss_storage                                    */
} hss_traces;

/* The midpoints to process and the traces each one uses.  Midpoint g uses traces
 * index ? index[begin[g] .. begin[g]+count[g]) : [begin[g], begin[g]+count[g]),
 * i.e. the reference's per-midpoint FILT_* compaction (ComputeAlgorithm.cpp:107-153)
 * expressed as ranges into the resident line.  The list's order is the order of the
 * sums: HSS_EXACT equals the reference bit for bit when the list is in line order.  The
 * OCT kernels (grid, DE, stretch-free, adjoint) keep the offset-dependent half of the
 * traveltime while consecutive traces share a half offset (the same bits as computing it
 * again), so lists grouped by half offset run 1.8-3x faster; the non-exact variants of
 * the host classes pass them so, which changes the sums' rounding only.
 *
 * `order` (no counterpart in the reference; NULL = 0, 1, ..., n-1) is the order in which
 * a launch takes up the midpoints: launch slot j works on midpoint order[j].  It changes
 * nothing but the schedule -- midpoint g keeps its result row g, its workspace rows and
 * its DE random stream (stream_base + g) -- so results are the same bits in any order.
 * The launch deals slots to the chip's 8 XCDs in 8 contiguous ranges; an order that gives
 * every range the same total work (largest first) keeps one XCD from finishing last
 * (hss_host.h: xcd_balanced_order).  It should be a permutation of 0..n-1: slots whose
 * entry lies outside [0, n) do nothing, and a midpoint listed twice is computed twice. */
typedef struct hss_midpoints {
    int32_t        n;       /* number of midpoints                                  */
    const int32_t *begin;   /* [n] device                                           */
    const int32_t *count;   /* [n] device                                           */
    const float   *m0;      /* [n] device                                           */
    const int32_t *index;   /* device or NULL                                       */
    float          h0;      /* OCT reference half offset (Traveltime::getReferenceHalfoffset) */
    float          apm;     /* midpoint aperture (Gather::getApm)                   */
    const int32_t *order;   /* [n] device or NULL: launch order (see above)         */
} hss_midpoints;

/* ------------------------------------------------------------------------------ */
int hss_abi_version(void);
const char *hss_status_string(int status);
/* Number of HIP devices (CudaSingleHostRunner::getNumOfDevices,
 * cuda/src/execution/CudaSingleHostRunner.cu:18-29). */
int hss_device_count(int *count);

/* Diagnostics (no counterpart in the reference): the number of float inputs, over all
 * 2^32 bit patterns, on which the kernels' short correctly-rounded sequences differ from
 * the library's correctly rounded operations on the current device: sqrt for every
 * non-negative input, and t/dt from RN(1/dt) with one residual correction wherever either
 * quotient lies in [1, 2e9) -- the only quotients a window with tau >= 1 can accept.  The
 * semblance entry points use the short sequences for a dt only when this is 0 and
 * tau >= 1, and the library operations otherwise; results are the same bits either way.
 * Synchronous on `stream`. */
int hss_verify_index_ops(float dt, unsigned long long *mismatches, void *stream);

/* Diagnostics (no counterpart in the reference): the state of that check for dt on the
 * current device -- *state = 1 when it passed (the kernels use the short sequences),
 * 0 when it found a mismatch (they use the library operations), -1 when it has not run
 * to completion there yet.  A semblance call on a stream under capture does not run the
 * check (it cannot wait): it uses the library operations and leaves the state at -1, so
 * the first call outside capture runs it. */
int hss_index_ops_state(float dt, int *state);

/* Grid candidates in storage form: CMP {c}, ZO-CRS {c,A,B}, OCT {c,slope}, c = 4/v^2,
 * v-major (ZO-CRS: B fastest; OCT: slope fastest); value d of axis k is
 * lower[k] + d*step[k].  Writes prod(count) * np floats to d_params.
 * Replaces buildParameterArrayFor{CommonMidPoint,ZeroOffsetCommonReflectionSurface,
 * OffsetContinuationTrajectory} launched by CudaLinearSearchAlgorithm::initializeParameters
 * (cuda/src/semblance/algorithm/CudaLinearSearchAlgorithm.cu:165-223). */
int hss_build_parameter_array(int model, const float *lower, const float *step,
                              const int32_t *count, float *d_params, void *stream);

/* Grid ("linear") search semblance for a batch of midpoints, with the best-candidate
 * pick fused in.  For midpoint g, sample s and candidate p it evaluates the semblance
 * and stack of the reference's computeSemblancesFor<Model> and keeps the first
 * strict maximum over p as selectBestSemblancesFor<Model> does.
 *   d_params  [P][np] storage form (hss_build_parameter_array);
 *   d_result  [n][(2+np)][ns]: S, stack, v = 2/sqrt(c), then A,B / slope;
 *   d_sembl, d_stack [n][ns][P] every evaluated value (NULL to skip; ns * P must stay
 *             below 2^32, else HSS_ERR_INVALID_SIZE);
 *   d_not_used [n] OCT: number of (s,p,trace) triples the model rejected
 *             (NULL to skip; accumulated, caller zeroes it).
 * Replaces the launches in CudaLinearSearchAlgorithm::computeSemblanceAtGpuForMidpoint
 * (cuda/src/semblance/algorithm/CudaLinearSearchAlgorithm.cu:35-163). */
int hss_grid_semblance(const hss_traces *traces, const hss_midpoints *mids, int model,
                       const float *d_params, int32_t P, int variant,
                       float *d_result, float *d_sembl, float *d_stack,
                       unsigned long long *d_not_used, void *stream);

/* The plain per-gather grid search, kept as the baseline the batched kernel is measured
 * against: the reference's kernel structure -- per midpoint one launch with one thread per
 * (t0, candidate) writing d_sembl/d_stack [ns][P] (scratch, reused per midpoint), then one
 * thread per t0 picking the first strict maximum into d_result [n][(2+np)][ns].
 * FP32 storage, HSS_EXACT arithmetic (same bits as hss_grid_semblance with HSS_EXACT).
 * h_begin, h_count, h_m0 [n] are HOST arrays (the midpoint loop runs on the host).
 * Restates cmp/linear_search.cu:19-115 (zocrs/, oct/ twins) as launched per CDP by
 * CudaLinearSearchAlgorithm::computeSemblanceAtGpuForMidpoint
 * (cuda/src/semblance/algorithm/CudaLinearSearchAlgorithm.cu:35-163). */
int hss_grid_semblance_plain(const hss_traces *traces, int model, const float *d_params,
                             int32_t P, int32_t n, const int32_t *h_begin,
                             const int32_t *h_count, const float *h_m0, float h0, float apm,
                             float *d_sembl, float *d_stack, float *d_result, void *stream);

/* The launch shape the grid-search autotuner chose on the current device for (model, w,
 * variant, storage): *cpl = candidates per lane (1 or 2), *t0_tiles = 64-sample t0 tiles
 * per workgroup (1, 2 or 4; the workgroup's four waves split into that many tiles times
 * 4 / t0_tiles candidate groups), *staged = 1 when the workgroup stages each trace's span
 * in LDS for all its waves (CMP, one t0 tile per workgroup); all 0 when that key has not
 * been tuned yet.  No
 * counterpart in the reference.  hss_grid_semblance tunes on its first launch of at
 * least 4096 workgroups for a key (outside stream capture) by timing each shape (six
 * unstaged, and for CMP two staged) on the first launch slots of the caller's batch
 * (about 8192 workgroups) and keeping the fastest, a shape listed earlier (two candidates
 * per lane, one t0 tile, unstaged first) winning ties within 1%; the environment variable
 * HSS_AUTOTUNE=0 turns it off.  An untuned launch runs two candidates per lane, one t0
 * tile, staged for CMP (HSS_KAHAN: one candidate, unstaged).  Results do not depend on the
 * shape (bit for bit). */
int hss_grid_tuned_shape(int model, int w, int variant, int storage, int32_t *cpl,
                         int32_t *t0_tiles, int32_t *staged);
/* hss_grid_tuned_shape's candidates per lane only. */
int hss_grid_tuned_config(int model, int w, int variant, int storage, int32_t *cpl);

/* Pins the grid-search launch shape for (model, w, variant, storage) on the current
 * device: cpl = 1 or 2 candidates per lane, t0_tiles = 1, 2 or 4, staged = 0 or 1 (1 only
 * for CMP with t0_tiles = 1: HSS_ERR_INVALID_MODEL / HSS_ERR_INVALID_SIZE otherwise);
 * cpl = 0 forgets the key (the next large launch tunes it again).
 * hss_grid_set_config(.., cpl) pins (cpl, 1, 0). */
int hss_grid_set_shape(int model, int w, int variant, int storage, int32_t cpl,
                       int32_t t0_tiles, int32_t staged);
int hss_grid_set_config(int model, int w, int variant, int storage, int32_t cpl);

/* Bytes of device workspace hss_de_semblance needs for n_mid midpoints of ns samples
 * with populations of NP individuals (two generations of every population). */
int hss_de_workspace_size(int32_t n_mid, int32_t ns, int model, int32_t NP, uint64_t *bytes);

/* Differential evolution of the traveltime parameters for a batch of midpoints: for every
 * (midpoint, t0) an independent population of NP individuals in v space, bounds
 * lower/upper[np] (host arrays: velocity, then A,B or slope), G generations in total
 * (the start population counts as the first), F = 0.85, CR = 0.5, selection on strict '>'
 * of the semblance.  Uniforms come from Philox4x32-10 with key `seed`; midpoint g uses
 * stream stream_base + g, so a batch and per-midpoint calls with matching stream ids
 * give identical results.
 *   gen_begin, gen_end: the generations this call runs, 0 <= gen_begin < gen_end <= G
 *             ([0, G) is the whole run).  The populations stay in d_workspace between
 *             calls, so consecutive ranges give the same result as one call; the outputs
 *             below are written by the call that reaches gen_end = G;
 *   d_result  [n][(2+np)][ns]: S, stack, then the best individual's parameters (v space);
 *   d_pop_x   [n][ns][NP][np], d_pop_fx [n][ns][NP][2]: the final populations (NULL to skip);
 *   d_not_used [n] OCT: (s, individual, trace) triples the model rejected, summed over the
 *             generations run (NULL to skip; accumulated, caller zeroes it).
 * Replaces setupRandomSeed/startPopulations/mutatePopulations/crossoverPopulations/
 * nextGeneration/computeSemblancesFor<Model>/selectBestIndividualsFor<Model>, launched one
 * by one with a device sync each by CudaDifferentialEvolutionAlgorithm
 * (cuda/src/semblance/algorithm/CudaDifferentialEvolutionAlgorithm.cu:35-353) under the
 * generation loop of DifferentialEvolutionAlgorithm::computeSemblanceAndParametersForMidpoint
 * (common/src/semblance/algorithm/DifferentialEvolutionAlgorithm.cpp:21-81). */
int hss_de_semblance(const hss_traces *traces, const hss_midpoints *mids, int model,
                     const float *lower, const float *upper, int32_t NP, int32_t G,
                     int32_t gen_begin, int32_t gen_end, uint64_t seed, uint32_t stream_base,
                     int variant, void *d_workspace, uint64_t workspace_bytes,
                     float *d_result, float *d_pop_x, float *d_pop_fx,
                     unsigned long long *d_not_used, void *stream);

/* Size of the static device scratch array of hss_de_semblance_noalloc (64 MiB; one copy
 * per device that calls it, held until the process ends). */
#define HSS_DE_SCRATCH_BYTES (64ull << 20)

/* hss_de_semblance without a caller workspace (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  The populations live in a file-scope device
 * array of HSS_DE_SCRATCH_BYTES; the call allocates and frees nothing.  A batch whose
 * hss_de_workspace_size exceeds HSS_DE_SCRATCH_BYTES is refused with
 * HSS_ERR_SCRATCH_LIMIT before anything runs (split the midpoints into smaller batches).
 * Arguments and results are those of hss_de_semblance, bit for bit.  Every call on a
 * device shares that device's array and it keeps the last call's populations (a run split
 * over generation ranges resumes from them when no other call intervenes); calls that
 * overlap on one device, on any streams, race. */
int hss_de_semblance_noalloc(const hss_traces *traces, const hss_midpoints *mids, int model,
                             const float *lower, const float *upper, int32_t NP, int32_t G,
                             int32_t gen_begin, int32_t gen_end, uint64_t seed,
                             uint32_t stream_base, int variant, float *d_result, float *d_pop_x,
                             float *d_pop_fx, unsigned long long *d_not_used, void *stream);

/* Stretch-free re-stack for a batch of midpoints.  d_vparams [n][np][ns] holds, per
 * sample, the parameters picked by an earlier search (velocity in v space, then A,B or
 * slope: the per-CDP traces of the --stack-datafiles files); nvals[nN] (host, nN <= 64)
 * the sample shifts (the reference uses -10..10).
 *   d_result [n][3][ns]: S, stack, n;  d_sembl, d_stack [n][ns][nN] (NULL to skip);
 *   d_not_used [n] OCT rejections, accumulated (NULL to skip).
 * Replaces computeSemblancesFor<Model> (stretch_free) + selectBestSemblances launched by
 * CudaStretchFreeAlgorithm::computeSemblanceAtGpuForMidpoint
 * (cuda/src/semblance/algorithm/CudaStretchFreeAlgorithm.cu:28-126). */
int hss_stretch_free(const hss_traces *traces, const hss_midpoints *mids, int model,
                     const float *d_vparams, const float *nvals, int32_t nN, int variant,
                     float *d_result, float *d_sembl, float *d_stack,
                     unsigned long long *d_not_used, void *stream);

/* Semblance gradient (no counterpart in the reference): for every (midpoint g, sample s)
 * of a batch, S, stack and dS/d(parameters) at the point held in rows 2.. of d_points
 * [n][(2+np)][ns] (v space: velocity, then A,B or slope -- the layout of the d_result of
 * hss_grid_semblance / hss_de_semblance, so a search result can be passed directly).
 * Writes d_out [n][(2+np)][ns]: S, stack, dS/dv, then dS/dA, dS/dB or dS/dslope.
 * FMA arithmetic (as HSS_FAST).  The gradient is exact for the piecewise-smooth S away
 * from the seams where a window slides by one sample.  Points with v <= 0 give zeros. */
int hss_semblance_gradient(const hss_traces *traces, const hss_midpoints *mids, int model,
                           const float *d_points, float *d_out, void *stream);

/* Gauss-Newton polish (no counterpart in the reference): starting from rows 2.. of
 * d_start [n][(2+np)][ns] (e.g. a grid or DE result), `steps` Levenberg-Marquardt-damped
 * Gauss-Newton steps per (midpoint, sample) on the residual 1 - S in box-normalised
 * coordinates, inside [lower, upper] (host arrays, v space).  A step is kept only when S
 * increases, so S never decreases.  Writes d_result [n][(2+np)][ns] (S, stack,
 * parameters), which must not alias d_start.  Each step costs one gradient evaluation. */
int hss_polish(const hss_traces *traces, const hss_midpoints *mids, int model,
               const float *lower, const float *upper, int32_t steps,
               const float *d_start, float *d_result, void *stream);

/* Vector-Jacobian product of the semblance (no counterpart in the reference): with
 * d_grad_rows the d_out of hss_semblance_gradient ([n][(2+np)][ns]) and d_grad_S [n][ns]
 * the gradient of some loss with respect to S, writes
 * d_grad_points[g][m][s] = d_grad_S[g][s] * dS/dp_m [g][s]  ([n][np][ns]) -- the
 * backward of a torch.autograd.Function-style op whose forward is S(points). */
int hss_semblance_vjp(int32_t n, int32_t np, int32_t ns, const float *d_grad_S,
                      const float *d_grad_rows, float *d_grad_points, void *stream);

/* OCT trace selection: which candidates are tried at every t0. */
enum hss_oct_mode {
    HSS_OCT_PAIRS = 0,       /* d_pairs[npairs][2] {c, slope}, the same for every t0 (grid) */
    HSS_OCT_PER_SAMPLE = 1,  /* d_vparams[n][2][ns] {velocity, slope} of each t0 (stretch-free) */
    HSS_OCT_DE_START = 2     /* the start population of hss_de_semblance with the same
                                NP, lower, upper, seed and stream_base */
};

typedef struct hss_oct_candidates {
    int32_t mode;
    const float *d_pairs;
    int32_t npairs;
    const float *d_vparams;
    int32_t NP;
    float lower[2];
    float upper[2];
    uint64_t seed;
    uint32_t stream_base;
} hss_oct_candidates;

/* Trace selection masks over the whole line for n midpoints: d_mask[n][ntr] = 1 when the
 * trace is used for that midpoint.  CMP: m == m0; ZO-CRS: |m0 - m| <= apm; OCT: some t0
 * and candidate give a displaced midpoint within apm (`oct` says which candidates).
 * Replaces selectTracesFor<Model>[AndDifferentialEvolution] launched by the
 * selectTracesToBeUsedForMidpoint overrides (CudaLinearSearchAlgorithm.cu:225-324,
 * CudaDifferentialEvolutionAlgorithm.cu:112-183, CudaStretchFreeAlgorithm.cu:128-200). */
int hss_select_traces(const hss_traces *traces, int model, int32_t n, const float *d_m0,
                      float apm, float h0, const hss_oct_candidates *oct, uint8_t *d_mask,
                      void *stream);

#ifdef __cplusplus
}
#endif
#endif /* HSS_H */
