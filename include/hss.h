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
 *   - floating point in/out is float32; trace samples may be stored as float32,
 *     float16 or bfloat16 (hss_traces.storage) with a per-trace float32 scale.
 *
 * Model ids follow traveltime_t (common/include/capability.h:3-7).
 */
#ifndef HSS_H
#define HSS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HSS_ABI_VERSION 1

/* traveltime models (common/include/capability.h:3-7) */
enum hss_model { HSS_CMP = 0, HSS_ZOCRS = 1, HSS_OCT = 2 };

/* trace sample storage */
enum hss_storage { HSS_F32 = 0, HSS_F16 = 1, HSS_BF16 = 2 };

/* arithmetic variant of the window sums
 *   HSS_EXACT: the reference op order, no FMA contraction -- results are bit-identical
 *              to an IEEE FP32 evaluation of common/include/gpu/interface.h:26-55;
 *   HSS_FAST : FMA-contracted interpolation and sums, stack sum derived from the
 *              window components; within 1e-5 relative of HSS_EXACT. */
enum hss_variant { HSS_EXACT = 0, HSS_FAST = 1 };

/* argument errors (negative statuses) */
enum hss_status {
    HSS_OK = 0,
    HSS_ERR_INVALID_MODEL = -1,
    HSS_ERR_INVALID_WINDOW = -2,   /* w != 2*tau+1, or w > 32 (MAX_WINDOW_SIZE) */
    HSS_ERR_INVALID_SIZE = -3,     /* a count is negative, zero where it must not be, or too large */
    HSS_ERR_NULL_POINTER = -4,
    HSS_ERR_INVALID_STORAGE = -5,
    HSS_ERR_INVALID_VARIANT = -6,
    HSS_ERR_NO_DEVICE = -7
};

/* All traces of a line (or of one midpoint), resident on the device, in the order the
 * reference's Gather sorts them (by midpoint, then half offset; Trace.cpp:63-74).
 * The float arrays are the GatherData MDPNT / HLFOFFST / HLFOFFST_SQ columns
 * (ComputeAlgorithm.hpp:18-27, ComputeAlgorithm.cpp:31-61). */
typedef struct hss_traces {
    const void  *samples;   /* [ntr][ns] in `storage`                              */
    const float *scale;     /* [ntr] per-trace scale; NULL => 1 (required for F16) */
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
 * expressed as ranges into the resident line. */
typedef struct hss_midpoints {
    int32_t        n;       /* number of midpoints                                  */
    const int32_t *begin;   /* [n] device                                           */
    const int32_t *count;   /* [n] device                                           */
    const float   *m0;      /* [n] device                                           */
    const int32_t *index;   /* device or NULL                                       */
    float          h0;      /* OCT reference half offset (Traveltime::getReferenceHalfoffset) */
    float          apm;     /* midpoint aperture (Gather::getApm)                   */
} hss_midpoints;

/* ------------------------------------------------------------------------------ */
int hss_abi_version(void);
const char *hss_status_string(int status);
/* Number of HIP devices (CudaSingleHostRunner::getNumOfDevices,
 * cuda/src/execution/CudaSingleHostRunner.cu:18-29). */
int hss_device_count(int *count);

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
 *   d_sembl, d_stack [n][ns][P] every evaluated value (NULL to skip);
 *   d_not_used [n] OCT: number of (s,p,trace) triples the model rejected
 *             (NULL to skip; accumulated, caller zeroes it).
 * Replaces the launches in CudaLinearSearchAlgorithm::computeSemblanceAtGpuForMidpoint
 * (cuda/src/semblance/algorithm/CudaLinearSearchAlgorithm.cu:35-163). */
int hss_grid_semblance(const hss_traces *traces, const hss_midpoints *mids, int model,
                       const float *d_params, int32_t P, int variant,
                       float *d_result, float *d_sembl, float *d_stack,
                       unsigned long long *d_not_used, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* HSS_H */
