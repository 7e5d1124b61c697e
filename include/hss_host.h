/*
 * hss_host.h -- C API of libhss_host.so, the C++17 host layer (SU reader, traveltime
 * models, command line) of the MI355X semblance backend, for bindings and tests that
 * need the host logic without a GPU.  Only the calls in the last section touch a device.
 *
 * Every function returns 0 on success, -1 when the host layer raised
 * std::invalid_argument, -2 for std::logic_error, -3 for std::runtime_error and -4 for
 * any other exception; hss_host_last_error() gives the message.
 */
#ifndef HSS_HOST_H
#define HSS_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *hss_host_last_error(void);

/* Reads an SU file with Gather::readGatherFromFile semantics (common/src/model/Gather.cpp:
 * 49-126): lowest/highest may be NULL for the defaults (FLT_MIN, FLT_MAX).  Reports the
 * scalars; hss_host_gather_arrays copies the kept traces (sorted by midpoint, then half
 * offset) and the CDP map (ascending midpoint, with each CDP's kept-trace count). */
int hss_host_read_gather(const char *path, float aph, float apm, float tau, float azimuth_deg,
                         const float *lowest, const float *highest, int32_t *ns, float *dt,
                         int32_t *tau_index, int32_t *w, int32_t *ntr, int32_t *ncdp);
int hss_host_gather_arrays(float *mid, float *hoff, float *samples, float *cdp_m0,
                           int32_t *cdp_traces);

/* Parameter bounds TraveltimeBuilder derives (common/src/traveltime/TraveltimeBuilder.cpp:
 * 23-101): model 0 CMP, 1 ZO-CRS (only lower/upper[0] are read; v0, ain, bpctg used),
 * 2 OCT; out_lower/out_upper receive np values. */
int hss_host_traveltime_bounds(int model, const float *lower, const float *upper, int32_t n,
                               float v0, float ain, float bpctg, float *out_lower,
                               float *out_upper);

/* Writes a result set with the output layer of the programs (Dumper, the reference's
 * common/src/output/Dumper.cpp:13-150 re-implemented) for the gather last read by
 * hss_host_read_gather: the directory <out_path>/<data_name>_<method>_<traveltime word>_
 * <YYYYMMDD_HHMM>/ with one SU file per result (CDP header + ns float32 per CDP, in
 * midpoint order), one CSV per statistic, gather.txt (input_path as given) and
 * traveltime.txt -- the files SingleHostRunner::main writes after a run.  The traveltime is
 * built as TraveltimeBuilder builds it (model 0 CMP, 1 ZO-CRS with v0/ain/bpctg, 2 OCT
 * with h0).  values [ncdp][results][ns] and statistics [ncdp][5] follow the CDP map order.
 * The directory path goes to dir (NUL-terminated, truncated to dir_capacity). */
int hss_host_dump_results(const char *out_path, const char *input_path, const char *data_name,
                          const char *method, int model, const float *lower, const float *upper,
                          int32_t nbounds, float v0, float ain, float bpctg, float h0,
                          const float *values, const float *statistics, char *dir,
                          int32_t dir_capacity);

/* Parses a command line of program 0 (single_host_linear_search), 1 (single_host_de) or
 * 2 (single_host_stretch_free) and builds its traveltime, without reading data or
 * opening a device. */
int hss_host_parse_command_line(int program, int argc, const char **argv);

/* The launch order (hss_midpoints.order) the host layer gives a batch of n midpoints
 * with the given work (selected-trace counts): sorted by work and dealt to the chip's 8
 * XCDs in snake order, so that each XCD's eighth of the launch carries the same work,
 * largest first.  Writes a permutation of 0..n-1 to order[n].  No counterpart in the
 * reference (it launches one midpoint at a time). */
int hss_host_xcd_balanced_order(const int32_t *work, int32_t n, int32_t *order);

/* ---- per-midpoint tasks --------------------------------------------------------------
 * The reference's SPITS module as C calls (job manager: common/src/execution/
 * SpitzJobManager.cpp:17-42, worker: SpitzWorker.cpp:10-40, committer: SpitzCommitter.cpp:
 * 40-99), without the SPITS runtime.  A job is one command line of program 0
 * (single_host_linear_search), 1 (single_host_de) or 2 (single_host_stretch_free): open
 * parses it, reads the line and creates the output directory.  next_task hands out one
 * task per CDP in ascending midpoint order (*length 0 when none is left); a task record is
 * one float32 midpoint.  A worker (one device; hss_host_worker_* are the only calls here
 * that need a GPU) turns a task into a result record: the midpoint, the midpoint's
 * result arrays (results x ns float32, as the traveltime lists them) and its five
 * statistics (float32, StatisticResult order).  commit_task stores a record (any order,
 * any worker); commit writes the files the programs write and returns the directory.
 * One job per process at a time (the line is process-wide, as the reference's Gather).
 * record_sizes reports the two record sizes in bytes and the number of tasks. */
typedef struct hss_host_job hss_host_job;
typedef struct hss_host_worker hss_host_worker;
int hss_host_job_open(int program, int argc, const char **argv, hss_host_job **job);
int hss_host_job_record_sizes(const hss_host_job *job, int32_t *task_bytes,
                              int32_t *result_bytes, int32_t *tasks);
int hss_host_job_next_task(hss_host_job *job, void *task, int32_t capacity, int32_t *length);
int hss_host_job_commit_task(hss_host_job *job, const void *result, int32_t length);
int hss_host_job_commit(hss_host_job *job, char *dir, int32_t dir_capacity);
int hss_host_job_close(hss_host_job *job);
int hss_host_worker_open(hss_host_job *job, int32_t device, hss_host_worker **worker);
int hss_host_worker_run(hss_host_worker *worker, const void *task, int32_t length,
                        void *result, int32_t capacity, int32_t *result_length);
int hss_host_worker_close(hss_host_worker *worker);

/* ---- device calls (need a GPU) ------------------------------------------------------
 * SemblanceOp (host/include/semblance_op.hpp), the torch.autograd.Function-shaped C++ op
 * over the adjoint: forward(d_points [n][np][ns]) -> d_S [n][ns], then
 * backward(d_gradS [n][ns]) -> d_gradPoints [n][np][ns].  Synchronous.  No counterpart
 * in the reference. */
typedef struct hss_traces hss_traces;
typedef struct hss_midpoints hss_midpoints;
int hss_host_semblance_op(const hss_traces *traces, const hss_midpoints *mids, int model,
                          const float *d_points, const float *d_gradS, float *d_S,
                          float *d_gradPoints, void *stream);

#ifdef __cplusplus
}
#endif
#endif
