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
