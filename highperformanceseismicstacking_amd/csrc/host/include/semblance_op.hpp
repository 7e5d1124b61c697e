// A differentiable semblance op with the shape of a torch.autograd.Function: forward maps
// the parameter points of a midpoint batch to their semblances and keeps what backward
// needs; backward maps dL/dS to dL/dpoints.  Both run on the device through the C ABI
// (hss_semblance_gradient computes S and dS/dpoints in one launch; hss_semblance_vjp
// applies the chain rule).  No counterpart in the reference.
#pragma once

#include <hip/hip_runtime_api.h>

#include "data.hpp"
#include "hss.h"

class SemblanceOp {
  public:
    // traces/mids/model describe the batch; the structs must outlive the op's calls.
    SemblanceOp(const hss_traces &traces, const hss_midpoints &mids, int model, hipStream_t stream);

    // d_points [n][np][ns] (v space) -> d_S [n][ns]; saves dS/dpoints for backward.
    void forward(const float *d_points, float *d_S);

    // d_gradS [n][ns] -> d_gradPoints [n][np][ns] (after forward).
    void backward(const float *d_gradS, float *d_gradPoints);

    int parameters() const { return np_; }

  private:
    hss_traces traces_;
    hss_midpoints mids_;
    int model_, np_, ns_;
    hipStream_t stream_;
    DeviceBuffer rows_;   // [n][(2+np)][ns]: the points on the way in
    DeviceBuffer saved_;  // [n][(2+np)][ns]: S, stack, dS/dpoints from forward
    bool haveForward_ = false;
};
