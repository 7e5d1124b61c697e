#include "semblance_op.hpp"

#include <stdexcept>

SemblanceOp::SemblanceOp(const hss_traces &traces, const hss_midpoints &mids, int model,
                         hipStream_t stream)
    : traces_(traces), mids_(mids), model_(model), ns_(traces.ns), stream_(stream) {
    np_ = model == HSS_CMP ? 1 : (model == HSS_ZOCRS ? 3 : (model == HSS_OCT ? 2 : -1));
    if (np_ < 0) throw std::invalid_argument("SemblanceOp: unknown traveltime model");
    const size_t bytes = static_cast<size_t>(mids.n) * (2 + np_) * ns_ * sizeof(float);
    rows_.resize(bytes);
    saved_.resize(bytes);
}

void SemblanceOp::forward(const float *d_points, float *d_S) {
    const size_t plane = static_cast<size_t>(ns_) * sizeof(float);
    // rows 2.. of each midpoint's block are the points; rows 0 and 1 are not read
    for (int g = 0; g < mids_.n; g++) {
        char *dst = static_cast<char *>(rows_.get()) + (static_cast<size_t>(g) * (2 + np_) + 2) * plane;
        const char *src = reinterpret_cast<const char *>(d_points) + static_cast<size_t>(g) * np_ * plane;
        HIP_ASSERT(hipMemcpyAsync(dst, src, np_ * plane, hipMemcpyDeviceToDevice, stream_));
    }
    HSS_ASSERT(hss_semblance_gradient(&traces_, &mids_, model_, rows_.as<float>(), saved_.as<float>(),
                                      stream_));
    for (int g = 0; g < mids_.n; g++) {
        const char *src = static_cast<const char *>(saved_.get()) + static_cast<size_t>(g) * (2 + np_) * plane;
        HIP_ASSERT(hipMemcpyAsync(reinterpret_cast<char *>(d_S) + static_cast<size_t>(g) * plane, src,
                                  plane, hipMemcpyDeviceToDevice, stream_));
    }
    haveForward_ = true;
}

void SemblanceOp::backward(const float *d_gradS, float *d_gradPoints) {
    if (!haveForward_) throw std::logic_error("SemblanceOp::backward before forward");
    HSS_ASSERT(hss_semblance_vjp(mids_.n, np_, ns_, d_gradS, saved_.as<float>(), d_gradPoints,
                                 stream_));
}
