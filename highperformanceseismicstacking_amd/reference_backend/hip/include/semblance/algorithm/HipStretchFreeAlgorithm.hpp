// Stretch-free re-stack on the HIP backend (replaces CudaStretchFreeAlgorithm,
// cuda/src/semblance/algorithm/CudaStretchFreeAlgorithm.cu): one hss_stretch_free call per
// midpoint with the shifts -10..10 and the midpoint's parameter traces the reference read
// into nonStretchFreeParameters.
#pragma once

#include "common/include/semblance/algorithm/StretchFreeAlgorithm.hpp"
#include "hip/include/semblance/algorithm/HipMidpointBatch.hpp"

class HipStretchFreeAlgorithm : public StretchFreeAlgorithm {
    public:
        HipStretchFreeAlgorithm(
            shared_ptr<Traveltime> traveltime,
            shared_ptr<DeviceContext> context,
            DataContainerBuilder* dataBuilder,
            unsigned int threadCount,
            const vector<string>& files
        );

        void computeSemblanceAtGpuForMidpoint(float m0) override;
        void selectTracesToBeUsedForMidpoint(float m0) override;

    private:
        HipMidpointBatch batch;
};
