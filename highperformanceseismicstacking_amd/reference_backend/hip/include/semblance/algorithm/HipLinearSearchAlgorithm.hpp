// Grid search on the HIP backend (replaces CudaLinearSearchAlgorithm,
// cuda/src/semblance/algorithm/CudaLinearSearchAlgorithm.cu): the parameter grid from
// hss_build_parameter_array, the trace selection as the reference selects (host mask for
// CMP / ZO-CRS, hss_select_traces for OCT over the reference's 1024 random samples), the
// selected traces copied by the reference's own copyOnlySelectedTracesToDevice, and one
// hss_grid_semblance call per midpoint with the best-candidate pick fused in.
#pragma once

#include "common/include/semblance/algorithm/LinearSearchAlgorithm.hpp"
#include "hip/include/semblance/algorithm/HipMidpointBatch.hpp"

class HipLinearSearchAlgorithm : public LinearSearchAlgorithm {
    public:
        HipLinearSearchAlgorithm(
            shared_ptr<Traveltime> traveltime,
            shared_ptr<DeviceContext> context,
            DataContainerBuilder* dataBuilder,
            unsigned int threadCount
        );

        void computeSemblanceAtGpuForMidpoint(float m0) override;
        void initializeParameters() override;
        void selectTracesToBeUsedForMidpoint(float m0) override;

    private:
        HipMidpointBatch batch;
        vector<float> octSamples;   // {c, slope} pairs of the OCT selection
};
