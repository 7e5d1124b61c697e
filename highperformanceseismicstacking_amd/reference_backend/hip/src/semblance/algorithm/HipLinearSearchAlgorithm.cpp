#include "hip/include/semblance/algorithm/HipLinearSearchAlgorithm.hpp"
#include "hip/include/execution/HipUtils.hpp"
#include "hip/include/semblance/data/HipDataContainer.hpp"

#include <random>

HipLinearSearchAlgorithm::HipLinearSearchAlgorithm(
    shared_ptr<Traveltime> traveltime,
    shared_ptr<DeviceContext> context,
    DataContainerBuilder* dataBuilder,
    unsigned int threadCount
) : LinearSearchAlgorithm(traveltime, context, dataBuilder, threadCount) {
}

// buildParameterArrayFor* (CudaLinearSearchAlgorithm.cu:165-223) as one library call.
void HipLinearSearchAlgorithm::initializeParameters() {
    hipStream_t stream = hipBackendStream(deviceContext);
    float lower[3] = {0.0f, 0.0f, 0.0f}, step[3] = {0.0f, 0.0f, 0.0f};
    int32_t count[3] = {1, 1, 1};
    for (unsigned int p = 0; p < traveltime->getNumberOfParameters(); p++) {
        lower[p] = traveltime->getLowerBoundForParameter(p);
        step[p] = discretizationStep[p];
        count[p] = static_cast<int32_t>(discretizationGranularity[p]);
    }
    HSS_CHECK(hss_build_parameter_array(traveltime->getModel(), lower, step, count,
                                        hipBackendPointer(deviceParameterArray), stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    if (traveltime->getModel() == OCT) {
        // the reference's OCT selection candidates (CudaLinearSearchAlgorithm.cu:268-292):
        // 1024 draws per parameter from a default-seeded default_random_engine, all
        // velocities first (stored as c = 4/v^2), then all slopes
        const unsigned int pop = 1024;
        octSamples.assign(2 * pop, 0.0f);
        default_random_engine engine;
        for (unsigned int p = 0; p < 2; p++) {
            uniform_real_distribution<float> draw(traveltime->getLowerBoundForParameter(p),
                                                  traveltime->getUpperBoundForParameter(p));
            for (unsigned int i = 0; i < pop; i++) {
                float x = draw(engine);
                if (p == 0) x = 4.0f / (x * x);
                octSamples[2 * i + p] = x;
            }
        }
    }
}

void HipLinearSearchAlgorithm::selectTracesToBeUsedForMidpoint(float m0) {
    hipStream_t stream = hipBackendStream(deviceContext);
    Gather* gather = Gather::getInstance();
    hss_oct_candidates oct{};
    unique_ptr<DataContainer> pairs;
    if (traveltime->getModel() == OCT) {
        pairs.reset(dataFactory->build(static_cast<unsigned int>(octSamples.size()), deviceContext));
        pairs->copyFrom(octSamples);
        oct.mode = HSS_OCT_PAIRS;
        oct.d_pairs = hipBackendPointer(pairs);
        oct.npairs = static_cast<int32_t>(octSamples.size() / 2);
    }
    const vector<unsigned char> mask = batch.selectTraces(
        deviceFilteredTracesDataMap, traveltime->getModel(), m0, gather->getApm(),
        traveltime->getReferenceHalfoffset(), traveltime->getModel() == OCT ? &oct : nullptr, stream);
    copyOnlySelectedTracesToDevice(mask);
}

void HipLinearSearchAlgorithm::computeSemblanceAtGpuForMidpoint(float m0) {
    // no selected trace: the results stay zero (CudaLinearSearchAlgorithm.cu:37-40)
    if (!filteredTracesCount) return;
    hipStream_t stream = hipBackendStream(deviceContext);
    Gather* gather = Gather::getInstance();
    batch.reset(stream);
    const hss_traces traces = batch.traces(deviceFilteredTracesDataMap, filteredTracesCount);
    const hss_midpoints mids = batch.midpoint(m0, filteredTracesCount,
                                              traveltime->getReferenceHalfoffset(),
                                              gather->getApm(), stream);
    HSS_CHECK(hss_grid_semblance(&traces, &mids, traveltime->getModel(),
                                 hipBackendPointer(deviceParameterArray),
                                 static_cast<int32_t>(getTotalNumberOfParameters()),
                                 hipBackendVariant(), hipBackendPointer(deviceResultArray),
                                 hipBackendPointer(commonResultDeviceArrayMap[SemblanceCommonResult::SEMBL]),
                                 hipBackendPointer(commonResultDeviceArrayMap[SemblanceCommonResult::STACK]),
                                 batch.notUsed(), stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    batch.storeNotUsed(deviceNotUsedCountArray.get(), stream);
}
