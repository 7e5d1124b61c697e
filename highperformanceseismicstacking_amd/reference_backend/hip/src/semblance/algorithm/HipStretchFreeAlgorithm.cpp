#include "hip/include/semblance/algorithm/HipStretchFreeAlgorithm.hpp"
#include "hip/include/execution/HipUtils.hpp"

HipStretchFreeAlgorithm::HipStretchFreeAlgorithm(
    shared_ptr<Traveltime> traveltime,
    shared_ptr<DeviceContext> context,
    DataContainerBuilder* dataBuilder,
    unsigned int threadCount,
    const vector<string>& files
) : StretchFreeAlgorithm(traveltime, context, dataBuilder, threadCount, files) {
}

void HipStretchFreeAlgorithm::selectTracesToBeUsedForMidpoint(float m0) {
    hipStream_t stream = hipBackendStream(deviceContext);
    Gather* gather = Gather::getInstance();
    hss_oct_candidates oct{};
    if (traveltime->getModel() == OCT) {
        // the parameters read for this midpoint, one pair per sample
        oct.mode = HSS_OCT_PER_SAMPLE;
        oct.d_vparams = hipBackendPointer(nonStretchFreeParameters[m0]);
    }
    const vector<unsigned char> mask = batch.selectTraces(
        deviceFilteredTracesDataMap, traveltime->getModel(), m0, gather->getApm(),
        traveltime->getReferenceHalfoffset(), traveltime->getModel() == OCT ? &oct : nullptr, stream);
    copyOnlySelectedTracesToDevice(mask);
}

void HipStretchFreeAlgorithm::computeSemblanceAtGpuForMidpoint(float m0) {
    if (!filteredTracesCount) return;
    hipStream_t stream = hipBackendStream(deviceContext);
    Gather* gather = Gather::getInstance();
    // the shifts of StretchFreeAlgorithm::setUp, -10 .. 10
    const unsigned int nN = getTotalNumberOfParameters();
    vector<float> shifts(nN);
    for (unsigned int i = 0; i < nN; i++) shifts[i] = static_cast<float>(static_cast<int>(i) - static_cast<int>(nN / 2));
    batch.reset(stream);
    const hss_traces traces = batch.traces(deviceFilteredTracesDataMap, filteredTracesCount);
    const hss_midpoints mids = batch.midpoint(m0, filteredTracesCount,
                                              traveltime->getReferenceHalfoffset(),
                                              gather->getApm(), stream);
    HSS_CHECK(hss_stretch_free(&traces, &mids, traveltime->getModel(),
                               hipBackendPointer(nonStretchFreeParameters[m0]), shifts.data(),
                               static_cast<int32_t>(nN), hipBackendVariant(),
                               hipBackendPointer(deviceResultArray),
                               hipBackendPointer(commonResultDeviceArrayMap[SemblanceCommonResult::SEMBL]),
                               hipBackendPointer(commonResultDeviceArrayMap[SemblanceCommonResult::STACK]),
                               batch.notUsed(), stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    batch.storeNotUsed(deviceNotUsedCountArray.get(), stream);
}
