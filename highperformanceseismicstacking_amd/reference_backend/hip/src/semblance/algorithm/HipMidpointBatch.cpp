#include "hip/include/semblance/algorithm/HipMidpointBatch.hpp"
#include "hip/include/execution/HipUtils.hpp"
#include "hip/include/semblance/data/HipDataContainer.hpp"
#include "hip/include/semblance/data/HipDeviceContext.hpp"

hipStream_t hipBackendStream(const shared_ptr<DeviceContext>& context) {
    const auto* hip = dynamic_cast<const HipDeviceContext*>(context.get());
    if (!hip) throw logic_error("the HIP backend needs a HipDeviceContext");
    return hip->getStream();
}

float* hipBackendPointer(const unique_ptr<DataContainer>& c) {
    const auto* hip = dynamic_cast<const HipDataContainer*>(c.get());
    if (!hip) throw logic_error("the HIP backend needs HipDataContainer arrays");
    return hip->getAddress();
}

HipMidpointBatch::~HipMidpointBatch() {
    if (dBegin) (void)hipFree(dBegin);
    if (dCount) (void)hipFree(dCount);
    if (dM0) (void)hipFree(dM0);
    if (dNotUsed) (void)hipFree(dNotUsed);
    if (dMask) (void)hipFree(dMask);
}

vector<unsigned char> HipMidpointBatch::selectTraces(
    unordered_map<GatherData, unique_ptr<DataContainer>>& data, int model, float m0, float apm,
    float h0, const hss_oct_candidates* oct, hipStream_t stream) {
    ensure();
    const Gather* g = Gather::getInstance();
    const size_t n = g->getTotalTracesCount();
    vector<unsigned char> mask(n, 0);
    if (!n) return mask;
    if (maskBytes < n) {
        if (dMask) HIP_CHECK(hipFree(dMask));
        dMask = nullptr;
        HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&dMask), n));
        maskBytes = n;
    }
    hss_traces line{};
    line.samples = hipBackendPointer(data[GatherData::MDPNT]);   // not read by the selection
    line.mid = hipBackendPointer(data[GatherData::MDPNT]);
    line.hoff = hipBackendPointer(data[GatherData::HLFOFFST]);
    line.hoff2 = hipBackendPointer(data[GatherData::HLFOFFST_SQ]);
    line.ntr = static_cast<int32_t>(n);
    line.ns = static_cast<int32_t>(g->getSamplesPerTrace());
    line.dt = g->getSamplePeriodInSeconds();
    line.tau = g->getTauIndexDisplacement();
    line.w = static_cast<int32_t>(g->getWindowSize());
    line.storage = HSS_F32;
    HIP_CHECK(hipMemcpy(dM0, &m0, sizeof(m0), hipMemcpyHostToDevice));
    HSS_CHECK(hss_select_traces(&line, model, 1, dM0, apm, h0, oct, dMask, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    HIP_CHECK(hipMemcpy(mask.data(), dMask, n, hipMemcpyDeviceToHost));
    return mask;
}

void HipMidpointBatch::ensure() {
    if (dBegin) return;
    HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&dBegin), sizeof(int32_t)));
    HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&dCount), sizeof(int32_t)));
    HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&dM0), sizeof(float)));
    HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&dNotUsed), sizeof(unsigned long long)));
}

hss_traces HipMidpointBatch::traces(unordered_map<GatherData, unique_ptr<DataContainer>>& data,
                                    unsigned int count) const {
    const Gather* g = Gather::getInstance();
    hss_traces t{};
    t.samples = hipBackendPointer(data[GatherData::FILT_SAMPL]);
    t.scale = nullptr;
    t.mid = hipBackendPointer(data[GatherData::FILT_MDPNT]);
    t.hoff = hipBackendPointer(data[GatherData::FILT_HLFOFFST]);
    t.hoff2 = hipBackendPointer(data[GatherData::FILT_HLFOFFST_SQ]);
    t.ntr = static_cast<int32_t>(count);
    t.ns = static_cast<int32_t>(g->getSamplesPerTrace());
    t.dt = g->getSamplePeriodInSeconds();
    t.tau = g->getTauIndexDisplacement();
    t.w = static_cast<int32_t>(g->getWindowSize());
    t.storage = HSS_F32;
    return t;
}

hss_midpoints HipMidpointBatch::midpoint(float m0, unsigned int count, float h0, float apm,
                                         hipStream_t stream) {
    ensure();
    const int32_t begin = 0, n = static_cast<int32_t>(count);
    HIP_CHECK(hipMemcpy(dBegin, &begin, sizeof(begin), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(dCount, &n, sizeof(n), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(dM0, &m0, sizeof(m0), hipMemcpyHostToDevice));
    (void)stream;
    hss_midpoints m{};
    m.n = 1;
    m.begin = dBegin;
    m.count = dCount;
    m.m0 = dM0;
    m.index = nullptr;
    m.h0 = h0;
    m.apm = apm;
    m.order = nullptr;
    return m;
}

void HipMidpointBatch::reset(hipStream_t stream) {
    ensure();
    HIP_CHECK(hipMemsetAsync(dNotUsed, 0, sizeof(unsigned long long), stream));
}

void HipMidpointBatch::storeNotUsed(DataContainer* notUsedArray, hipStream_t stream) const {
    unsigned long long total = 0;
    HIP_CHECK(hipMemcpyAsync(&total, dNotUsed, sizeof(total), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    auto* hip = dynamic_cast<HipDataContainer*>(notUsedArray);
    if (!hip) throw logic_error("the HIP backend needs HipDataContainer arrays");
    if (!hip->getElementCount()) return;
    const float value = static_cast<float>(total);
    HIP_CHECK(hipMemcpy(hip->getAddress(), &value, sizeof(value), hipMemcpyHostToDevice));
}
