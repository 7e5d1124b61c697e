#include "hip/include/semblance/data/HipDataContainer.hpp"
#include "hip/include/execution/HipUtils.hpp"

#include <string>

HipDataContainer::HipDataContainer(unsigned int count, shared_ptr<DeviceContext> context)
    : DataContainer(count, context) {
    allocate();
}

HipDataContainer::~HipDataContainer() {
    if (address) (void)hipFree(address);
    address = nullptr;
}

void HipDataContainer::allocate() {
    address = nullptr;
    if (elementCount) {
        HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&address), elementCount * sizeof(float)));
        reset();
    }
}

void HipDataContainer::deallocate() {
    if (address) HIP_CHECK(hipFree(address));
    address = nullptr;
}

void HipDataContainer::copyFrom(const vector<float>& sourceArray) {
    if (sourceArray.size() != elementCount) {
        throw invalid_argument("source holds " + to_string(sourceArray.size()) +
                               " elements, the device array " + to_string(elementCount));
    }
    copyFromWithOffset(sourceArray, 0);
}

void HipDataContainer::copyFromWithOffset(const vector<float>& sourceArray, unsigned int offset) {
    if (static_cast<size_t>(offset) + sourceArray.size() > elementCount) {
        throw invalid_argument("copy past the end of the device array");
    }
    if (sourceArray.empty()) return;
    HIP_CHECK(hipMemcpy(address + offset, sourceArray.data(), sourceArray.size() * sizeof(float),
                        hipMemcpyHostToDevice));
}

void HipDataContainer::pasteTo(vector<float>& targetArray) {
    if (targetArray.size() != elementCount) {
        throw invalid_argument("target holds " + to_string(targetArray.size()) +
                               " elements, the device array " + to_string(elementCount));
    }
    if (!elementCount) return;
    HIP_CHECK(hipMemcpy(targetArray.data(), address, elementCount * sizeof(float),
                        hipMemcpyDeviceToHost));
}

void HipDataContainer::reset() {
    if (elementCount) HIP_CHECK(hipMemset(address, 0, elementCount * sizeof(float)));
}
