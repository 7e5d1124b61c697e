// HIP device context, float containers and raw buffers (see data.hpp).
#include "data.hpp"

#include <sstream>

#include "hss.h"
#include "logger.hpp"

void hipAssert(hipError_t code, const char *file, int line) {
    if (code != hipSuccess) {
        std::ostringstream s;
        s << "HIP error detected at " << file << "::" << line << " with message "
          << hipGetErrorString(code);
        LOGE(s.str());
        throw std::runtime_error(s.str());
    }
}

void hssAssert(int status, const char *file, int line) {
    if (status != HSS_OK) {
        std::ostringstream s;
        s << "hss error " << status << " at " << file << "::" << line << ": "
          << hss_status_string(status);
        LOGE(s.str());
        if (status < 0) throw std::invalid_argument(s.str());
        throw std::runtime_error(s.str());
    }
}

// ---------------------------------------------------------------------------------
HipDeviceContext::HipDeviceContext(unsigned int devId) : DeviceContext(devId) {
    activate();
    HIP_ASSERT(hipStreamCreate(&stream));
}

HipDeviceContext::~HipDeviceContext() {
    if (stream) {
        (void)hipSetDevice(static_cast<int>(deviceId));
        (void)hipStreamDestroy(stream);
    }
}

void HipDeviceContext::activate() const {
    LOGH("Activating GPU#" << deviceId);
    HIP_ASSERT(hipSetDevice(static_cast<int>(deviceId)));
}

void HipDeviceContext::synchronize() const { HIP_ASSERT(hipStreamSynchronize(stream)); }

DeviceContextBuilder *HipDeviceContextBuilder::getInstance() {
    static HipDeviceContextBuilder builder;
    return &builder;
}

// ---------------------------------------------------------------------------------
HipDataContainer::HipDataContainer(unsigned int count, std::shared_ptr<DeviceContext> context)
    : DataContainer(count, std::move(context)) {
    allocate();
}

// No exception may leave a destructor: a failing free is ignored here (deallocate()
// reports it when called directly).
HipDataContainer::~HipDataContainer() {
    if (address) (void)hipFree(address);
}

hipStream_t HipDataContainer::stream() const {
    auto *ctx = dynamic_cast<HipDeviceContext *>(deviceContext.get());
    return ctx ? ctx->getStream() : nullptr;
}

void HipDataContainer::allocate() {
    LOGH("Allocating data container for " << elementCount << " elements.");
    address = nullptr;
    if (elementCount) HIP_ASSERT(hipMalloc(reinterpret_cast<void **>(&address), elementCount * sizeof(float)));
    reset();
}

void HipDataContainer::deallocate() {
    if (address) {
        LOGH("Deallocating data container");
        HIP_ASSERT(hipFree(address));
        address = nullptr;
    }
}

void HipDataContainer::copyFrom(const std::vector<float> &source) {
    if (source.size() != elementCount) {
        throw std::invalid_argument("source holds " + std::to_string(source.size()) +
                                    " elements, the device buffer " + std::to_string(elementCount));
    }
    copyFromWithOffset(source, 0);
}

void HipDataContainer::copyFromWithOffset(const std::vector<float> &source, unsigned int offset) {
    if (static_cast<size_t>(offset) + source.size() > elementCount) {
        throw std::invalid_argument("Copy exceeds the allocated device array.");
    }
    if (source.empty()) return;
    HIP_ASSERT(hipMemcpyAsync(address + offset, source.data(), source.size() * sizeof(float),
                              hipMemcpyHostToDevice, stream()));
    HIP_ASSERT(hipStreamSynchronize(stream()));
}

void HipDataContainer::pasteTo(std::vector<float> &target) {
    if (target.size() != elementCount) {
        throw std::invalid_argument("target holds " + std::to_string(target.size()) +
                                    " elements, the device buffer " + std::to_string(elementCount));
    }
    if (!elementCount) return;
    HIP_ASSERT(hipMemcpyAsync(target.data(), address, elementCount * sizeof(float),
                              hipMemcpyDeviceToHost, stream()));
    HIP_ASSERT(hipStreamSynchronize(stream()));
}

void HipDataContainer::reset() {
    if (!elementCount) return;
    HIP_ASSERT(hipMemsetAsync(address, 0, elementCount * sizeof(float), stream()));
    HIP_ASSERT(hipStreamSynchronize(stream()));
}

DataContainerBuilder *HipDataContainerBuilder::getInstance() {
    static HipDataContainerBuilder builder;
    return &builder;
}

// ---------------------------------------------------------------------------------
DeviceBuffer::~DeviceBuffer() {
    if (ptr) (void)hipFree(ptr);
}

void DeviceBuffer::resize(size_t bytes) {
    if (bytes == bytes_ && ptr) return;
    if (ptr) HIP_ASSERT(hipFree(ptr));
    ptr = nullptr;
    bytes_ = bytes;
    if (bytes) HIP_ASSERT(hipMalloc(&ptr, bytes));
}

void DeviceBuffer::upload(const void *host, size_t bytes, hipStream_t stream) {
    if (bytes > bytes_) throw std::invalid_argument("upload larger than the device buffer");
    if (!bytes) return;
    HIP_ASSERT(hipMemcpyAsync(ptr, host, bytes, hipMemcpyHostToDevice, stream));
    HIP_ASSERT(hipStreamSynchronize(stream));
}

void DeviceBuffer::download(void *host, size_t bytes, hipStream_t stream) const {
    if (bytes > bytes_) throw std::invalid_argument("download larger than the device buffer");
    if (!bytes) return;
    HIP_ASSERT(hipMemcpyAsync(host, ptr, bytes, hipMemcpyDeviceToHost, stream));
    HIP_ASSERT(hipStreamSynchronize(stream));
}
