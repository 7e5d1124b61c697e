#include "hip/include/semblance/data/HipDeviceContext.hpp"
#include "hip/include/execution/HipUtils.hpp"

HipDeviceContext::HipDeviceContext(unsigned int devId) : DeviceContext(devId) {
    HIP_CHECK(hipSetDevice(static_cast<int>(devId)));
    HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
}

HipDeviceContext::~HipDeviceContext() {
    if (stream) (void)hipStreamDestroy(stream);
}

void HipDeviceContext::activate() const {
    HIP_CHECK(hipSetDevice(static_cast<int>(deviceId)));
}
