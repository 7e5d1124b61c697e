#include "hip/include/semblance/data/HipDeviceContextBuilder.hpp"
#include "hip/include/semblance/data/HipDeviceContext.hpp"

DeviceContextBuilder* HipDeviceContextBuilder::getInstance() {
    static HipDeviceContextBuilder builder;
    return &builder;
}

DeviceContext* HipDeviceContextBuilder::build(unsigned int devId) {
    return new HipDeviceContext(devId);
}
