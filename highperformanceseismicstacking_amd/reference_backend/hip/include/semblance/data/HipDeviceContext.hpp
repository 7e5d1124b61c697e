// DeviceContext of the HIP backend (common/include/semblance/data/DeviceContext.hpp):
// activate() makes the device current; every call of the backend goes to the context's
// own stream and waits for it before returning, as the reference's calls are synchronous.
#pragma once

#include <hip/hip_runtime.h>

#include "common/include/semblance/data/DeviceContext.hpp"

class HipDeviceContext : public DeviceContext {
    public:
        explicit HipDeviceContext(unsigned int devId);
        ~HipDeviceContext() override;
        void activate() const override;
        hipStream_t getStream() const { return stream; }

    private:
        hipStream_t stream = nullptr;
};
