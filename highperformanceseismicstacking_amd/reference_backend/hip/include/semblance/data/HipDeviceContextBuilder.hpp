#pragma once

#include "common/include/semblance/data/DeviceContextBuilder.hpp"

#include <memory>

class HipDeviceContextBuilder : public DeviceContextBuilder {
    public:
        static DeviceContextBuilder* getInstance();
        DeviceContext* build(unsigned int devId) override;
};
