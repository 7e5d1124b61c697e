#pragma once

#include "common/include/semblance/data/DataContainerBuilder.hpp"

class HipDataContainerBuilder : public DataContainerBuilder {
    public:
        static DataContainerBuilder* getInstance();
        DataContainer* build(unsigned int allocatedCount, shared_ptr<DeviceContext> deviceContext) override;
};
