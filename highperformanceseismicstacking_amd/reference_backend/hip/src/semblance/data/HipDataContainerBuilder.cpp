#include "hip/include/semblance/data/HipDataContainerBuilder.hpp"
#include "hip/include/semblance/data/HipDataContainer.hpp"

DataContainerBuilder* HipDataContainerBuilder::getInstance() {
    static HipDataContainerBuilder builder;
    return &builder;
}

DataContainer* HipDataContainerBuilder::build(unsigned int allocatedCount,
                                              shared_ptr<DeviceContext> deviceContext) {
    return new HipDataContainer(allocatedCount, deviceContext);
}
