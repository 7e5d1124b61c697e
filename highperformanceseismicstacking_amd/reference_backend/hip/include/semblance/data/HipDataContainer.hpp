// DataContainer of the HIP backend (common/include/semblance/data/DataContainer.hpp): a
// float array in device memory; sizes in elements; copyFrom / pasteTo refuse a size
// mismatch with std::invalid_argument (CudaDataContainer.cu:31-59).
#pragma once

#include "common/include/semblance/data/DataContainer.hpp"

class HipDataContainer : public DataContainer {
    public:
        HipDataContainer(unsigned int elementCount, shared_ptr<DeviceContext> context);
        ~HipDataContainer() override;

        float* getAddress() const { return address; }

        void allocate() override;
        void deallocate() override;
        void copyFrom(const vector<float>& sourceArray) override;
        void copyFromWithOffset(const vector<float>& sourceArray, unsigned int offset) override;
        void pasteTo(vector<float>& targetArray) override;
        void reset() override;

    private:
        float* address = nullptr;
};
