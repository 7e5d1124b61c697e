// The C-ABI view of one midpoint of the reference's ComputeAlgorithm: its FILT_* device
// arrays (the selected traces, copied by copyOnlySelectedTracesToDevice) as an
// hss_traces, and a one-midpoint hss_midpoints (begin 0, count = the selected traces).
#pragma once

#include "common/include/semblance/algorithm/ComputeAlgorithm.hpp"
#include "hss.h"

#include <hip/hip_runtime.h>

#include <unordered_map>

class HipMidpointBatch {
    public:
        ~HipMidpointBatch();
        // traces of the current midpoint (filtered arrays of the algorithm)
        hss_traces traces(unordered_map<GatherData, unique_ptr<DataContainer>>& data,
                          unsigned int count) const;
        // the one-midpoint batch for m0 (device arrays uploaded on `stream`)
        hss_midpoints midpoint(float m0, unsigned int count, float h0, float apm, hipStream_t stream);
        // the trace mask of one midpoint over the whole line, from the line's resident
        // MDPNT / HLFOFFST arrays (hss_select_traces reads the geometry only)
        vector<unsigned char> selectTraces(unordered_map<GatherData, unique_ptr<DataContainer>>& data,
                                           int model, float m0, float apm, float h0,
                                           const hss_oct_candidates* oct, hipStream_t stream);
        // the OCT rejection count of the last launch (device), zeroed by reset()
        unsigned long long* notUsed() const { return dNotUsed; }
        void reset(hipStream_t stream);
        // writes the rejection total into element 0 of the reference's notUsed array
        void storeNotUsed(DataContainer* notUsedArray, hipStream_t stream) const;

    private:
        void ensure();
        int32_t* dBegin = nullptr;
        int32_t* dCount = nullptr;
        float* dM0 = nullptr;
        unsigned long long* dNotUsed = nullptr;
        uint8_t* dMask = nullptr;
        size_t maskBytes = 0;
};

// The device address of a HipDataContainer.
float* hipBackendPointer(const unique_ptr<DataContainer>& container);

hipStream_t hipBackendStream(const shared_ptr<DeviceContext>& context);
