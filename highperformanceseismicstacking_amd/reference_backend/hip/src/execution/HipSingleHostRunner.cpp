#include "hip/include/execution/HipSingleHostRunner.hpp"
#include "hip/include/execution/HipUtils.hpp"
#include "hip/include/semblance/algorithm/HipComputeAlgorithmBuilder.hpp"
#include "hip/include/semblance/data/HipDeviceContextBuilder.hpp"

HipSingleHostRunner::HipSingleHostRunner(Parser* parser)
    : SingleHostRunner(parser, HipComputeAlgorithmBuilder::getInstance(),
                       HipDeviceContextBuilder::getInstance()) {
}

unsigned int HipSingleHostRunner::getNumOfDevices() const {
    int count = 0;
    const int status = hss_device_count(&count);
    if (status == HSS_ERR_NO_DEVICE || count == 0) {
        throw runtime_error("no HIP device is visible to this process");
    }
    HSS_CHECK(status);
    return static_cast<unsigned int>(count);
}
