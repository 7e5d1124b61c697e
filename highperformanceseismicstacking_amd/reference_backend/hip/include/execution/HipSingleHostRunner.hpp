// The runner of the HIP backend's programs (replaces CudaSingleHostRunner,
// cuda/src/execution/CudaSingleHostRunner.cu:14-29): the reference's SingleHostRunner with
// the HIP algorithm and device-context builders injected.
#pragma once

#include "common/include/execution/SingleHostRunner.hpp"
#include "common/include/parser/Parser.hpp"

class HipSingleHostRunner : public SingleHostRunner {
    public:
        explicit HipSingleHostRunner(Parser* parser);
        unsigned int getNumOfDevices() const override;
};
