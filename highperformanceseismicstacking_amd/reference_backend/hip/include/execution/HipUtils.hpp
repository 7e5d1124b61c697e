// Status checks of the HIP backend: a failing HIP call or hss entry point becomes a
// std::runtime_error, as CUDA_ASSERT does for the CUDA backend
// (cuda/src/execution/CudaUtils.cu:8-15), so SingleHostRunner::main keeps its exit codes.
#pragma once

#include <hip/hip_runtime.h>

#include <sstream>
#include <stdexcept>

#include "hss.h"

#define HIP_CHECK(expr) hipBackendCheckHip((expr), #expr, __FILE__, __LINE__)
#define HSS_CHECK(expr) hipBackendCheckHss((expr), #expr, __FILE__, __LINE__)

inline void hipBackendCheckHip(hipError_t e, const char *what, const char *file, int line) {
    if (e != hipSuccess) {
        std::ostringstream s;
        s << file << ":" << line << ": " << what << ": " << hipGetErrorString(e);
        throw std::runtime_error(s.str());
    }
}

inline void hipBackendCheckHss(int status, const char *what, const char *file, int line) {
    if (status != HSS_OK) {
        std::ostringstream s;
        s << file << ":" << line << ": " << what << ": " << hss_status_string(status);
        throw std::runtime_error(s.str());
    }
}

// The arithmetic variant the backend asks the library for: HSS_VARIANT=exact gives the
// reference's operation order (results equal an IEEE FP32 evaluation of the reference's
// kernels bit for bit), anything else the contracted sums (within 1e-5).
int hipBackendVariant();
