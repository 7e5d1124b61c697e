#include "hip/include/execution/HipUtils.hpp"

#include <cstdlib>
#include <cstring>

int hipBackendVariant() {
    const char* v = std::getenv("HSS_VARIANT");
    return (v && std::strcmp(v, "exact") == 0) ? HSS_EXACT : HSS_FAST;
}
