// Instantiates the adjoint (gradient and polish) kernels for the CMP traveltime.
#include "adjoint.hpp"

namespace hss {
int launch_adjoint_cmp(const AdjointArgs &a, int n_mid, int storage, bool polish, hipStream_t st) {
    return launch_adjoint<HSS_CMP>(a, n_mid, storage, polish, st);
}
}  // namespace hss
