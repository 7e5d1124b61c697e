// Instantiates the adjoint (gradient and polish) kernels for the ZOCRS traveltime.
#include "adjoint.hpp"

namespace hss {
int launch_adjoint_zocrs(const AdjointArgs &a, int n_mid, int storage, bool polish, hipStream_t st) {
    return launch_adjoint<HSS_ZOCRS>(a, n_mid, storage, polish, st);
}
}  // namespace hss
