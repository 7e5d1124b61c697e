// Instantiates the adjoint (gradient and polish) kernels for the OCT traveltime.
#include "adjoint.hpp"

namespace hss {
int launch_adjoint_oct(const AdjointArgs &a, int n_mid, int storage, bool polish, hipStream_t st) {
    return launch_adjoint<HSS_OCT>(a, n_mid, storage, polish, st);
}
}  // namespace hss
