// Differential-evolution instantiations for the OCT traveltime (see de_impl.hpp).
#include "de_impl.hpp"

namespace hss {
int launch_de_oct(const DeArgs &a, int n_mid, int variant, int storage, hipStream_t stream) {
    return launch_de<HSS_OCT>(a, n_mid, variant, storage, stream);
}
}  // namespace hss
