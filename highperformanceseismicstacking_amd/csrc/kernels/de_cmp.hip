// Differential-evolution instantiations for the CMP traveltime (see de_impl.hpp).
#include "de_impl.hpp"

namespace hss {
int launch_de_cmp(const DeArgs &a, int n_mid, int variant, int storage, hipStream_t stream) {
    return launch_de<HSS_CMP>(a, n_mid, variant, storage, stream);
}
}  // namespace hss
