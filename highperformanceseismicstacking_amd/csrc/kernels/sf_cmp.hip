// Stretch-free instantiations for the CMP traveltime (see sf_impl.hpp).
#include "sf_impl.hpp"

namespace hss {
int launch_sf_cmp(const SfArgs &a, int n_mid, int variant, int storage, hipStream_t stream) {
    return launch_sf<HSS_CMP>(a, n_mid, variant, storage, stream);
}
}  // namespace hss
