// Stretch-free instantiations for the ZOCRS traveltime (see sf_impl.hpp).
#include "sf_impl.hpp"

namespace hss {
int launch_sf_zocrs(const SfArgs &a, int n_mid, int variant, int storage, hipStream_t stream) {
    return launch_sf<HSS_ZOCRS>(a, n_mid, variant, storage, stream);
}
}  // namespace hss
