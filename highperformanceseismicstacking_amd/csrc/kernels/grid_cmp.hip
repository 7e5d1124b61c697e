// Grid-search semblance instantiations for the CMP traveltime (see grid_impl.hpp).
#include "grid_impl.hpp"

namespace hss {
int launch_grid_cmp(const GridArgs &a, int n_slots, int variant, int storage, int cpl,
                    hipStream_t stream) {
    return launch_grid<HSS_CMP>(a, n_slots, variant, storage, cpl, stream);
}
}  // namespace hss
