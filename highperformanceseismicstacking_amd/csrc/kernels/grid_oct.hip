// Grid-search semblance instantiations for the OCT traveltime (see grid_impl.hpp).
#include "grid_impl.hpp"

namespace hss {
int launch_grid_oct(const GridArgs &a, int n_slots, int variant, int storage, int cpl,
                    hipStream_t stream) {
    return launch_grid<HSS_OCT>(a, n_slots, variant, storage, cpl, stream);
}
}  // namespace hss
