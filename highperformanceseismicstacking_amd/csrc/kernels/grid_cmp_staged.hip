// CMP grid search with workgroup-staged traces: instantiations and launcher
// (see grid_staged.hpp).
#include "grid_staged.hpp"

namespace hss {
int launch_grid_cmp_staged(const GridArgs &a, int n_slots, int variant, int storage, int cpl,
                           hipStream_t stream) {
    if (a.tt != 1) return HSS_ERR_INVALID_SIZE;
    int st = HSS_OK;
    const unsigned blocks = (unsigned)n_slots * (unsigned)a.ntiles;
    if (blocks == 0) return HSS_OK;
    int rc = dispatch_storage(storage, [&](auto S) {
        st = dispatch_variant(variant, [&](auto V) {
            dispatch_window(a.e.w, fast_index_verified(a.e.dt, a.e.tau, stream), [&](auto Wc) {
                constexpr int WW = decltype(Wc)::value, VV = decltype(V)::value,
                              SS = decltype(S)::value;
                if (cpl == 2 && WW > 0)
                    hipLaunchKernelGGL((grid_staged_kernel<WW, VV, SS, 2>), dim3(blocks),
                                       dim3(kGridThreads), 0, stream, a);
                else
                    hipLaunchKernelGGL((grid_staged_kernel<WW, VV, SS, 1>), dim3(blocks),
                                       dim3(kGridThreads), 0, stream, a);
            });
        });
    });
    if (rc != HSS_OK) return rc;
    if (st != HSS_OK) return st;
    return hip_status(hipGetLastError());
}
}  // namespace hss
