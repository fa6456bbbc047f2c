// batch_s0.hip -- batched stage kernels of the (phi, Pi) system specialised
// for forward RK4 stage 0 (compile-time stage kind), for every variant; a
// translation unit of its own so the build compiles it in parallel.
#include "batch_kernel.hpp"

namespace tsp {
hipError_t launch_batched_phipi_s0(int variant, const BatchArgs &b, int tile_y, int grid, int max_abl_T, int halo_ring,
                                   hipStream_t stream) {
    return launch_batched_phipi_mode<kStage0>(variant, b, tile_y, grid, max_abl_T, halo_ring, stream);
}
} // namespace tsp
