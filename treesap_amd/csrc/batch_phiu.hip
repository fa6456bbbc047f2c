// batch_phiu.hip -- batched stage kernels of the (phi, u) system for every
// variant (generic stage body); a translation unit of its own so the build
// compiles it in parallel.
#include "batch_kernel.hpp"

namespace tsp {
hipError_t launch_batched_phiu(int variant, const BatchArgs &b, int tile_y, int grid, int max_abl_T, int halo_ring,
                               hipStream_t stream) {
    return launch_batched_mode<SysPhiU, -1>(variant, b, tile_y, grid, max_abl_T, halo_ring, stream);
}
} // namespace tsp
