// inst_phipi_s3.hip -- fused stage kernels of the SysPhiPi system specialised for
// forward RK4 stage 3 (compile-time stage kind), for every supported
// (stencil width, KO order) pair; a translation unit of its own so the build
// compiles it in parallel with the generic kernels.
#include "stage_kernel.hpp"

namespace tsp {
LaunchFn lookup_phipi_s3(int n, int r) { return lookup_fused<SysPhiPi, false, kStage3>(n, r); }
} // namespace tsp
