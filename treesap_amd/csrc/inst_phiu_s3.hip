// inst_phiu_s3.hip -- fused stage kernels of the SysPhiU system specialised for
// forward RK4 stage 3 (compile-time stage kind), for every supported
// (stencil width, KO order) pair; a translation unit of its own so the build
// compiles it in parallel with the generic kernels.
#include "stage_kernel.hpp"

namespace tsp {
LaunchFn lookup_phiu_s3(int n, int r) { return lookup_fused<SysPhiU, false, kStage3>(n, r); }
} // namespace tsp
