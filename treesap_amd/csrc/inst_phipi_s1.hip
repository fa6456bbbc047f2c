// inst_phipi_s1.hip -- fused stage kernels of the SysPhiPi system specialised for
// the middle forward RK4 stages 1 and 2 (compile-time stage kind: init and acc
// read, arg_next written), for every supported (stencil width, KO order) pair;
// a translation unit of its own so the build compiles it in parallel.
#include "stage_kernel.hpp"

namespace tsp {
LaunchFn lookup_phipi_s1(int n, int r) { return lookup_fused<SysPhiPi, false, kStage1>(n, r); }
} // namespace tsp
