// inst_phipi.hip -- fused stage kernels of the SysPhiPi system for every supported
// (stencil width, KO order) pair; one translation unit per system (the
// transposed system of the adjoint has its own) so the build compiles them in
// parallel.
#include "stage_kernel.hpp"

namespace tsp {
LaunchFn lookup_phipi(int n, int r) { return lookup_fused<SysPhiPi>(n, r); }
} // namespace tsp
