// inst_phipi.hip -- fused stage kernels of the SysPhiPi system for every supported
// (stencil width, KO order) pair; one translation unit per system so the
// build compiles them in parallel.
#include "stage_kernel.hpp"

namespace tsp {
LaunchFn lookup_phipi(int n, int r) { return lookup_fused<SysPhiPi>(n, r); }
} // namespace tsp

namespace tsp {
LaunchFn lookup_phipi_adjoint(int n, int r) { return lookup_fused<SysPhiPiT, true>(n, r); }
} // namespace tsp
