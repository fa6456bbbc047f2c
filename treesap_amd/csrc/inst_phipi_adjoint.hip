// inst_phipi_adjoint.hip -- fused stage kernels of the transposed (phi, Pi)
// system (SysPhiPiT) that the discrete adjoint runs, for every supported
// (stencil width, KO order) pair.
#include "stage_kernel.hpp"

namespace tsp {
LaunchFn lookup_phipi_adjoint(int n, int r) { return lookup_fused<SysPhiPiT, true>(n, r); }
} // namespace tsp
