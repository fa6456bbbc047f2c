// inst_nonlinear_s0.hip -- fused stage kernels of the SysNonLinear system specialised for
// forward RK4 stage 0 (compile-time stage kind), for every supported
// (stencil width, KO order) pair; a translation unit of its own so the build
// compiles it in parallel with the generic kernels.
#include "stage_kernel.hpp"

namespace tsp {
LaunchFn lookup_nonlinear_s0(int n, int r) { return lookup_fused<SysNonLinear, false, kStage0>(n, r); }
} // namespace tsp
