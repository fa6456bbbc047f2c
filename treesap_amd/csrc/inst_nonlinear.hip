// inst_nonlinear.hip -- fused stage kernels of the SysNonLinear system for every supported
// (stencil width, KO order) pair; one translation unit per system so the
// build compiles them in parallel.
#include "stage_kernel.hpp"

namespace tsp {
LaunchFn lookup_nonlinear(int n, int r) { return lookup_fused<SysNonLinear>(n, r); }
} // namespace tsp
