// inst_phiu.hip -- fused stage kernels of the SysPhiU system for every supported
// (stencil width, KO order) pair; one translation unit per system so the
// build compiles them in parallel.
#include "stage_kernel.hpp"

namespace tsp {
LaunchFn lookup_phiu(int n, int r) { return lookup_fused<SysPhiU>(n, r); }
} // namespace tsp
