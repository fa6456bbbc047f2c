// Stage-pair kernels of the (phi, Pi) wave for every (stencil width, KO order)
// with KO dissipation (the pair kernel reads every component with stencils).
#include "pair_kernel.hpp"

namespace tsp {

PairFn lookup_pair_phipi(int n, int r) {
#define TSP_CASE(NN, RR)                                                                                              \
    if (n == NN && r == RR) return &launch_pair<SysPhiPi, NN, RR>;
    TSP_CASE(1, 1)
    TSP_CASE(1, 2)
    TSP_CASE(2, 2)
    TSP_CASE(2, 3)
    TSP_CASE(3, 3)
    TSP_CASE(3, 4)
    TSP_CASE(4, 4)
#undef TSP_CASE
    return nullptr;
}

size_t pair_lds_bytes_phipi(int n, int r) {
#define TSP_CASE(NN, RR)                                                                                              \
    if (n == NN && r == RR) return sizeof(double) * StagePair<SysPhiPi, NN, RR, 0>::LDS_DOUBLES;
    TSP_CASE(1, 1)
    TSP_CASE(1, 2)
    TSP_CASE(2, 2)
    TSP_CASE(2, 3)
    TSP_CASE(3, 3)
    TSP_CASE(3, 4)
    TSP_CASE(4, 4)
#undef TSP_CASE
    return 0;
}

} // namespace tsp
