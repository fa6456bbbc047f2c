// user_rhs_traits.hpp -- system_traits mappings of the user right-hand sides of
// user_rhs.hpp onto the systems the device library carries, so that the
// programs including this header run them on the library's tiled kernels. The
// drop-in layer verifies every such mapping when a model is mapped (bitwise,
// on random fields at cells spread over the domain) and throws on a mismatch.
// Programs that do not include it run the same functors on the generic kernel.
#pragma once

#include "device/model_map.hpp"
#include "user_rhs.hpp"

namespace treesap {
template <int g> struct system_traits<FirstOrderWave<g>> {
    static constexpr int system = TSP_SYS_WAVE_PHIU, stencil_n = g;
    static double gamma(const FirstOrderWave<g> &) { return 0.0; }
};
template <int o, int g> struct system_traits<ScalarWave<o, g>> {
    static constexpr int system = TSP_SYS_WAVE_PHIPI, stencil_n = o;
    static double gamma(const ScalarWave<o, g> &) { return 0.0; }
};
template <int g> struct system_traits<SelfDampedWave<g>> {
    static constexpr int system = TSP_SYS_NONLINEAR, stencil_n = g;
    static double gamma(const SelfDampedWave<g> &r) { return r.gamma; }
};
template <int o, int g> struct system_traits<VacuumADM<o, g>> {
    static constexpr int system = TSP_SYS_TENSOR_ADM, stencil_n = o;
    static double gamma(const VacuumADM<o, g> &) { return 0.0; }
};
// deliberately wrong mappings: verification must reject them
template <int g> struct system_traits<Advection<g>> {
    static constexpr int system = TSP_SYS_WAVE_PHIPI, stencil_n = g;
    static double gamma(const Advection<g> &) { return 0.0; }
};
template <int g> struct system_traits<ForcedWave<g>> {
    static constexpr int system = TSP_SYS_WAVE_PHIPI, stencil_n = g;
    static double gamma(const ForcedWave<g> &) { return 0.0; }
};
template <int g> struct system_traits<PatchedWave<g>> {
    static constexpr int system = TSP_SYS_WAVE_PHIPI, stencil_n = g;
    static double gamma(const PatchedWave<g> &) { return 0.0; }
};
} // namespace treesap
