// generic_runner.hpp -- steps a model whose right-hand side the library does
// not carry (TSP_SYS_CUSTOM) through tsp_rk4_steps_custom /
// tsp_euler_steps_custom with the stage kernel of device/generic.hpp, which is
// compiled from the model's functor. That kernel exists only where the
// including translation unit is compiled by a HIP compiler; a host-only build
// that steps such a model fails to compile with a message saying so.
#pragma once

#include "../amrex_types.hpp"
#include "solver.hpp"

#include <cstdint>

namespace treesap::device {

// what the stage callback needs: a copy of the (decorated) functor and the geometry
template <class RHS> struct GenericStage {
    RHS rhs;
    amrex::GeometryData geom;
};

template <class RHS> int stage_callback(void *user, const tsp_stage_view *v, void *stream);

template <class> constexpr bool generic_needs_hip = false;

template <class RHS> class GenericRunner {
  public:
    GenericRunner(const RHS &rhs, const amrex::GeometryData &g) : st_{rhs, g} {}
    void rk4_steps(Solver &s, std::int64_t n) {
#if defined(__HIP__)
        check(tsp_rk4_steps_custom(s.handle(), n, &stage_callback<RHS>, &st_), "tsp_rk4_steps_custom");
#else
        static_assert(generic_needs_hip<RHS>, "treesap: this right-hand side is not one the device library carries; "
                                              "its stage kernel is compiled from the functor, so compile this "
                                              "program with hipcc (or map the functor with system_traits)");
#endif
    }
    void euler_steps(Solver &s, std::int64_t n) {
#if defined(__HIP__)
        check(tsp_euler_steps_custom(s.handle(), n, &stage_callback<RHS>, &st_), "tsp_euler_steps_custom");
#else
        static_assert(generic_needs_hip<RHS>, "treesap: this right-hand side is not one the device library carries; "
                                              "its stage kernel is compiled from the functor, so compile this "
                                              "program with hipcc (or map the functor with system_traits)");
#endif
    }

  private:
    GenericStage<RHS> st_;
};

} // namespace treesap::device

#if defined(__HIP__)
#include "generic.hpp"
#endif
