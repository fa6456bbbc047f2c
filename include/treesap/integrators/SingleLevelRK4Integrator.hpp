// SingleLevelRK4Integrator.hpp -- classical four-stage Runge-Kutta on one
// level (integrators/SingleLevelRK4Integrator.hpp:50-169). Each step is four
// fused stage launches of the device library: stage s evaluates the decorated
// RHS k_s on arg (init for s = 0), accumulates acc = init + b_0 k_0 (s = 0) or
// acc += b_s k_s, and forms arg_next = init + a_{s+1} k_s for s < 3, with
// b = dt * {1/6, 1/3, 1/3, 1/6} and a = dt * {0, 1/2, 1/2, 1}; ghost cells of
// each stage's source are filled first (FillPatch: periodic images, then the
// boundary functor).
#pragma once

#include "../SingleLevelIntegrator.hpp"

namespace treesap {

template <typename PDEModelType>
    requires PDEModelConcept<PDEModelType>
class SingleLevelRK4Integrator : public SingleLevelIntegrator<PDEModelType> {
    using Base = SingleLevelIntegrator<PDEModelType>;
    using Base::attach_boundary;
    using Base::custom_boundary;
    using Base::generic;
    using Base::library_kernels;
    using Base::require_model_dt;
    using Base::solver;

  public:
    SingleLevelRK4Integrator(PDEModelType model_, const amrex::IntVect &block_size)
        : Base(std::move(model_), block_size) {}

    // one RK4 step of the whole level
    amrex::Real computeNewState(amrex::Real delta_time) override {
        require_model_dt(delta_time);
        attach_boundary();
        if constexpr (library_kernels) solver.rk4_steps(1);
        else generic->rk4_steps(solver, 1);
        return delta_time;
    }

  protected:
    // graph replay where it applies: not under a custom boundary functor, whose
    // fills each pass their own time
    void device_steps(std::int64_t n, bool graph) override {
        attach_boundary();
        if constexpr (!library_kernels) generic->rk4_steps(solver, n);
        else if (graph && !custom_boundary) solver.graph_rk4_steps(n);
        else solver.rk4_steps(n);
    }
};

} // namespace treesap
