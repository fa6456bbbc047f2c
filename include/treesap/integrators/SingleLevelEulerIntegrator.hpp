// SingleLevelEulerIntegrator.hpp -- forward Euler on one level
// (integrators/SingleLevelEulerIntegrator.hpp:27-91): new = old + dt F(old),
// one fused stage launch per step after the ghost fill of old.
#pragma once

#include "../SingleLevelIntegrator.hpp"

namespace treesap {

template <typename PDEModelType>
    requires PDEModelConcept<PDEModelType>
class SingleLevelEulerIntegrator : public SingleLevelIntegrator<PDEModelType> {
    using Base = SingleLevelIntegrator<PDEModelType>;
    using Base::attach_boundary;
    using Base::generic;
    using Base::library_kernels;
    using Base::require_model_dt;
    using Base::solver;

  public:
    SingleLevelEulerIntegrator(PDEModelType model_, const amrex::IntVect &block_size)
        : Base(std::move(model_), block_size) {}

    amrex::Real computeNewState(amrex::Real delta_time) override {
        require_model_dt(delta_time);
        attach_boundary();
        if constexpr (library_kernels) solver.euler_steps(1);
        else generic->euler_steps(solver, 1);
        return delta_time;
    }

  protected:
    void device_steps(std::int64_t n, bool) override {
        attach_boundary();
        if constexpr (library_kernels) solver.euler_steps(n);
        else generic->euler_steps(solver, n);
    }
};

} // namespace treesap
