// Integrator.hpp -- base of the time integrators: owns the model, counts
// steps and accumulates time (Integrator.hpp:53-148). step() runs
// pre_step(), then dt = step_impl(), steps += 1, time += dt, then post_step(),
// each inside a ScopedTimer; initialise() runs initialise_impl(). The timers
// record on the stream timing_stream() names (the device solver's stream in
// the single-level integrators).
#pragma once

#include "pde-model/PDEModelConcepts.hpp"
#include "util/ScopedTimer.hpp"

#include <cstddef>
#include <utility>

namespace treesap {

template <typename PDEModelType>
    requires PDEModelConcept<PDEModelType>
class Integrator {
  protected:
    static constexpr int ncomp = PDEModelType::ncomp;
    static constexpr int ngrow = PDEModelType::ngrow;
    using RHSType = typename PDEModelType::rhs_t;
    using InitType = typename PDEModelType::init_t;
    using BCType = typename PDEModelType::bc_t;

  public:
    explicit Integrator(PDEModelType model_) : model(std::move(model_)) {}
    virtual ~Integrator() = default;

    void initialise() {
        ScopedTimer::StreamGuard on(timing_stream());
        ScopedTimer timer("Integrator::initialise()");
        initialise_impl();
    }

    void step() {
        ScopedTimer::StreamGuard on(timing_stream());
        {
            ScopedTimer timer("Integrator::step() [PRE]");
            pre_step();
        }
        {
            ScopedTimer timer("Integrator::step() [STEP]");
            const amrex::Real dt = step_impl();
            steps++;
            time += dt;
        }
        {
            ScopedTimer timer("Integrator::step() [POST]");
            post_step();
        }
    }

    // Read access to the counters (the reference keeps them protected).
    amrex::Real get_time() const { return time; }
    std::size_t get_steps() const { return steps; }
    const PDEModelType &get_model() const { return model; }

  protected:
    virtual void initialise_impl() = 0;
    virtual amrex::Real step_impl() = 0;
    virtual void pre_step() {}
    virtual void post_step() {}
    // hipStream_t of the integrator's device work (nullptr: none)
    virtual void *timing_stream() const { return nullptr; }

    PDEModelType model;
    std::size_t steps = 0;
    amrex::Real time = 0.0;
};

} // namespace treesap
