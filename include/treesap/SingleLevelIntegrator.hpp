// SingleLevelIntegrator.hpp -- integrator on one uniform level whose state
// lives in device memory behind the C ABI (SingleLevelIntegrator.hpp:22-187).
//
// Construction maps the model onto a device solver (device/model_map.hpp) and
// allocates its registers. Right-hand sides the library carries (registered, or
// mapped by system_traits) run on its tiled stage kernels; any other functor
// runs on a stage kernel compiled from it (device/generic.hpp), which needs
// this program to be compiled by a HIP compiler.
// As in the reference, the level is split into boxes of at most block_size
// cells per axis (SingleLevelIntegrator.hpp:47-49,172-176), each box with its
// own registers, and FillPatch exchanges ghosts between the boxes before every
// stage (tsp_level_*); results do not depend on the split. The (phi, Pi),
// (phi, u) and NonLinear systems run split levels; the tensor system and
// functors on the generic kernel run the level as one box, which gives the
// same results.
// The model's boundary functor: NoOpBC and DirichletBC are the library's own;
// any other BCConcept functor runs on a kernel compiled from it
// (device/generic_bc.hpp) after each fill's periodic step, on one box, with
// steps taken eagerly (each fill passes its own time to the functor).
// initialise() evaluates the initial condition at cell centres
// x = prob_lo + (i + 0.5) dx (SingleLevelIntegrator.hpp:75-93): a FunctionInit
// on the device (its expressions compiled to postfix programs), any other
// initial-condition functor on the host, uploaded.
// pre_step() exports a plotfile when one is due. The FillPatch before each
// stage (SingleLevelIntegrator.hpp:153-163) and the swap of old and new state
// after the step happen inside the library's step.
//
// Additions without a reference counterpart: advance(n) runs n steps as
// replays of a captured device graph (bit-identical to n calls of step()),
// get_state()/set_state() move the valid state between host and device, and
// autotune() picks the stage-kernel tiling for this shape.
#pragma once

#include "Integrator.hpp"
#include "device/generic_bc.hpp"
#include "device/generic_runner.hpp"
#include "device/model_map.hpp"
#include "device/solver.hpp"
#include "init/FunctionInit.hpp"
#include "util/PlotExporter.hpp"

#include <cstdint>
#include <optional>
#include <stdexcept>
#include <vector>

namespace treesap {

template <typename PDEModelType>
    requires PDEModelConcept<PDEModelType>
class SingleLevelIntegrator : public Integrator<PDEModelType> {
  protected:
    using Base = Integrator<PDEModelType>;
    using Base::ncomp;
    using Base::ngrow;
    using typename Base::InitType;
    using Base::model;
    using Base::steps;
    using Base::time;

  public:
    using RHSType = typename PDEModelType::rhs_t;
    // whether the library's tiled kernels run this model (else the generic kernel)
    static constexpr bool library_kernels = device::runs_library_kernels<PDEModelType>;
    // whether the boundary functor runs on a kernel compiled from it
    static constexpr bool custom_boundary = device::custom_boundary<PDEModelType>;

    SingleLevelIntegrator(PDEModelType model_, const amrex::IntVect &block_size)
        : Base(std::move(model_)), solver(make_solver(model, block_size)) {
        if constexpr (!library_kernels) generic.emplace(model.get_rhs(), model.get_geometry().data());
        if constexpr (custom_boundary)
            boundary.emplace(model.get_bc(), model.get_geometry().data(), model.get_bc_recs());
        ScopedTimer::set_timing_stream(solver.stream());
    }
    ~SingleLevelIntegrator() override {
        if (ScopedTimer::timing_stream() == solver.stream()) ScopedTimer::set_timing_stream(nullptr);
    }

    void configureExporter(const fs::path &path, int export_interval) {
        const amrex::Box &dom = model.get_geometry().Domain();
        std::vector<amrex::Box> boxes;
        for (const auto &b : solver.box_bounds()) // domain-relative lo, hi of every box of the level
            boxes.emplace_back(amrex::IntVect(dom.smallEnd(0) + b[0], dom.smallEnd(1) + b[1], dom.smallEnd(2) + b[2]),
                               amrex::IntVect(dom.smallEnd(0) + b[3], dom.smallEnd(1) + b[4], dom.smallEnd(2) + b[5]));
        exporter.emplace([this](std::vector<amrex::Real> &v) { v = get_state(); }, model.get_geometry(),
                         model.get_comp_names(), path, export_interval, std::move(boxes));
        interval_ = export_interval;
    }

    void initialise_impl() override {
        const InitType init = model.get_init();
        const amrex::Geometry &geom = model.get_geometry();
        if constexpr (is_function_init<InitType>::value) {
            // the expressions run on the device at every cell centre
            std::vector<tsp_expr_op> ops;
            std::vector<std::int32_t> start;
            init.program(ops, start);
            const amrex::Box &dom = geom.Domain();
            const std::int32_t index_lo[3] = {dom.smallEnd(0), dom.smallEnd(1), dom.smallEnd(2)};
            solver.init_expr(ops.data(), start.data(), index_lo);
            return;
        }
        const auto dx = geom.CellSizeArray();
        const auto lo = geom.ProbLoArray();
        const amrex::Box &dom = geom.Domain();
        const int nx = dom.length(0), ny = dom.length(1), nz = dom.length(2);
        std::vector<amrex::Real> u((std::size_t)nx * ny * nz * ncomp);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const int ia = dom.smallEnd(0) + i, ja = dom.smallEnd(1) + j, ka = dom.smallEnd(2) + k;
                    const amrex::Real x = lo[0] + (ia + 0.5) * dx[0];
                    const amrex::Real y = lo[1] + (ja + 0.5) * dx[1];
                    const amrex::Real z = lo[2] + (ka + 0.5) * dx[2];
                    static_for<ncomp>([&]<int c>() {
                        u[(((std::size_t)c * nz + k) * ny + j) * nx + i] = init.template operator()<c>(x, y, z);
                    });
                }
        solver.set_state(u.data());
    }

    // nsteps calls of step(), with the device steps between exports replayed from a captured graph
    void advance(std::int64_t nsteps) {
        ScopedTimer::StreamGuard on(solver.stream());
        while (nsteps > 0) {
            std::int64_t chunk = nsteps;
            if (exporter) {
                ScopedTimer timer("Integrator::step() [PRE]");
                pre_step();
                chunk = std::min<std::int64_t>(nsteps, interval_ - (std::int64_t)(steps % interval_));
            }
            {
                ScopedTimer timer("Integrator::advance()");
                device_steps(chunk, true);
            }
            const amrex::Real dt = model.dt_hint();
            for (std::int64_t q = 0; q < chunk; ++q) {
                steps++;
                time += dt;
            }
            nsteps -= chunk;
        }
    }

    std::vector<amrex::Real> get_state() const {
        std::vector<amrex::Real> u((std::size_t)solver.cells() * ncomp);
        solver.get_state(u.data());
        return u;
    }
    void set_state(const std::vector<amrex::Real> &u) {
        if ((std::int64_t)u.size() != solver.cells() * ncomp)
            throw std::invalid_argument("treesap: set_state needs ncomp * cells values");
        solver.set_state(u.data());
    }
    tsp_kernel_cfg autotune(int steps_per_candidate = 3) {
        attach_boundary();
        return solver.autotune(steps_per_candidate);
    }
    device::Solver &device_solver() { return solver; }

  protected:
    void pre_step() final override {
        if (exporter) exporter->exportIfNecessary((std::int64_t)steps, time);
    }
    amrex::Real step_impl() final override {
        return computeNewState(model.dt_hint());
    }
    void post_step() final override {}
    void *timing_stream() const final override { return solver.stream(); }

    virtual amrex::Real computeNewState(amrex::Real dt_hint) = 0;
    virtual void device_steps(std::int64_t n, bool graph) = 0;

    void require_model_dt(amrex::Real dt) const {
        if (!(dt == model.dt_hint()))
            throw std::invalid_argument("treesap: the device solver steps with the model's dt_hint()");
    }

    static device::Solver make_solver(const PDEModelType &m, const amrex::IntVect &block_size) {
        for (int a = 0; a < 3; ++a)
            if (block_size[a] < 1) throw std::invalid_argument("treesap: block_size entries must be positive");
        const int block[3] = {block_size[0], block_size[1], block_size[2]};
        return device::Solver(device::make_desc(m), block);
    }

    // before stepping: a custom boundary functor's callback, with this integrator's copy
    void attach_boundary() {
        if constexpr (custom_boundary) boundary->attach(solver);
    }

    mutable device::Solver solver;
    std::optional<device::GenericRunner<RHSType>> generic; // models on the generic kernel
    std::optional<device::BCRunner<typename PDEModelType::bc_t>> boundary; // custom boundary functors

  private:
    std::optional<PlotExporter> exporter;
    std::int64_t interval_ = 1;
};

} // namespace treesap
