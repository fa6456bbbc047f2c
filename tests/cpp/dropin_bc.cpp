// dropin_bc.cpp -- TreESaP-style programs whose boundary functor the device
// library does not carry (neither NoOpBC nor DirichletBC): compiled with hipcc,
// the functor runs on a kernel compiled from it (device/generic_bc.hpp) after
// each fill's periodic step. Each scenario writes into <outdir>/<scenario>/:
//   init.bin   the state after initialise(), a[((c*nz+k)*ny+j)*nx+i]
//   final.bin  the state after the device steps
//   host.bin   the same steps of the same functors on the host (host_rk4.hpp)
//   noop.bin   the host steps with ghost memory left untouched instead (NoOpBC),
//              to show the boundary functor matters
//   noalloc.bin, noalloc_host.bin (library_rk4): two more steps from the final
//              state through tsp_rk4_steps_noalloc, and the same on the host
//   refusals.txt (library_rk4): the status codes of tsp_graph_rk4_steps and of
//              tsp_set_kernel with ghost_mode 0 on that solver
//   meta.txt   "<integrator> <steps> <time as %.17g> <system id> <bc id>"
//
//   dropin_bc <outdir> [scenario ...]
#include "treesap.hpp"
#include "host_rk4.hpp"
#include "user_rhs_traits.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

using namespace treesap;
namespace fs = std::filesystem;

// A driven boundary: component 0 takes amp * t * (1 - t) * x at the ghost
// cell's centre x (time- and position-dependent); every other component copies
// the nearest valid cell, scaled by 1 where its boundary record says foextrap
// on the high z face and by 0.5 otherwise (so the records reach the device).
// Only valid cells are read.
template <int nc, int ng> struct DrivenBC {
    static constexpr int ncomp = nc;
    static constexpr int ngrow = ng;
    amrex::Real amp = 0.0;

    AMREX_GPU_HOST_DEVICE
    void operator()(amrex::IntVect const &iv, amrex::Array4<amrex::Real> const &dest, int dcomp, int numcomp,
                    amrex::GeometryData const &geom, amrex::Real time, const amrex::BCRec *bcr, int bcomp,
                    int) const {
        const amrex::Box &dom = geom.Domain();
        amrex::IntVect in = iv;
        for (int d = 0; d < 3; ++d)
            in[d] = iv[d] < dom.smallEnd(d) ? dom.smallEnd(d) : iv[d] > dom.bigEnd(d) ? dom.bigEnd(d) : iv[d];
        const amrex::Real x = geom.ProbLo(0) + (iv[0] - dom.smallEnd(0) + 0.5) * geom.CellSize(0);
        dest(iv, dcomp) = amp * time * (1.0 - time) * x;
        for (int c = 1; c < numcomp; ++c) {
            const amrex::Real f = bcr[bcomp + c].hi(2) == amrex::BCType::foextrap ? 1.0 : 0.5;
            dest(iv, dcomp + c) = f * dest(in, dcomp + c);
        }
    }
};
static_assert(BCConcept<DrivenBC<2, 2>>);

struct Quadratic {
    constexpr amrex::Real operator()(amrex::Real s) const { return s * s; }
};

static void write_bin(const fs::path &p, const std::vector<double> &v) {
    std::ofstream f(p, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
    if (!f) throw std::runtime_error("cannot write " + p.string());
}

template <class Integ>
static void dump(const fs::path &dir, Integ &integ, const std::vector<double> &init, const char *kind) {
    write_bin(dir / "init.bin", init);
    write_bin(dir / "final.bin", integ.get_state());
    std::ofstream m(dir / "meta.txt");
    char buf[200];
    std::snprintf(buf, sizeof buf, "%s %zu %.17g %d %d\n", kind, integ.get_steps(), integ.get_time(),
                  integ.device_solver().desc().system, integ.device_solver().desc().bc);
    m << buf;
}

static amrex::Geometry geom_at(amrex::IntVect lo, int nx, int ny, int nz, std::array<double, 3> hi,
                               std::array<int, 3> per) {
    const amrex::RealBox rb({0.0, 0.0, 0.0}, hi);
    return amrex::Geometry(amrex::Box(lo, lo + amrex::IntVect(nx - 1, ny - 1, nz - 1)), &rb, amrex::CoordSys::cartesian,
                           per.data());
}

// boundary records: foextrap on the high z face for the odd components
template <int nc> static amrex::Vector<amrex::BCRec> records() {
    amrex::Vector<amrex::BCRec> r;
    for (int c = 0; c < nc; ++c) {
        const int hz = c % 2 ? amrex::BCType::foextrap : amrex::BCType::int_dir;
        r.emplace_back(amrex::BCType::int_dir, amrex::BCType::int_dir, amrex::BCType::int_dir, amrex::BCType::int_dir,
                       amrex::BCType::int_dir, hz);
    }
    return r;
}

// host stepping of the same right-hand side and boundary functor (or NoOp)
template <class RHS, class BC>
static std::vector<double> host_steps(const RHS &rhs, const BC &bc, const amrex::Geometry &geom, double dt,
                                      const amrex::Vector<amrex::BCRec> &recs, const std::vector<double> &u0,
                                      std::size_t steps, bool euler, bool with_bc = true) {
    HostStepper<RHS> host(rhs, geom, dt, false, nullptr);
    const amrex::GeometryData gd = geom.data();
    if (with_bc)
        host.custom_bc = [bc, gd, recs](const amrex::IntVect &iv, const amrex::Array4<double> &a, double t) {
            bc(iv, a, 0, RHS::ncomp, gd, t, recs.data(), 0, 0);
        };
    host.set_state(u0);
    for (std::size_t s = 0; s < steps; ++s) euler ? host.euler_step() : host.rk4_step();
    return host.get_state();
}

// (phi, Pi) with KO on the library's tiled kernels (ScalarWave is mapped by
// user_rhs_traits.hpp), periodic in x, the driven boundary on the y and z
// faces; a block size smaller than the domain (the level runs as one box)
static void scenario_library(const fs::path &dir, bool euler) {
    const auto geom = geom_at({0, 0, 0}, 28, 20, 24, {1.0, 0.8, 0.9}, {1, 0, 0});
    using RHSType = KODecorator<ScalarWave<2, 2>, 2>;
    using BCType = DrivenBC<2, 2>;
    using ModelType = PDEModel<RHSType, FunctionInit<2>, BCType>;
    static_assert(device::runs_library_kernels<ModelType> && device::custom_boundary<ModelType>);
    FunctionInit<2> init({"exp(-((x-0.4)^2 + (y-0.45)^2 + (z-0.5)^2)/0.15^2)", "0.2*sin(2*pi*x)*y"});
    const BCType bc{3.0};
    const auto recs = records<2>();
    const double dt = 0.2 / 28;
    ModelType model({"phi", "Pi"}, geom, dt, RHSType{{}, 0.1}, init, bc, recs);
    const RHSType rhs = model.get_rhs();
    std::vector<double> u0;
    std::size_t steps = 0;
    if (euler) {
        SingleLevelEulerIntegrator<ModelType> integ(ModelType(model), {16, 16, 16});
        integ.initialise();
        u0 = integ.get_state();
        for (int s = 0; s < 3; ++s) integ.step();
        integ.advance(2);
        dump(dir, integ, u0, "euler");
        steps = integ.get_steps();
    } else {
        SingleLevelRK4Integrator<ModelType> integ(ModelType(model), {16, 16, 16});
        integ.autotune(1);
        integ.initialise();
        u0 = integ.get_state();
        for (int s = 0; s < 2; ++s) integ.step();
        integ.advance(3); // eager steps: each fill passes its own time
        dump(dir, integ, u0, "rk4");
        steps = integ.get_steps();
        // two more steps through the allocation-free entry point (the integrator's
        // boundary callback stays attached), from the final state and time
        const std::vector<double> s = integ.get_state();
        const std::size_t bytes = s.size() * sizeof(double);
        double *din = nullptr, *dout = nullptr;
        if (hipMalloc((void **)&din, bytes) != hipSuccess || hipMalloc((void **)&dout, bytes) != hipSuccess)
            throw std::runtime_error("hipMalloc failed");
        if (hipMemcpy(din, s.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("hipMemcpy failed");
        device::check(tsp_rk4_steps_noalloc(integ.device_solver().handle(), din, dout, 2), "tsp_rk4_steps_noalloc");
        integ.device_solver().sync(); // the steps are enqueued on the solver's (non-blocking) stream
        std::vector<double> na(s.size());
        if (hipMemcpy(na.data(), dout, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("hipMemcpy failed");
        (void)hipFree(din);
        (void)hipFree(dout);
        write_bin(dir / "noalloc.bin", na);
        HostStepper<RHSType> host(rhs, geom, dt, false, nullptr);
        const amrex::GeometryData gd = geom.data();
        host.custom_bc = [bc, gd, recs](const amrex::IntVect &iv, const amrex::Array4<double> &a, double tt) {
            bc(iv, a, 0, RHSType::ncomp, gd, tt, recs.data(), 0, 0);
        };
        host.time = integ.get_time();
        host.set_state(s);
        for (int q = 0; q < 2; ++q) host.rk4_step();
        write_bin(dir / "noalloc_host.bin", host.get_state());
        // what the C ABI refuses under a custom boundary functor: graph replay
        // (it would freeze the fill times) and the virtual-ghost mode
        tsp_solver *h = integ.device_solver().handle();
        const int graph_status = tsp_graph_rk4_steps(h, 2);
        tsp_kernel_cfg cfg{};
        device::check(tsp_get_kernel(h, &cfg), "tsp_get_kernel");
        cfg.ghost_mode = 0;
        const int mode_status = tsp_set_kernel(h, &cfg);
        std::ofstream(dir / "refusals.txt") << graph_status << " " << mode_status << "\n";
    }
    write_bin(dir / "host.bin", host_steps(rhs, bc, geom, dt, recs, u0, steps, euler));
    write_bin(dir / "noop.bin", host_steps(rhs, bc, geom, dt, recs, u0, steps, euler, false));
}

// the linearised ADM system on the library's tensor kernel, non-periodic
static void scenario_tensor(const fs::path &dir) {
    const auto geom = geom_at({0, 0, 0}, 18, 16, 14, {1.0, 0.9, 0.8}, {0, 0, 0});
    using RHSType = KODecorator<VacuumADM<2, 2>, 2>;
    using BCType = DrivenBC<12, 2>;
    using ModelType = PDEModel<RHSType, FunctionInit<12>, BCType>;
    static_assert(device::runs_library_kernels<ModelType> && device::custom_boundary<ModelType>);
    amrex::Array<std::string, 12> exprs;
    for (int c = 0; c < 12; ++c)
        exprs[c] = "0.01*sin(2*pi*(" + std::to_string(c + 1) + "*x + y)) * cos(2*pi*" + std::to_string((c % 3) + 1) + "*z)";
    const BCType bc{0.5};
    const auto recs = records<12>();
    const double dt = 0.2 / 18;
    ModelType model({}, geom, dt, RHSType{{}, 0.1}, FunctionInit<12>(exprs), bc, recs);
    const RHSType rhs = model.get_rhs();
    SingleLevelRK4Integrator<ModelType> integ(std::move(model), {64, 64, 64});
    integ.initialise();
    const auto u0 = integ.get_state();
    integ.advance(3);
    dump(dir, integ, u0, "rk4");
    write_bin(dir / "host.bin", host_steps(rhs, bc, geom, dt, recs, u0, 3, false));
    write_bin(dir / "noop.bin", host_steps(rhs, bc, geom, dt, recs, u0, 3, false, false));
}

// a functor the library does not carry (the generic kernel) under an absorbing
// layer, on a domain whose first cell is not (0, 0, 0): cells, the layer's
// bulk box and the boundary functor all see the domain's own indices
static void scenario_generic_shifted(const fs::path &dir) {
    const auto geom = geom_at({5, -3, 7}, 22, 18, 20, {1.1, 0.9, 1.0}, {0, 1, 0});
    using RHSType = ABLDecorator<CurlWave<3>, Quadratic>;
    using Data = typename RHSType::ABLData;
    using BCType = DrivenBC<4, 3>;
    using ModelType = PDEModel<RHSType, FunctionInit<4>, BCType>;
    static_assert(!device::runs_library_kernels<ModelType> && device::custom_boundary<ModelType>);
    FunctionInit<4> init({"exp(-((x-0.5)^2 + (y-0.4)^2 + (z-0.6)^2)/0.2^2)", "0.2*sin(2*pi*z)", "0.1*x",
                          "0.3*cos(2*pi*x)*y"});
    const BCType bc{2.0};
    const auto recs = records<4>();
    const double dt = 0.2 / 22;
    const RHSType rhs{CurlWave<3>{}, geom, Data({{0.05, 0.0, -0.02, 0.0}}, 4, 1.5, Quadratic{})};
    ModelType model({"phi", "E_0", "E_1", "E_2"}, geom, dt, rhs, init, bc, recs);
    SingleLevelRK4Integrator<ModelType> integ(std::move(model), {64, 64, 64});
    integ.initialise();
    const auto u0 = integ.get_state();
    for (int s = 0; s < 2; ++s) integ.step();
    integ.advance(2);
    dump(dir, integ, u0, "rk4");
    write_bin(dir / "host.bin", host_steps(rhs, bc, geom, dt, recs, u0, 4, false));
    write_bin(dir / "noop.bin", host_steps(rhs, bc, geom, dt, recs, u0, 4, false, false));
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <outdir> [scenario ...]\n", argv[0]);
        return 2;
    }
    const fs::path out = argv[1];
    std::vector<std::string> which(argv + 2, argv + argc);
    if (which.empty()) which = {"library_rk4", "library_euler", "tensor", "generic_shifted"};
    for (const auto &w : which) {
        const fs::path dir = out / w;
        fs::create_directories(dir);
        std::printf("scenario %s\n", w.c_str());
        std::fflush(stdout);
        try {
            if (w == "library_rk4") scenario_library(dir, false);
            else if (w == "library_euler") scenario_library(dir, true);
            else if (w == "tensor") scenario_tensor(dir);
            else if (w == "generic_shifted") scenario_generic_shifted(dir);
            else {
                std::fprintf(stderr, "unknown scenario %s\n", w.c_str());
                return 2;
            }
        } catch (const std::exception &e) {
            std::fprintf(stderr, "scenario %s failed: %s\n", w.c_str(), e.what());
            return 1;
        }
    }
    return 0;
}
