// dropin_gpu.cpp -- TreESaP-style programs run through the C++ drop-in
// integrators on the GPU. Each scenario builds a model from user-written RHS
// functors, initialises it, steps it, and writes into <outdir>/<scenario>/:
//   desc.bin   the tsp_desc the model was mapped to (raw struct bytes)
//   init.bin   the state after initialise(), a[((c*nz+k)*ny+j)*nx+i]
//   final.bin  the state after the steps
//   meta.txt   "<integrator> <steps> <time as %.17g> <boxes of the level>"
// The tests replay desc + init.bin on the CPU oracle and compare final.bin.
// The "timer" scenario (not run by default) prints device-timed scopes.
//
//   dropin_gpu <outdir> [scenario ...]
#include "treesap.hpp"
#include "user_rhs_traits.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

using namespace treesap;
namespace fs = std::filesystem;

struct Quadratic {
    constexpr amrex::Real operator()(amrex::Real s) const { return s * s; }
};

static void write_bin(const fs::path &p, const void *data, std::size_t bytes) {
    std::ofstream f(p, std::ios::binary);
    f.write(static_cast<const char *>(data), (std::streamsize)bytes);
    if (!f) throw std::runtime_error("cannot write " + p.string());
}

template <class Integ> static void dump(const fs::path &dir, Integ &integ, const std::vector<double> &init,
                                        const char *kind) {
    const std::vector<double> fin = integ.get_state();
    const tsp_desc &d = integ.device_solver().desc();
    write_bin(dir / "desc.bin", &d, sizeof d);
    write_bin(dir / "init.bin", init.data(), init.size() * sizeof(double));
    write_bin(dir / "final.bin", fin.data(), fin.size() * sizeof(double));
    std::ofstream m(dir / "meta.txt");
    char buf[128];
    std::snprintf(buf, sizeof buf, "%s %zu %.17g %d\n", kind, integ.get_steps(), integ.get_time(),
                  integ.device_solver().boxes());
    m << buf;
}

static amrex::Geometry box_geom(int nx, int ny, int nz, double lo, double hi, int per) {
    const int p[3] = {per, per, per};
    const amrex::RealBox rb({lo, lo, lo}, {hi, hi, hi});
    return amrex::Geometry(amrex::Box({0, 0, 0}, {nx - 1, ny - 1, nz - 1}), &rb, amrex::CoordSys::cartesian, p);
}

// (phi, u) wave, Dirichlet zero on a non-periodic [-2, 2]^3 split into 2 x 2 x 2
// boxes, plotfiles every 4 steps
static void scenario_first_order(const fs::path &dir) {
    const amrex::Geometry geom = box_geom(32, 32, 32, -2.0, 2.0, 0);
    constexpr int ngrow = 1;
    const amrex::Real dx = geom.CellSize(0);
    const amrex::Real dt = std::sqrt(8.0 / 3.0) * dx;
    using RHSType = FirstOrderWave<ngrow>;
    using InitType = FunctionInit<RHSType::ncomp>;
    using BCType = DirichletBC<RHSType::ncomp, RHSType::ngrow>;
    using ModelType = PDEModel<RHSType, InitType, BCType>;
    InitType init({"exp(-(x^2 + y^2 + z^2)/0.5^2)", "0.0", "0.0", "0.0"});
    ModelType model({"phi", "u_0", "u_1", "u_2"}, geom, dt, RHSType{}, init, BCType{0.0, 0.0, 0.0, 0.0});
    SingleLevelRK4Integrator<ModelType> integ(std::move(model), {16, 16, 16});
    integ.initialise();
    const auto u0 = integ.get_state();
    integ.configureExporter(dir / "plt", 4);
    for (int s = 0; s < 5; ++s) integ.step();
    integ.advance(7);
    dump(dir, integ, u0, "rk4");
}

// ABL(KO(user (phi, Pi))) through ABLPDEModel, n = 4, KO order 4, 8-cell quadratic layer
static void scenario_abl_ko(const fs::path &dir) {
    const amrex::Geometry geom = box_geom(48, 48, 48, 0.0, 1.0, 0);
    using RHSType = KODecorator<ScalarWave<4, 4>, 4>;
    using InitType = FunctionInit<2>;
    using ModelType = ABLPDEModel<RHSType, InitType, Quadratic>;
    using Data = typename ABLDecorator<RHSType, Quadratic>::ABLData;
    InitType init({"0.8*exp(-((x-0.45)^2 + (y-0.55)^2 + (z-0.5)^2)/0.1^2) - 0.5*exp(-((x-0.6)^2 + (y-0.4)^2 + "
                   "(z-0.55)^2)/0.07^2)",
                   "0.0"});
    ModelType model({"phi", "Pi"}, geom, 0.25 / 48, RHSType{{}, 0.2}, Data({{0.0, 0.0}}, 8, 2.0, Quadratic{}), init);
    SingleLevelRK4Integrator<ModelType> integ(std::move(model), {32, 32, 32});
    integ.initialise();
    const auto u0 = integ.get_state();
    integ.advance(10);
    dump(dir, integ, u0, "rk4");
}

// (phi, Pi) wave with KO on a periodic ragged box split into boxes of at most
// 7 x 9 x 64 cells (a level of 3 x 2 x 1 boxes), forward Euler steps, plotfiles
// every 5 steps
static void scenario_level_euler(const fs::path &dir) {
    const int p[3] = {1, 1, 1};
    const amrex::RealBox rb({0.0, 0.0, 0.0}, {1.0, 0.9, 0.8});
    const amrex::Geometry geom(amrex::Box({0, 0, 0}, {19, 17, 15}), &rb, amrex::CoordSys::cartesian, p);
    using RHSType = KODecorator<ScalarWave<2, 3>, 3>;
    using InitType = FunctionInit<2>;
    using ModelType = PDEModel<RHSType, InitType>;
    InitType init({"sin(2*pi*x)*cos(2*pi*y/0.9) + 0.3*sin(2*pi*z/0.8)", "0.5*cos(2*pi*(x + y))"});
    ModelType model({"phi", "Pi"}, geom, 0.002, RHSType{{}, 0.1}, init);
    SingleLevelEulerIntegrator<ModelType> integ(std::move(model), {7, 9, 64});
    integ.initialise();
    const auto u0 = integ.get_state();
    integ.configureExporter(dir / "plt", 5);
    for (int s = 0; s < 6; ++s) integ.step();
    dump(dir, integ, u0, "euler");
}

// nonlinear wave with a workspace temporary, KO, periodic ragged box split into
// boxes of at most 16 x 20 x 64 cells, forward Euler
static void scenario_nonlinear_euler(const fs::path &dir) {
    const int p[3] = {1, 1, 1};
    const amrex::RealBox rb({0.0, 0.0, 0.0}, {1.0, 0.9, 0.8});
    const amrex::Geometry geom(amrex::Box({0, 0, 0}, {39, 35, 31}), &rb, amrex::CoordSys::cartesian, p);
    using RHSType = KODecorator<SelfDampedWave<2>, 2>;
    using InitType = FunctionInit<2>;
    using ModelType = PDEModel<RHSType, InitType>;
    InitType init({"sin(2*pi*x)*cos(2*pi*y/0.9) + 0.3*sin(2*pi*z/0.8)", "0.5*cos(2*pi*(x + y))"});
    ModelType model({"phi", "pi"}, geom, 0.004, RHSType{{0.05}, 0.1}, init);
    SingleLevelEulerIntegrator<ModelType> integ(std::move(model), {16, 20, 64}); // a level of 3 x 2 x 1 boxes
    integ.initialise();
    const auto u0 = integ.get_state();
    for (int s = 0; s < 8; ++s) integ.step();
    dump(dir, integ, u0, "euler");
}

// linearised ADM (12 components) with KO and an ABL through ABLPDEModel
static void scenario_tensor(const fs::path &dir) {
    const amrex::Geometry geom = box_geom(24, 24, 24, 0.0, 1.0, 0);
    using RHSType = KODecorator<VacuumADM<2, 2>, 2>;
    using InitType = FunctionInit<12>;
    using ModelType = ABLPDEModel<RHSType, InitType, Quadratic>;
    using Data = typename ABLDecorator<RHSType, Quadratic>::ABLData;
    amrex::Array<std::string, 12> exprs;
    for (int c = 0; c < 12; ++c)
        exprs[c] = "0.01*sin(2*pi*(" + std::to_string(c + 1) + "*x + " + std::to_string((c % 3) + 1) +
                   "*y)) * cos(2*pi*" + std::to_string((c % 4) + 1) + "*z)";
    amrex::GpuArray<amrex::Real, 12> tgt{};
    for (int c = 0; c < 12; ++c) tgt[c] = 0.0;
    ModelType model({}, geom, 0.25 / 24, RHSType{{}, 0.1}, Data(tgt, 4, 1.0, Quadratic{}), InitType(exprs));
    SingleLevelRK4Integrator<ModelType> integ(std::move(model), {24, 24, 24});
    integ.initialise();
    const auto u0 = integ.get_state();
    integ.step();
    integ.advance(3);
    dump(dir, integ, u0, "rk4");
}

// ScopedTimer on the device: 120 graph-replayed steps are queued without
// waiting, then a scope enqueues 12 more; the scope's time covers its own 12
// steps, not the queue ahead of it. Prints "host_us <host time of both>"
// and the timers' lines (no files).
static void scenario_timer(const fs::path &) {
    const amrex::Geometry geom = box_geom(96, 96, 96, 0.0, 1.0, 1);
    using RHSType = KODecorator<ScalarWave<4, 4>, 4>;
    using ModelType = PDEModel<RHSType, FunctionInit<2>>;
    ModelType model({"phi", "Pi"}, geom, 0.25 / 96, RHSType{{}, 0.2},
                    FunctionInit<2>({"exp(-((x-0.5)^2 + (y-0.5)^2 + (z-0.5)^2)/0.1^2)", "0.0"}));
    SingleLevelRK4Integrator<ModelType> integ(std::move(model), {96, 96, 96});
    integ.initialise();
    integ.advance(4); // captures the replayed graphs before timing
    device::Solver &sv = integ.device_solver();
    sv.sync();
    ScopedTimer::enabled = true;
    const auto t0 = std::chrono::steady_clock::now();
    sv.graph_rk4_steps(120);
    {
        ScopedTimer t("scope of 12 steps");
        sv.graph_rk4_steps(12);
    }
    const double host_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    for (int s = 0; s < 3; ++s) integ.step();
    integ.advance(8);
    ScopedTimer::printOverview();
    ScopedTimer::enabled = false;
    std::printf("host_us %.1f\n", host_us);
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <outdir> [scenario ...]\n", argv[0]);
        return 2;
    }
    const fs::path out = argv[1];
    std::vector<std::string> which(argv + 2, argv + argc);
    if (which.empty()) which = {"first_order", "abl_ko", "nonlinear_euler", "tensor", "level_euler"};
    for (const auto &w : which) {
        const fs::path dir = out / w;
        fs::create_directories(dir);
        std::printf("scenario %s\n", w.c_str());
        std::fflush(stdout);
        try {
            if (w == "first_order") scenario_first_order(dir);
            else if (w == "abl_ko") scenario_abl_ko(dir);
            else if (w == "nonlinear_euler") scenario_nonlinear_euler(dir);
            else if (w == "tensor") scenario_tensor(dir);
            else if (w == "level_euler") scenario_level_euler(dir);
            else if (w == "timer") scenario_timer(dir);
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
