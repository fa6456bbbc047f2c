// dropin_generic.cpp -- TreESaP-style programs whose right-hand sides the
// device library does not carry: compiled with hipcc, the drop-in integrators
// run them on the stage kernel compiled from the functor (device/generic.hpp).
// Each scenario writes into <outdir>/<scenario>/:
//   init.bin   the state after initialise(), a[((c*nz+k)*ny+j)*nx+i]
//   final.bin  the state after the device steps
//   host.bin   the same steps of the same functor on the host (host_rk4.hpp),
//              for scenarios without an oracle system
//   meta.txt   "<integrator> <steps> <time as %.17g> <system id of the descriptor>"
//
//   dropin_generic <outdir> [scenario ...]
#include "treesap.hpp"
#include "host_rk4.hpp"
#include "user_rhs.hpp"

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
    char buf[160];
    std::snprintf(buf, sizeof buf, "%s %zu %.17g %d\n", kind, integ.get_steps(), integ.get_time(),
                  integ.device_solver().desc().system);
    m << buf;
}

static amrex::Geometry box_geom(int nx, int ny, int nz, std::array<double, 3> hi, std::array<int, 3> per) {
    const amrex::RealBox rb({0.0, 0.0, 0.0}, hi);
    return amrex::Geometry(amrex::Box({0, 0, 0}, {nx - 1, ny - 1, nz - 1}), &rb, amrex::CoordSys::cartesian, per.data());
}

// FirstOrderWave (examples/01-wave-eqn style, no system_traits here) with KO,
// Dirichlet on a non-periodic box: the oracle's (phi, u) system
static void scenario_first_order(const fs::path &dir) {
    const auto geom = box_geom(24, 20, 22, {1.0, 1.0, 1.0}, {0, 0, 0});
    using RHSType = KODecorator<FirstOrderWave<2>, 2>;
    using BCType = DirichletBC<4, 2>;
    using ModelType = PDEModel<RHSType, FunctionInit<4>, BCType>;
    FunctionInit<4> init({"exp(-((x-0.5)^2 + (y-0.45)^2 + (z-0.55)^2)/0.15^2)", "0.1*x*y", "0.0", "-0.05*z"});
    ModelType model({"phi", "u_0", "u_1", "u_2"}, geom, 0.25 / 24, RHSType{{}, 0.1}, init,
                    BCType{0.1, -0.2, 0.3, 0.05});
    SingleLevelRK4Integrator<ModelType> integ(std::move(model), {64, 64, 64});
    static_assert(!device::runs_library_kernels<ModelType>);
    integ.initialise();
    const auto u0 = integ.get_state();
    for (int s = 0; s < 3; ++s) integ.step();
    integ.advance(3);
    dump(dir, integ, u0, "rk4");
}

// ABL(KO(linearised ADM)) through ABLPDEModel: the oracle's tensor system
static void scenario_abl_tensor(const fs::path &dir) {
    const auto geom = box_geom(20, 18, 16, {1.0, 0.9, 0.8}, {0, 0, 0});
    using RHSType = KODecorator<VacuumADM<2, 2>, 2>;
    using ModelType = ABLPDEModel<RHSType, FunctionInit<12>, Quadratic>;
    using Data = typename ABLDecorator<RHSType, Quadratic>::ABLData;
    amrex::Array<std::string, 12> exprs;
    for (int c = 0; c < 12; ++c)
        exprs[c] = "0.01*sin(2*pi*(" + std::to_string(c + 1) + "*x + y)) * cos(2*pi*" + std::to_string((c % 3) + 1) + "*z)";
    amrex::GpuArray<amrex::Real, 12> tgt{};
    ModelType model({}, geom, 0.2 / 20, RHSType{{}, 0.1}, Data(tgt, 3, 1.5, Quadratic{}), FunctionInit<12>(exprs));
    SingleLevelRK4Integrator<ModelType> integ(std::move(model), {64, 64, 64});
    integ.initialise();
    const auto u0 = integ.get_state();
    integ.advance(3);
    dump(dir, integ, u0, "rk4");
}

// nonlinear wave with a workspace temporary, KO, periodic ragged box, forward Euler
static void scenario_nonlinear_euler(const fs::path &dir) {
    const auto geom = box_geom(26, 22, 18, {1.0, 0.9, 0.8}, {1, 1, 1});
    using RHSType = KODecorator<SelfDampedWave<2>, 2>;
    using ModelType = PDEModel<RHSType, FunctionInit<2>>;
    FunctionInit<2> init({"sin(2*pi*x)*cos(2*pi*y/0.9) + 0.3*sin(2*pi*z/0.8)", "0.5*cos(2*pi*(x + y))"});
    ModelType model({"phi", "pi"}, geom, 0.004, RHSType{{0.05}, 0.1}, init);
    SingleLevelEulerIntegrator<ModelType> integ(std::move(model), {64, 64, 64});
    integ.initialise();
    const auto u0 = integ.get_state();
    for (int s = 0; s < 5; ++s) integ.step();
    dump(dir, integ, u0, "euler");
}

// third derivatives, curl (Levi-Civita), antisymmetric part, cyclic sum,
// position and time dependence; mixed periodicity with Dirichlet values: the
// host stepping of the same functor is the reference
static void scenario_curl(const fs::path &dir, bool euler) {
    const auto geom = box_geom(22, 18, 20, {1.1, 0.9, 1.0}, {1, 0, 1});
    using RHSType = CurlWave<3>;
    using BCType = DirichletBC<4, 3>;
    using ModelType = PDEModel<RHSType, FunctionInit<4>, BCType>;
    FunctionInit<4> init({"exp(-((x-0.5)^2 + (y-0.4)^2 + (z-0.6)^2)/0.2^2)", "0.2*sin(2*pi*z)", "0.1*x",
                          "0.3*cos(2*pi*x)*y"});
    const BCType bc{0.01, 0.02, -0.03, 0.04};
    const double dt = 0.2 / 22;
    ModelType model({"phi", "E_0", "E_1", "E_2"}, geom, dt, RHSType{}, init, bc);
    std::vector<double> u0;
    std::size_t steps = 0;
    if (euler) {
        SingleLevelEulerIntegrator<ModelType> integ(ModelType(model), {64, 64, 64});
        integ.initialise();
        u0 = integ.get_state();
        for (int s = 0; s < 4; ++s) integ.step();
        dump(dir, integ, u0, "euler");
        steps = integ.get_steps();
    } else {
        SingleLevelRK4Integrator<ModelType> integ(ModelType(model), {64, 64, 64});
        integ.initialise();
        u0 = integ.get_state();
        for (int s = 0; s < 2; ++s) integ.step();
        integ.advance(3);
        dump(dir, integ, u0, "rk4");
        steps = integ.get_steps();
    }
    HostStepper<RHSType> host(RHSType{}, geom, dt, true, bc.boundary_values.data());
    host.set_state(u0);
    for (std::size_t s = 0; s < steps; ++s) euler ? host.euler_step() : host.rk4_step();
    write_bin(dir / "host.bin", host.get_state());
}

// a wave whose source term lives in the upper half of the box, with an
// absorbing layer: the host stepping is the reference
static void scenario_patched(const fs::path &dir) {
    const auto geom = box_geom(24, 16, 18, {1.0, 1.0, 1.0}, {0, 1, 0});
    using RHSType = PatchedWave<2>;
    using ModelType = ABLPDEModel<RHSType, FunctionInit<2>, Quadratic>;
    using Data = typename ABLDecorator<RHSType, Quadratic>::ABLData;
    FunctionInit<2> init({"exp(-((x-0.6)^2 + (y-0.5)^2 + (z-0.5)^2)/0.2^2)", "0.0"});
    const double dt = 0.25 / 24;
    ModelType model({"phi", "Pi"}, geom, dt, RHSType{}, Data({{0.0, 0.0}}, 4, 2.0, Quadratic{}), init);
    const auto rhs = model.get_rhs();
    SingleLevelRK4Integrator<ModelType> integ(std::move(model), {64, 64, 64});
    integ.initialise();
    const auto u0 = integ.get_state();
    integ.advance(5);
    dump(dir, integ, u0, "rk4");
    const double zero[2] = {0.0, 0.0};
    HostStepper<std::remove_cv_t<decltype(rhs)>> host(rhs, geom, dt, true, zero);
    host.set_state(u0);
    for (int s = 0; s < 5; ++s) host.rk4_step();
    write_bin(dir / "host.bin", host.get_state());
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <outdir> [scenario ...]\n", argv[0]);
        return 2;
    }
    const fs::path out = argv[1];
    std::vector<std::string> which(argv + 2, argv + argc);
    if (which.empty()) which = {"first_order", "abl_tensor", "nonlinear_euler", "curl_rk4", "curl_euler", "patched"};
    for (const auto &w : which) {
        const fs::path dir = out / w;
        fs::create_directories(dir);
        std::printf("scenario %s\n", w.c_str());
        std::fflush(stdout);
        try {
            if (w == "first_order") scenario_first_order(dir);
            else if (w == "abl_tensor") scenario_abl_tensor(dir);
            else if (w == "nonlinear_euler") scenario_nonlinear_euler(dir);
            else if (w == "curl_rk4") scenario_curl(dir, false);
            else if (w == "curl_euler") scenario_curl(dir, true);
            else if (w == "patched") scenario_patched(dir);
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
