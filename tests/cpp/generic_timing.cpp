// generic_timing.cpp -- throughput of the stage kernel compiled from a user
// functor (device/generic.hpp): the (phi, Pi) wave of ScalarWave<4, 4> under
// KODecorator<., 4> (strength 0.2) and an ABLPDEModel layer of 16 cells
// (quadratic profile, max 2.0) on an N^3 non-periodic unit cube with dt = dx/4;
// the same arithmetic as the registered wave256 workload, run through the
// generic kernel because ScalarWave has no system_traits mapping. Prints one
// JSON line: cell-updates/s of graph-replayed RK4 steps (host clock around
// `steps` steps between device synchronisations, best of three).
//
//   generic_timing [N=256] [steps=20]
#include "treesap.hpp"
#include "user_rhs.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

using namespace treesap;

struct Quadratic {
    constexpr amrex::Real operator()(amrex::Real s) const { return s * s; }
};

int main(int argc, char **argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 256;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 20;
    const amrex::RealBox rb({0.0, 0.0, 0.0}, {1.0, 1.0, 1.0});
    int per[3] = {0, 0, 0};
    const amrex::Geometry geom(amrex::Box({0, 0, 0}, {n - 1, n - 1, n - 1}), &rb, amrex::CoordSys::cartesian, per);
    using RHSType = KODecorator<ScalarWave<4, 4>, 4>;
    using ModelType = ABLPDEModel<RHSType, FunctionInit<2>, Quadratic>;
    using Data = typename ABLDecorator<RHSType, Quadratic>::ABLData;
    static_assert(!device::runs_library_kernels<ModelType>, "the functor must run on the generic kernel");
    try {
        FunctionInit<2> init({"exp(-((x-0.5)^2 + (y-0.45)^2 + (z-0.55)^2)/0.1^2)", "0.0"});
        ModelType model({"phi", "Pi"}, geom, 0.25 / n, RHSType{{}, 0.2}, Data({{0.0, 0.0}}, 16, 2.0, Quadratic{}),
                        init);
        SingleLevelRK4Integrator<ModelType> integ(std::move(model), {n, n, n});
        integ.initialise();
        integ.advance(4); // captures the graphs of both step parities
        if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("hipDeviceSynchronize");
        double best = 1e300;
        for (int rep = 0; rep < 3; ++rep) {
            const auto t0 = std::chrono::steady_clock::now();
            integ.advance(steps);
            if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("hipDeviceSynchronize");
            best = std::min(best, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        }
        const auto st = integ.get_state();
        bool finite = true;
        for (double v : st) finite = finite && std::isfinite(v);
        std::printf("{\"cells\": %lld, \"steps\": %d, \"cell_updates_per_s\": %.6e, \"ms_per_step\": %.6f, "
                    "\"finite\": %s}\n",
                    (long long)n * n * n, steps, (double)n * n * n * steps / best, 1e3 * best / steps,
                    finite ? "true" : "false");
    } catch (const std::exception &e) {
        std::fprintf(stderr, "generic_timing: %s\n", e.what());
        return 1;
    }
    return 0;
}
