// dropin_host.cpp -- host-only checks of the C++ drop-in layer (no device
// calls): right-hand-side recognition, model -> descriptor mapping, tensor
// alias maps, the expression parser and the plotfile writer; plus an `eval`
// mode that evaluates a decorated registered RHS on a state read from a file,
// for comparison with the CPU oracle.
//
//   dropin_host selftest
//   dropin_host eval <case> <nx> <ny> <nz> <px> <py> <pz> <bc> <in.bin> <out.bin>
//   dropin_host expr <file>   (lines "x y z<TAB>expression"; prints each value)
// The eval input is the valid state a[((c*nz+k)*ny+j)*nx+i] (doubles); ghost
// cells are filled like the library's FillPatch (periodic images, then the
// boundary functor outside non-periodic faces, with Dirichlet value 0).
#include "treesap.hpp"
#include "user_rhs_traits.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

using namespace treesap;

static int g_failures = 0;
#define EXPECT(cond)                                                                                                   \
    do {                                                                                                               \
        if (!(cond)) {                                                                                                 \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);                                     \
            ++g_failures;                                                                                              \
        }                                                                                                              \
    } while (0)

template <class F> static bool throws(F &&f) {
    try {
        f();
    } catch (const std::exception &) {
        return true;
    }
    return false;
}

static amrex::Geometry cube(int n, double lo, double hi, int per) {
    const int p[3] = {per, per, per};
    const amrex::RealBox rb({lo, lo, lo}, {hi, hi, hi});
    return amrex::Geometry(amrex::Box({0, 0, 0}, {n - 1, n - 1, n - 1}), &rb, amrex::CoordSys::cartesian, p);
}

struct Quadratic {
    constexpr amrex::Real operator()(amrex::Real s) const { return s * s; }
};
struct Smoothstep {
    constexpr amrex::Real operator()(amrex::Real s) const { return s * s * (3.0 - 2.0 * s); }
};

static void test_aliases() {
    using namespace tensor;
    // Sym2: xx, yx, yy, zx, zy, zz
    EXPECT(SymRank2Alias<0>::get_relative({{0, 0}}) == 0);
    EXPECT(SymRank2Alias<0>::get_relative({{1, 0}}) == 1 && SymRank2Alias<0>::get_relative({{0, 1}}) == 1);
    EXPECT(SymRank2Alias<0>::get_relative({{1, 1}}) == 2);
    EXPECT(SymRank2Alias<0>::get_relative({{2, 0}}) == 3 && SymRank2Alias<0>::get_relative({{2, 1}}) == 4);
    EXPECT(SymRank2Alias<0>::get_relative({{2, 2}}) == 5);
    EXPECT(SymRank2Alias<6>::get_absolute({{2, 1}}) == 10 && SymRank2Alias<6>::next == 12);
    // RankN: first index fastest
    EXPECT(Rank2Alias<0>::get_relative({{2, 1}}) == 2 + 3 * 1);
    EXPECT(Rank3Alias<1>::get_absolute({{1, 2, 0}}) == 1 + 1 + 3 * 2);
    EXPECT(VectorAlias<1>::restore_absolute(3)[0] == 2);
    static_assert(ScalarAlias<0>::next == 1 && VectorAlias<1>::next == 4 && Rank4Alias<0>::ncomp == 81);
    // contraction order and operations on constants
    const double c = contract<2>([]<int a, int b>() { return double(10 * a + b); });
    EXPECT(c == 99.0); // sum_{a,b} (10a + b) = 90 + 9
    EXPECT((symmetrise<0, 1>([]<int a, int b>() { return double(a - b); })) == 0.0);
    EXPECT((antisymmetrise<0, 1>([]<int a, int b>() { return double(a - b); })) == -1.0);
    EXPECT((tensor::levi_civita<2, 1, 0>) == -1 && (tensor::kronecker<1, 1>) == 1);
}

static void test_stencil_weights() {
    using namespace stencil_impl::weights;
    EXPECT(first<1>(1) == 0.5 && first<1>(-1) == -0.5);
    EXPECT(first<2>(1) == 2.0 / 3.0 && first<2>(2) == -1.0 / 12.0);
    EXPECT(second<1>(0) == -2.0 && second<2>(0) == -2.5 && second<2>(2) == -1.0 / 12.0);
    EXPECT(third<2>(1) == -1.0 && third<2>(2) == 0.5);
    // polynomial exactness of D1 (degree 2n) and D2 (2n+1) on a 1-D field
    std::vector<double> f(9 * 9 * 9);
    const amrex::Array4<double> a(f.data(), {-4, -4, -4}, {5, 5, 5}, 1);
    for (int k = -4; k < 5; ++k)
        for (int j = -4; j < 5; ++j)
            for (int i = -4; i < 5; ++i) a(i, j, k) = std::pow(0.5 * i, 4) + 0.25 * j;
    const amrex::Array4<const double> ca(a);
    const auto g = cube(2, 0.0, 1.0, 1); // dx = 0.5
    // d/dx (x^4) at x = 0.5 is 4 * 0.125 = 0.5; n = 2 is exact for degree 4
    EXPECT(std::fabs(stencil::derivative<2, 0>(ca, 0, amrex::IntVect(1, 0, 0), g.data()) - 0.5) < 1e-13);
    EXPECT(std::fabs(stencil::derivative<2, const double, 1>(ca, 0, amrex::IntVect(1, 0, 0), g.data()) - 0.5) < 1e-13); // 0.25 j = 0.5 y
    EXPECT(std::fabs(stencil::derivative<2, 0, 0>(ca, 0, amrex::IntVect(1, 0, 0), g.data()) - 3.0) < 1e-12);
}

static void test_recognition() {
    const auto g = cube(16, 0.0, 1.0, 1);
    using device::recognise;
    // system_traits mappings of user functors, verified bit for bit over the domain
    auto r1 = recognise(FirstOrderWave<1>{}, g.data(), 2);
    EXPECT(r1.system == TSP_SYS_WAVE_PHIU && r1.stencil_n == 1);
    auto r3 = recognise(FirstOrderWave<3>{}, g.data(), 2);
    EXPECT(r3.system == TSP_SYS_WAVE_PHIU && r3.stencil_n == 3);
    auto r2 = recognise(ScalarWave<2, 3>{}, g.data(), 2);
    EXPECT(r2.system == TSP_SYS_WAVE_PHIPI && r2.stencil_n == 2);
    auto r4 = recognise(ScalarWave<4, 4>{}, g.data(), 2);
    EXPECT(r4.system == TSP_SYS_WAVE_PHIPI && r4.stencil_n == 4);
    auto rn = recognise(SelfDampedWave<2>{0.3}, g.data(), 2);
    EXPECT(rn.system == TSP_SYS_NONLINEAR && rn.stencil_n == 2 && rn.gamma == 0.3);
    auto rt = recognise(VacuumADM<3, 3>{}, g.data(), 2);
    EXPECT(rt.system == TSP_SYS_TENSOR_ADM && rt.stencil_n == 3);
    auto rt2 = recognise(VacuumADM<1, 2>{}, g.data(), 2);
    EXPECT(rt2.system == TSP_SYS_TENSOR_ADM && rt2.stencil_n == 1);
    // wrong mappings are rejected: another operator, a time dependence, and a
    // position-dependent source that matches the registered system at every
    // cell of the lower half of the box (the corner region a local check sees)
    EXPECT(throws([&] { recognise(Advection<1>{}, g.data(), 2); }));
    EXPECT(throws([&] { recognise(ForcedWave<2>{}, g.data(), 2); }));
    EXPECT(throws([&] { recognise(PatchedWave<2>{}, g.data(), 2); }));
    {
        std::vector<double> f(12 * 12 * 12 * 2);
        for (std::size_t q = 0; q < f.size(); ++q) f[q] = std::sin(0.37 * (double)q);
        const amrex::Array4<const double> arr(f.data(), {-2, -2, -2}, {10, 10, 10}, 2);
        bool same = true;
        for (int k = 0; k < 8; ++k)
            for (int j = 0; j < 8; ++j)
                for (int i = 0; i < 8; ++i) {
                    PatchedWave<2>::Workspace w1{};
                    systems::PhiPiWaveRHS<2, 2>::Workspace w2{};
                    const double a = PatchedWave<2>{}.operator()<1>({i, j, k}, 0.0, arr, g.data(), w1);
                    const double b = systems::PhiPiWaveRHS<2, 2>{}.operator()<1>({i, j, k}, 0.0, arr, g.data(), w2);
                    same = same && a == b;
                }
        EXPECT(same);
    }
    // library functors map without probing
    auto rl = recognise(systems::NonLinearWaveRHS<3, 4>{0.7}, g.data(), 2);
    EXPECT(rl.system == TSP_SYS_NONLINEAR && rl.stencil_n == 3 && rl.gamma == 0.7);
    static_assert(!device::is_mapped<CurlWave<3>> && device::is_mapped<ScalarWave<2, 2>>);
}

static void test_descriptors() {
    // ABL(KO(user)) through ABLPDEModel with a quadratic profile, as in the 256^3 workload
    {
        using RHS = KODecorator<ScalarWave<4, 4>, 4>;
        using Model = ABLPDEModel<RHS, ConstInit<2>, Quadratic>;
        const auto g = cube(64, 0.0, 1.0, 0);
        Model m({"phi", "Pi"}, g, 0.25 / 64, RHS{{}, 0.2},
                typename ABLDecorator<RHS, Quadratic>::ABLData({{0.0, 0.0}}, 16, 2.0, Quadratic{}), ConstInit<2>{0.0});
        const tsp_desc d = device::make_desc(m);
        EXPECT(d.system == TSP_SYS_WAVE_PHIPI && d.stencil_n == 4 && d.ghost == 4);
        EXPECT(d.ko_order == 4 && d.ko_strength == 0.2);
        EXPECT(d.abl_thickness == 16 && d.abl_max == 2.0 && d.abl_profile == TSP_PROFILE_QUADRATIC);
        EXPECT(d.bc == TSP_BC_DIRICHLET && d.target[0] == 0.0 && d.target[1] == 0.0);
        EXPECT(d.n[0] == 64 && d.n[2] == 64 && d.periodic[0] == 0 && d.prob_hi[1] == 1.0);
        EXPECT(d.dt == 0.25 / 64);
    }
    // periodic, KO order above the stencil half-width, NoOp BC
    {
        using RHS = KODecorator<ScalarWave<2, 3>, 3>;
        using Model = PDEModel<RHS, ConstInit<2>>;
        Model m({"phi", "Pi"}, cube(32, -1.0, 1.0, 1), 0.01, RHS{{}, 0.1}, ConstInit<2>{1.0});
        const tsp_desc d = device::make_desc(m);
        EXPECT(d.system == TSP_SYS_WAVE_PHIPI && d.stencil_n == 2 && d.ghost == 3 && d.ko_order == 3);
        EXPECT(d.bc == TSP_BC_NOOP && d.abl_thickness == 0 && d.periodic[1] == 1 && d.prob_lo[0] == -1.0);
    }
    // Dirichlet values and tensor system
    {
        using RHS = VacuumADM<2, 2>;
        using BC = DirichletBC<12, 2>;
        BC bc{};
        for (int c = 0; c < 12; ++c) bc.boundary_values[c] = 0.5 * c;
        PDEModel<RHS, ConstInit<12>, BC> m({}, cube(24, 0.0, 1.0, 0), 0.001, RHS{}, ConstInit<12>{0.0}, bc);
        const tsp_desc d = device::make_desc(m);
        EXPECT(d.system == TSP_SYS_TENSOR_ADM && d.bc == TSP_BC_DIRICHLET && d.target[11] == 5.5);
    }
    // a functor the library carries no kernel for maps onto the generic kernel, decorators included
    {
        using RHS = KODecorator<CurlWave<3>, 2>;
        using BC = DirichletBC<4, 3>;
        PDEModel<RHS, ConstInit<4>, BC> m({}, cube(20, 0.0, 1.0, 0), 0.002, RHS{{}, 0.1}, ConstInit<4>{0.0},
                                          BC{0.5, 0.25, 0.0, -1.0});
        const tsp_desc d = device::make_desc(m);
        EXPECT(d.system == TSP_SYS_CUSTOM && d.ncomp == 4 && d.ghost == 3 && d.ko_order == 0);
        EXPECT(d.abl_thickness == 0 && d.bc == TSP_BC_DIRICHLET && d.target[0] == 0.5 && d.target[3] == -1.0);
        EXPECT(d.dt == 0.002 && d.n[1] == 20);
    }
    // unsupported profile and mismatched targets are rejected
    {
        using RHS = ScalarWave<2, 2>;
        using Model = ABLPDEModel<RHS, ConstInit<2>, Smoothstep>;
        Model m({"phi", "Pi"}, cube(32, 0.0, 1.0, 0), 0.01, RHS{},
                typename ABLDecorator<RHS, Smoothstep>::ABLData({{0.0, 0.0}}, 4, 1.0, Smoothstep{}), ConstInit<2>{0.0});
        EXPECT(throws([&] { device::make_desc(m); }));
    }
    {
        using RHS = ABLDecorator<ScalarWave<2, 2>, Quadratic>;
        using Model = PDEModel<RHS, ConstInit<2>, DirichletBC<2, 2>>;
        const auto g = cube(32, 0.0, 1.0, 0);
        Model m({"phi", "Pi"}, g, 0.01,
                RHS(ScalarWave<2, 2>{}, g, RHS::ABLData({{0.0, 0.0}}, 4, 1.0, Quadratic{})), ConstInit<2>{0.0},
                DirichletBC<2, 2>{{1.0, 0.0}});
        EXPECT(throws([&] { device::make_desc(m); }));
    }
}

static void test_function_init() {
    FunctionInit<3> f({"exp(-(x^2 + y^2 + z^2)/0.1^2)", "-x^2 + 2^3^2 - 3*y/2", "atan2(y, x) + pow(z, 2) + max(x, -1)"});
    const double x = 0.3, y = -0.2, z = 0.7;
    EXPECT(f.operator()<0>(x, y, z) == std::exp(-(std::pow(x, 2.0) + std::pow(y, 2.0) + std::pow(z, 2.0)) / std::pow(0.1, 2.0)));
    EXPECT(f.operator()<1>(x, y, z) == -std::pow(x, 2.0) + std::pow(2.0, std::pow(3.0, 2.0)) - 3 * y / 2);
    EXPECT(f.operator()<2>(x, y, z) == std::atan2(y, x) + std::pow(z, 2.0) + std::fmax(x, -1.0));
    EXPECT(throws([] { FunctionInit<1>({"x + (y"}); }));
    EXPECT(throws([] { FunctionInit<1>({"w * 2"}); }));
    EXPECT(throws([] { FunctionInit<1>({"foo(x)"}); }));
    static_assert(std::is_trivially_copyable_v<FunctionInit<3>>);
}

static void test_plotfile(const std::string &dir) {
    const auto g = cube(4, 0.0, 2.0, 0);
    std::vector<double> state(2 * 64);
    for (std::size_t i = 0; i < state.size(); ++i) state[i] = 0.25 * (double)i - 3.0;
    PlotExporter px([&](std::vector<double> &v) { v = state; }, g, {"phi", "Pi"}, dir + "/plt", 2);
    px.exportIfNecessary(1, 0.5); // not due
    px.exportIfNecessary(4, 0.5);
    EXPECT(!std::filesystem::exists(dir + "/plt00001"));
    EXPECT(std::filesystem::exists(dir + "/plt00004/Header"));
    std::ifstream h(dir + "/plt00004/Header");
    std::string line;
    std::getline(h, line);
    EXPECT(line == "HyperCLaw-V1.1");
    std::getline(h, line);
    EXPECT(line == "2");
    std::ifstream fab(dir + "/plt00004/Level_0/Cell_D_00000", std::ios::binary);
    std::getline(fab, line);
    EXPECT(line == "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))((0,0,0) (3,3,3) (0,0,0)) 2");
    std::vector<double> back(state.size());
    fab.read(reinterpret_cast<char *>(back.data()), (std::streamsize)(back.size() * sizeof(double)));
    EXPECT(back == state);
}

// ---- eval mode ---------------------------------------------------------------

template <class RHS> struct EvalCase {
    static constexpr int G = RHS::ngrow;
    static void run(const RHS &rhs, const amrex::Geometry &geom, int bc, const std::vector<double> &in,
                    std::vector<double> &out) {
        constexpr int nc = RHS::ncomp;
        const amrex::Box dom = geom.Domain();
        const int nx = dom.length(0), ny = dom.length(1), nz = dom.length(2);
        const amrex::Dim3 b{-G, -G, -G}, e{nx + G, ny + G, nz + G};
        std::vector<double> reg((std::size_t)(nx + 2 * G) * (ny + 2 * G) * (nz + 2 * G) * nc, 0.0);
        const amrex::Array4<double> a(reg.data(), b, e, nc);
        for (int c = 0; c < nc; ++c)
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) a(i, j, k, c) = in[(((std::size_t)c * nz + k) * ny + j) * nx + i];
        const int n[3] = {nx, ny, nz};
        for (int k = -G; k < nz + G; ++k)
            for (int j = -G; j < ny + G; ++j)
                for (int i = -G; i < nx + G; ++i) {
                    int s[3] = {i, j, k};
                    bool inside = true, phys = false;
                    for (int d = 0; d < 3; ++d)
                        if (s[d] < 0 || s[d] >= n[d]) {
                            inside = false;
                            if (geom.isPeriodic(d)) s[d] += s[d] < 0 ? n[d] : -n[d];
                            else phys = true;
                        }
                    if (inside) continue;
                    for (int c = 0; c < nc; ++c)
                        if (phys) {
                            if (bc == 1) a(i, j, k, c) = 0.0;
                        } else {
                            a(i, j, k, c) = a(s[0], s[1], s[2], c);
                        }
                }
        const amrex::Array4<const double> ca(a);
        out.assign((std::size_t)nx * ny * nz * nc, 0.0);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    typename RHS::Workspace ws{};
                    static_for<nc>([&]<int c>() {
                        out[(((std::size_t)c * nz + k) * ny + j) * nx + i] =
                            rhs.template operator()<c>(amrex::IntVect(i, j, k), 0.0, ca, geom.data(), ws);
                    });
                }
    }
};

// decorated registered systems, named by case id (the tests use the same list)
static bool eval_case(const std::string &id, const amrex::Geometry &geom, int bc, const std::vector<double> &in,
                      std::vector<double> &out) {
    auto abl = [&](auto inner, int T, double mx) {
        using Inner = decltype(inner);
        using D = ABLDecorator<Inner, Quadratic>;
        amrex::GpuArray<double, Inner::ncomp> tgt{};
        for (int c = 0; c < Inner::ncomp; ++c) tgt[c] = 0.0;
        return D(inner, geom, typename D::ABLData(tgt, T, mx, Quadratic{}));
    };
    using namespace systems;
    if (id == "phipi_n2_g3_ko3_abl4") {
        EvalCase<decltype(abl(KODecorator<PhiPiWaveRHS<2, 3>, 3>{{}, 0.1}, 4, 2.0))>::run(
            abl(KODecorator<PhiPiWaveRHS<2, 3>, 3>{{}, 0.1}, 4, 2.0), geom, bc, in, out);
    } else if (id == "phipi_n4_g4_ko4_abl5") {
        EvalCase<decltype(abl(KODecorator<PhiPiWaveRHS<4, 4>, 4>{{}, 0.2}, 5, 2.0))>::run(
            abl(KODecorator<PhiPiWaveRHS<4, 4>, 4>{{}, 0.2}, 5, 2.0), geom, bc, in, out);
    } else if (id == "phiu_n1_g1_ko1") {
        EvalCase<KODecorator<PhiUWaveRHS<1, 1>, 1>>::run({{}, 0.05}, geom, bc, in, out);
    } else if (id == "phiu_n3_g3_abl3") {
        EvalCase<decltype(abl(PhiUWaveRHS<3, 3>{}, 3, 1.5))>::run(abl(PhiUWaveRHS<3, 3>{}, 3, 1.5), geom, bc, in, out);
    } else if (id == "nonlinear_n2_g2_ko2") {
        EvalCase<KODecorator<NonLinearWaveRHS<2, 2>, 2>>::run({{0.3}, 0.1}, geom, bc, in, out);
    } else if (id == "tensor_n2_g2_ko2_abl3") {
        EvalCase<decltype(abl(KODecorator<LinearADMRHS<2, 2>, 2>{{}, 0.1}, 3, 1.0))>::run(
            abl(KODecorator<LinearADMRHS<2, 2>, 2>{{}, 0.1}, 3, 1.0), geom, bc, in, out);
    } else if (id == "tensor_n3_g3") {
        EvalCase<LinearADMRHS<3, 3>>::run({}, geom, bc, in, out);
    } else {
        return false;
    }
    return true;
}

int main(int argc, char **argv) {
    const std::string mode = argc > 1 ? argv[1] : "selftest";
    if (mode == "selftest") {
        const std::string dir = argc > 2 ? argv[2] : "/tmp";
        test_aliases();
        test_stencil_weights();
        test_recognition();
        test_descriptors();
        test_function_init();
        test_plotfile(dir);
        if (g_failures) {
            std::fprintf(stderr, "%d check(s) failed\n", g_failures);
            return 1;
        }
        std::printf("drop-in host self-test passed\n");
        return 0;
    }
    if (mode == "eval" && argc == 12) {
        const std::string id = argv[2];
        const int nx = std::atoi(argv[3]), ny = std::atoi(argv[4]), nz = std::atoi(argv[5]);
        const int per[3] = {std::atoi(argv[6]), std::atoi(argv[7]), std::atoi(argv[8])};
        const int bc = std::atoi(argv[9]);
        const amrex::RealBox rb({0.0, 0.0, 0.0}, {1.0, 1.0, 1.0});
        const amrex::Geometry geom(amrex::Box({0, 0, 0}, {nx - 1, ny - 1, nz - 1}), &rb, 0, per);
        std::ifstream fin(argv[10], std::ios::binary);
        std::vector<double> in;
        double v;
        while (fin.read(reinterpret_cast<char *>(&v), sizeof v)) in.push_back(v);
        std::vector<double> out;
        if (!eval_case(id, geom, bc, in, out)) {
            std::fprintf(stderr, "unknown case %s\n", id.c_str());
            return 2;
        }
        if (out.size() != in.size()) {
            std::fprintf(stderr, "size mismatch: %zu values in, %zu out\n", in.size(), out.size());
            return 2;
        }
        std::ofstream fout(argv[11], std::ios::binary);
        fout.write(reinterpret_cast<const char *>(out.data()), (std::streamsize)(out.size() * sizeof(double)));
        return fout ? 0 : 1;
    }
    if (mode == "expr" && argc == 3) {
        // one "x y z<TAB>expression" per line of the file; prints the value of each
        // and, after a tab, its postfix device program as "code value" pairs
        std::ifstream fin(argv[2]);
        std::string line;
        while (std::getline(fin, line)) {
            const auto tab = line.find('\t');
            if (tab == std::string::npos) return 2;
            double x = 0, y = 0, z = 0;
            if (std::sscanf(line.substr(0, tab).c_str(), "%lf %lf %lf", &x, &y, &z) != 3) return 2;
            try {
                const FunctionInit<1> f({line.substr(tab + 1)});
                std::vector<tsp_expr_op> ops;
                std::vector<std::int32_t> start;
                f.program(ops, start);
                std::printf("%.17g\t", f.operator()<0>(x, y, z));
                for (const auto &op : ops) std::printf(" %d %.17g", op.code, op.value);
                std::printf("\n");
            } catch (const std::exception &e) {
                std::printf("error %s\n", e.what());
            }
        }
        return 0;
    }
    std::fprintf(stderr, "usage: %s selftest [dir] | eval <case> nx ny nz px py pz bc in out | expr <file>\n", argv[0]);
    return 2;
}
