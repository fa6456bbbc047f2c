// model_map.hpp -- turns a TreESaP PDE model into the descriptor of the C ABI
// (tsp_desc), i.e. decides which device kernel runs the model.
//
// A model's right-hand side is a compile-time functor (RHSConcept,
// pde-model/PDEModelConcepts.hpp:44-64). The device library carries kernels
// for the systems of systems/registered_rhs.hpp, each with derivative
// half-width n = 1..4, optionally wrapped as ABLDecorator<KODecorator<user>>
// (the nesting ABLPDEModel produces when handed a KODecorator). The mapping:
//
//   decorators   unwrapped at compile time: KODecorator<R, r> gives ko_order r
//                and its strength, ABLDecorator<R, P> its thickness, maximum,
//                targets and profile. Other nestings (KO outside ABL, repeated
//                decorators) are rejected at compile time: their rounding
//                differs from the kernels'.
//   profile      must equal s, s*s or s*s*s on every argument the layer uses
//                (checked on the host, bit for bit).
//   user RHS     either named through a specialisation of system_traits, or
//                recognised: the functor is evaluated on the host on a seeded
//                random field at 4^3 cells and two times, every component in
//                order with one Workspace per cell, and compared with each
//                registered functor of the same ncomp and every n <= ngrow
//                (NonLinear uses the functor's `gamma` member). A match within
//                1e-12 of the component's largest magnitude selects that
//                system; no match throws.
//   BC           DirichletBC -> TSP_BC_DIRICHLET with its values, NoOpBC ->
//                TSP_BC_NOOP; other functors are rejected at compile time. With
//                both an ABL and a DirichletBC the two value sets must agree
//                (the device keeps one target per component).
//   geometry     cells per axis, prob_lo/hi, periodicity; ghost = RHS::ngrow.
#pragma once

#include "../pde-model/ABLPDEModel.hpp"
#include "../pde-model/PDEModel.hpp"
#include "../pde-model/bcs/DirichletBC.hpp"
#include "../pde-model/bcs/NoOpBC.hpp"
#include "../pde-model/rhs-decorators/ABLDecorator.hpp"
#include "../pde-model/rhs-decorators/KODecorator.hpp"
#include "../systems/registered_rhs.hpp"
#include "../util/static_for.hpp"
#include "treesap_hip.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <typeinfo>
#include <vector>

namespace treesap {

// Optional explicit mapping of an (undecorated) RHS type onto a device system:
//   template <int g> struct treesap::system_traits<MyRHS<g>> {
//       static constexpr int system = TSP_SYS_WAVE_PHIU;   // a tsp_system id
//       static constexpr int stencil_n = g;                  // derivative half-width
//       static double gamma(const MyRHS<g> &) { return 0.0; } // NonLinear coefficient
//   };
// Types without one are recognised by probing (see above).
template <class RHS> struct system_traits {};

template <int n, int g> struct system_traits<systems::PhiPiWaveRHS<n, g>> {
    static constexpr int system = TSP_SYS_WAVE_PHIPI, stencil_n = n;
    static double gamma(const systems::PhiPiWaveRHS<n, g> &) { return 0.0; }
};
template <int n, int g> struct system_traits<systems::PhiUWaveRHS<n, g>> {
    static constexpr int system = TSP_SYS_WAVE_PHIU, stencil_n = n;
    static double gamma(const systems::PhiUWaveRHS<n, g> &) { return 0.0; }
};
template <int n, int g> struct system_traits<systems::NonLinearWaveRHS<n, g>> {
    static constexpr int system = TSP_SYS_NONLINEAR, stencil_n = n;
    static double gamma(const systems::NonLinearWaveRHS<n, g> &r) { return r.gamma; }
};
template <int n, int g> struct system_traits<systems::LinearADMRHS<n, g>> {
    static constexpr int system = TSP_SYS_TENSOR_ADM, stencil_n = n;
    static double gamma(const systems::LinearADMRHS<n, g> &) { return 0.0; }
};

} // namespace treesap

namespace treesap::device {

struct Recognised {
    int system = -1;
    int stencil_n = 0;
    double gamma = 0.0;
};

namespace detail {

template <class> constexpr bool always_false = false;

// SplitMix64 (Steele, Lea, Flood 2014), uniform doubles in [-1, 1).
struct SplitMix64 {
    std::uint64_t s;
    std::uint64_t next() {
        std::uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0; }
};

// Evaluate every component of `f` in order, one Workspace per cell, at the
// P^3 cells of a random field with H ghost layers.
template <class F, int H, int P>
std::vector<double> probe_values(const F &f, const std::vector<double> &field, const amrex::GeometryData &g, double t) {
    constexpr int nc = F::ncomp;
    const amrex::Array4<const double> arr(field.data(), {-H, -H, -H}, {P + H, P + H, P + H}, nc);
    std::vector<double> out((std::size_t)P * P * P * nc);
    for (int k = 0; k < P; ++k)
        for (int j = 0; j < P; ++j)
            for (int i = 0; i < P; ++i) {
                const amrex::IntVect iv(i, j, k);
                typename F::Workspace ws{};
                treesap::static_for<nc>([&]<int c>() {
                    out[(((std::size_t)c * P + k) * P + j) * P + i] = f.template operator()<c>(iv, t, arr, g, ws);
                });
            }
    return out;
}

template <class A, class B> bool probe_equal(const A &a, const B &b, const amrex::GeometryData &g) {
    static_assert(A::ncomp == B::ncomp);
    constexpr int nc = A::ncomp;
    constexpr int H = A::ngrow > B::ngrow ? A::ngrow : B::ngrow;
    constexpr int P = 4;
    constexpr int W = P + 2 * H;
    std::vector<double> field((std::size_t)W * W * W * nc);
    SplitMix64 rng{0x7265636f676e6973ull};
    for (double &v : field) v = rng.uniform();
    for (double t : {0.375, 1.90625}) {
        const auto va = probe_values<A, H, P>(a, field, g, t);
        const auto vb = probe_values<B, H, P>(b, field, g, t);
        const std::size_t per = (std::size_t)P * P * P;
        for (int c = 0; c < nc; ++c) {
            double scale = 0.0;
            for (std::size_t q = 0; q < per; ++q) scale = std::fmax(scale, std::fabs(vb[c * per + q]));
            for (std::size_t q = 0; q < per; ++q) {
                const double x = va[c * per + q], y = vb[c * per + q];
                if (!std::isfinite(x) || !(std::fabs(x - y) <= 1e-12 * scale)) return false;
            }
        }
    }
    return true;
}

template <class U, int n> Recognised try_half_width(const U &u, const amrex::GeometryData &g) {
    constexpr int G = U::ngrow;
    if constexpr (U::ncomp == 2) {
        if constexpr (requires { { u.gamma } -> std::convertible_to<double>; }) {
            const systems::NonLinearWaveRHS<n, G> cand{static_cast<double>(u.gamma)};
            if (probe_equal(u, cand, g)) return {TSP_SYS_NONLINEAR, n, cand.gamma};
        }
        if (probe_equal(u, systems::PhiPiWaveRHS<n, G>{}, g)) return {TSP_SYS_WAVE_PHIPI, n, 0.0};
    } else if constexpr (U::ncomp == 4) {
        if (probe_equal(u, systems::PhiUWaveRHS<n, G>{}, g)) return {TSP_SYS_WAVE_PHIU, n, 0.0};
    } else if constexpr (U::ncomp == 12) {
        if (probe_equal(u, systems::LinearADMRHS<n, G>{}, g)) return {TSP_SYS_TENSOR_ADM, n, 0.0};
    }
    return {};
}

} // namespace detail

// Which device system evaluates the undecorated right-hand side `u`.
template <class U> Recognised recognise(const U &u, const amrex::GeometryData &g) {
    if constexpr (requires { system_traits<U>::system; }) {
        return {system_traits<U>::system, system_traits<U>::stencil_n, system_traits<U>::gamma(u)};
    } else {
        Recognised r;
        treesap::static_for<1, 5>([&]<int n>() {
            if constexpr (n <= U::ngrow)
                if (r.system < 0) r = detail::try_half_width<U, n>(u, g);
        });
        if (r.system < 0)
            throw std::invalid_argument(
                std::string("treesap: the right-hand side ") + typeid(U).name() + " (ncomp " + std::to_string(U::ncomp) +
                ", ngrow " + std::to_string(U::ngrow) +
                ") matches none of the systems the device library carries: (phi,Pi) wave, (phi,u) wave, "
                "NonLinear (phi,pi), linearised ADM (h,K), derivative half-width 1..4. Write it with "
                "systems/registered_rhs.hpp's arithmetic or specialise treesap::system_traits for it.");
        return r;
    }
}

// ---- decorators ------------------------------------------------------------

template <class R> struct decorators {
    using user_t = R;
    static constexpr bool has_ko = false, has_abl = false;
    static const user_t &user(const R &r) { return r; }
    static void fill(const R &, tsp_desc &) {}
};

template <class In, int r> struct decorators<KODecorator<In, r>> {
    static_assert(!decorators<In>::has_ko && !decorators<In>::has_abl,
                  "treesap: the device kernels evaluate ABLDecorator<KODecorator<user RHS>>; a KODecorator must "
                  "wrap the user RHS directly");
    using user_t = typename decorators<In>::user_t;
    static constexpr bool has_ko = true, has_abl = false;
    static const user_t &user(const KODecorator<In, r> &d) { return decorators<In>::user(d.rhs); }
    static void fill(const KODecorator<In, r> &d, tsp_desc &desc) {
        desc.ko_order = r;
        desc.ko_strength = d.strength;
        decorators<In>::fill(d.rhs, desc);
    }
};

template <class P> int profile_id(const P &p, int thickness) {
    const double inv_t = 1.0 / static_cast<double>(thickness);
    const int ids[3] = {TSP_PROFILE_LINEAR, TSP_PROFILE_QUADRATIC, TSP_PROFILE_CUBIC};
    for (int id : ids) {
        bool same = true;
        for (int d2 = 1; d2 <= 3 * thickness * thickness && same; ++d2) {
            const double x = amrex::Clamp(inv_t * std::sqrt(static_cast<double>(d2)), 0.0, 1.0);
            const double want = id == TSP_PROFILE_LINEAR ? x : id == TSP_PROFILE_QUADRATIC ? x * x : x * x * x;
            const double got = static_cast<double>(p(x));
            same = std::memcmp(&want, &got, sizeof(double)) == 0;
        }
        if (same) return id;
    }
    throw std::invalid_argument("treesap: the ABL damping profile is none of s, s*s, s*s*s (the profiles the device "
                                "library tabulates)");
}

template <class In, class P> struct decorators<ABLDecorator<In, P>> {
    static_assert(!decorators<In>::has_abl, "treesap: one ABLDecorator per model");
    using user_t = typename decorators<In>::user_t;
    static constexpr bool has_ko = decorators<In>::has_ko, has_abl = true;
    static const user_t &user(const ABLDecorator<In, P> &d) { return decorators<In>::user(d.rhs); }
    static void fill(const ABLDecorator<In, P> &d, tsp_desc &desc) {
        if (d.data.thickness <= 0) throw std::invalid_argument("treesap: ABL thickness must be positive");
        if (d.data.inv_thickness != 1.0 / static_cast<double>(d.data.thickness))
            throw std::invalid_argument("treesap: ABLData::inv_thickness must be 1/thickness");
        desc.abl_thickness = d.data.thickness;
        desc.abl_max = d.data.max_strength;
        desc.abl_profile = profile_id(d.data.profile, d.data.thickness);
        for (int c = 0; c < In::ncomp; ++c) desc.target[c] = d.data.target_values[c];
        decorators<In>::fill(d.rhs, desc);
    }
};

// ---- boundary functors -----------------------------------------------------

template <class BC> struct bc_map {
    static_assert(detail::always_false<BC>,
                  "treesap: the device library carries NoOpBC and DirichletBC as physical boundary functors");
};
template <int nc, int ng> struct bc_map<NoOpBC<nc, ng>> {
    static void fill(const NoOpBC<nc, ng> &, tsp_desc &d, bool) { d.bc = TSP_BC_NOOP; }
};
template <int nc, int ng> struct bc_map<DirichletBC<nc, ng>> {
    static void fill(const DirichletBC<nc, ng> &b, tsp_desc &d, bool has_abl) {
        d.bc = TSP_BC_DIRICHLET;
        for (int c = 0; c < nc; ++c) {
            if (has_abl && !(d.target[c] == b.boundary_values[c]))
                throw std::invalid_argument("treesap: ABL target values and Dirichlet boundary values differ; the "
                                            "device keeps one target per component");
            d.target[c] = b.boundary_values[c];
        }
    }
};

// ---- the descriptor --------------------------------------------------------

template <class Model> tsp_desc make_desc(const Model &m) {
    static_assert(Model::ncomp <= TSP_MAX_COMP, "treesap: at most 12 components");
    using Dec = decorators<typename Model::rhs_t>;
    const amrex::Geometry &geom = m.get_geometry();

    tsp_desc d;
    std::memset(&d, 0, sizeof(d));
    const Recognised rec = recognise(Dec::user(m.get_rhs()), geom.data());
    d.system = rec.system;
    d.stencil_n = rec.stencil_n;
    d.gamma = rec.gamma;
    d.ghost = Model::ngrow;
    for (int a = 0; a < 3; ++a) {
        d.n[a] = geom.Domain().length(a);
        d.prob_lo[a] = geom.ProbLo(a);
        d.prob_hi[a] = geom.ProbHi(a);
        d.periodic[a] = geom.isPeriodic(a) ? 1 : 0;
    }
    d.abl_profile = TSP_PROFILE_QUADRATIC;
    Dec::fill(m.get_rhs(), d);
    bc_map<typename Model::bc_t>::fill(m.get_bc(), d, Dec::has_abl);
    d.dt = m.dt_hint();
    return d;
}

} // namespace treesap::device
