// model_map.hpp -- turns a TreESaP PDE model into the descriptor of the C ABI
// (tsp_desc), i.e. decides which device kernel runs the model.
//
// A model's right-hand side is a compile-time functor (RHSConcept,
// pde-model/PDEModelConcepts.hpp:44-64). Two device paths:
//
//   library kernels  for the functors of systems/registered_rhs.hpp and for any
//                    functor a system_traits specialisation maps onto one of
//                    them, optionally wrapped as ABLDecorator<KODecorator<user>>
//                    (the nesting ABLPDEModel produces when handed a
//                    KODecorator): the tiled stage kernels of the library.
//                    Decorators are unwrapped at compile time (KODecorator<R, r>
//                    gives ko_order r and its strength, ABLDecorator<R, P> its
//                    thickness, maximum, targets and profile; other nestings do
//                    not compile). The profile must equal s, s*s or s*s*s on every
//                    argument the layer uses. A user's system_traits mapping is
//                    verified when the model is mapped: the functor and the
//                    registered one are evaluated on seeded random fields at
//                    cells spread over the whole domain (corners, faces, the
//                    absorbing layer, the bulk, random cells) at several times,
//                    every component in order with one Workspace per cell, and
//                    must agree bit for bit; otherwise the mapping throws.
//   generic kernel   every other functor (TSP_SYS_CUSTOM): the decorated functor
//                    itself is compiled into the stage kernel of
//                    device/generic.hpp; the library fills ghosts and keeps the
//                    registers and the clock.
//
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
// A user's mapping is verified when the model is mapped (see above); types
// without one run on the generic kernel (device/generic.hpp).
template <class RHS> struct system_traits {};

namespace device {
// the library's own functors: exact by construction, not probed
template <class RHS> constexpr bool is_builtin = false;
template <int n, int g> constexpr bool is_builtin<systems::PhiPiWaveRHS<n, g>> = true;
template <int n, int g> constexpr bool is_builtin<systems::PhiUWaveRHS<n, g>> = true;
template <int n, int g> constexpr bool is_builtin<systems::NonLinearWaveRHS<n, g>> = true;
template <int n, int g> constexpr bool is_builtin<systems::LinearADMRHS<n, g>> = true;

// whether the library's kernels run a functor (directly or through a system_traits mapping)
template <class RHS> constexpr bool is_mapped = requires { system_traits<RHS>::system; };
} // namespace device

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

// Cells of the domain at which a mapping is verified: the 8 corners, the 6
// face centres, the centre, a cell `inset` cells in from every corner (inside
// an absorbing layer of that depth) and 16 seeded random cells.
inline std::vector<amrex::IntVect> probe_cells(const amrex::Box &dom, int inset) {
    std::vector<amrex::IntVect> out;
    const amrex::IntVect lo = dom.smallEnd(), hi = dom.bigEnd();
    for (int c = 0; c < 8; ++c)
        out.push_back({(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]});
    const amrex::IntVect mid{(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2};
    out.push_back(mid);
    for (int a = 0; a < 3; ++a) {
        amrex::IntVect f = mid;
        f[a] = lo[a];
        out.push_back(f);
        f[a] = hi[a];
        out.push_back(f);
    }
    for (int c = 0; c < 8; ++c) {
        amrex::IntVect v;
        for (int a = 0; a < 3; ++a) {
            const int d = std::min(inset, (hi[a] - lo[a]) / 2);
            v[a] = ((c >> a) & 1) ? hi[a] - d : lo[a] + d;
        }
        out.push_back(v);
    }
    SplitMix64 rng{0x70726f6265636c6cull};
    for (int q = 0; q < 16; ++q) {
        amrex::IntVect v;
        for (int a = 0; a < 3; ++a) v[a] = lo[a] + (int)(rng.next() % (std::uint64_t)(hi[a] - lo[a] + 1));
        out.push_back(v);
    }
    return out;
}

// Whether functors `a` and `b` give bit-identical values (every component in
// order, one Workspace per cell) on seeded random fields around every probe
// cell, at several times. Returns a description of the first difference.
template <class A, class B>
bool probe_equal(const A &a, const B &b, const amrex::GeometryData &g, int inset, std::string *why) {
    static_assert(A::ncomp == B::ncomp);
    constexpr int nc = A::ncomp;
    constexpr int H = A::ngrow > B::ngrow ? A::ngrow : B::ngrow;
    constexpr int W = 2 * H + 1;
    std::vector<double> field((std::size_t)W * W * W * nc);
    SplitMix64 rng{0x7265636f676e6973ull};
    for (const amrex::IntVect &cell : probe_cells(g.Domain(), inset)) {
        for (double &v : field) v = rng.uniform();
        const amrex::Array4<const double> arr(field.data(), {cell[0] - H, cell[1] - H, cell[2] - H},
                                              {cell[0] + H + 1, cell[1] + H + 1, cell[2] + H + 1}, nc);
        for (double t : {0.0, 0.375, 1.90625}) {
            typename A::Workspace wa{};
            typename B::Workspace wb{};
            bool same = true;
            int bad = -1;
            double va = 0.0, vb = 0.0;
            treesap::static_for<nc>([&]<int c>() {
                const double x = a.template operator()<c>(cell, t, arr, g, wa);
                const double y = b.template operator()<c>(cell, t, arr, g, wb);
                if (same && std::memcmp(&x, &y, sizeof(double)) != 0) {
                    same = false;
                    bad = c;
                    va = x;
                    vb = y;
                }
            });
            if (!same) {
                if (why)
                    *why = "component " + std::to_string(bad) + " at cell (" + std::to_string(cell[0]) + "," +
                           std::to_string(cell[1]) + "," + std::to_string(cell[2]) + "), t=" + std::to_string(t) +
                           ": " + std::to_string(va) + " vs " + std::to_string(vb);
                return false;
            }
        }
    }
    return true;
}

template <int sys, int n, int g> struct registered_functor;
template <int n, int g> struct registered_functor<TSP_SYS_WAVE_PHIPI, n, g> {
    template <class U> static systems::PhiPiWaveRHS<n, g> make(double) { return {}; }
};
template <int n, int g> struct registered_functor<TSP_SYS_WAVE_PHIU, n, g> {
    template <class U> static systems::PhiUWaveRHS<n, g> make(double) { return {}; }
};
template <int n, int g> struct registered_functor<TSP_SYS_NONLINEAR, n, g> {
    template <class U> static systems::NonLinearWaveRHS<n, g> make(double gamma) { return {gamma}; }
};
template <int n, int g> struct registered_functor<TSP_SYS_TENSOR_ADM, n, g> {
    template <class U> static systems::LinearADMRHS<n, g> make(double) { return {}; }
};

} // namespace detail

// Which library system evaluates the undecorated right-hand side `u` (which
// must be mapped: registered, or named by a system_traits specialisation); a
// user's mapping is verified by probing, and a mismatch throws.
template <class U> Recognised recognise(const U &u, const amrex::GeometryData &g, int inset) {
    static_assert(is_mapped<U>, "recognise() is for mapped functors");
    using T = system_traits<U>;
    const Recognised r{T::system, T::stencil_n, T::gamma(u)};
    if constexpr (!is_builtin<U>) {
        static_assert(T::stencil_n >= 1 && T::stencil_n <= 4, "system_traits: stencil_n must be in 1..4");
        const auto ref = detail::registered_functor<T::system, T::stencil_n, U::ngrow>::template make<U>(r.gamma);
        static_assert(decltype(ref)::ncomp == U::ncomp, "system_traits: the mapped system has another ncomp");
        std::string why;
        if (!detail::probe_equal(u, ref, g, inset, &why))
            throw std::invalid_argument(std::string("treesap: system_traits maps the right-hand side ") + typeid(U).name() +
                                        " onto library system " + std::to_string(T::system) +
                                        " (stencil_n " + std::to_string(T::stencil_n) + "), but their values differ: " +
                                        why);
    }
    return r;
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

// Any functor but the two the library carries is applied on the device by a
// kernel compiled from it (device/generic_bc.hpp, through tsp_set_bc_callback).
template <class BC> struct bc_map {
    static constexpr bool custom = true;
    static void fill(const BC &, tsp_desc &d, bool) { d.bc = TSP_BC_CUSTOM; }
};
template <int nc, int ng> struct bc_map<NoOpBC<nc, ng>> {
    static constexpr bool custom = false;
    static void fill(const NoOpBC<nc, ng> &, tsp_desc &d, bool) { d.bc = TSP_BC_NOOP; }
};
template <int nc, int ng> struct bc_map<DirichletBC<nc, ng>> {
    static constexpr bool custom = false;
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

// Whether a model runs on the library's kernels (else on the generic kernel).
template <class Model> constexpr bool runs_library_kernels = is_mapped<typename decorators<typename Model::rhs_t>::user_t>;
// Whether a model's boundary functor runs on a kernel compiled from it.
template <class Model> constexpr bool custom_boundary = bc_map<typename Model::bc_t>::custom;

template <class Model> tsp_desc make_desc(const Model &m) {
    static_assert(Model::ncomp <= TSP_MAX_COMP, "treesap: at most 12 components");
    const amrex::Geometry &geom = m.get_geometry();
    tsp_desc d;
    std::memset(&d, 0, sizeof(d));
    d.ghost = Model::ngrow;
    for (int a = 0; a < 3; ++a) {
        d.n[a] = geom.Domain().length(a);
        d.prob_lo[a] = geom.ProbLo(a);
        d.prob_hi[a] = geom.ProbHi(a);
        d.periodic[a] = geom.isPeriodic(a) ? 1 : 0;
    }
    d.abl_profile = TSP_PROFILE_QUADRATIC;
    if constexpr (runs_library_kernels<Model>) {
        using Dec = decorators<typename Model::rhs_t>;
        Dec::fill(m.get_rhs(), d);
        const Recognised rec = recognise(Dec::user(m.get_rhs()), geom.data(), d.abl_thickness > 0 ? d.abl_thickness / 2 : 2);
        d.system = rec.system;
        d.stencil_n = rec.stencil_n;
        d.gamma = rec.gamma;
        bc_map<typename Model::bc_t>::fill(m.get_bc(), d, Dec::has_abl);
    } else {
        // the generic kernel evaluates the decorated functor itself
        d.system = TSP_SYS_CUSTOM;
        d.ncomp = Model::ncomp;
        bc_map<typename Model::bc_t>::fill(m.get_bc(), d, false);
    }
    d.dt = m.dt_hint();
    return d;
}

} // namespace treesap::device
