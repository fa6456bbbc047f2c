// user_rhs.hpp -- right-hand sides written the way a TreESaP user writes
// them (aliases, contract, derivative<>), with no reference to the library's
// registry. Mapped with system_traits (user_rhs_traits.hpp) they run on the
// library's kernels; unmapped, on the generic kernel compiled from them.
#pragma once

#include "pde-model/PDEModelConcepts.hpp"
#include "pde-model/tensor-alias/RankNAlias.hpp"
#include "pde-model/tensor-alias/SymRank2Alias.hpp"
#include "util/named_tensors.hpp"
#include "util/static_for.hpp"
#include "util/tensor_operations.hpp"
#include "util/tensor_stencil.hpp"

#include <cmath>

// d_t phi = div u, d_t u_a = d_a phi, stencil half-width = ghost width
template <int ngrow_> struct FirstOrderWave {
    static constexpr int ncomp = 4;
    static constexpr int ngrow = ngrow_;
    struct Workspace {};

    template <int comp>
    AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real operator()(const amrex::IntVect idx, amrex::Real,
                                                                    const amrex::Array4<const amrex::Real> &arr,
                                                                    const amrex::GeometryData &geom, Workspace &) const {
        using namespace treesap;
        using namespace tensor;
        using namespace stencil;
        ScalarAlias<0> phi{{arr}};
        VectorAlias<phi.next> u{{arr}};
        static_assert(u.next == ncomp);
        if constexpr (phi.contains(comp)) {
            return contract([&]<int a>() { return derivative<ngrow, a>(u, {a}, idx, geom); });
        } else if constexpr (u.contains(comp)) {
            constexpr int a = u.restore_absolute(comp)[0];
            return derivative<ngrow, a>(phi, {}, idx, geom);
        } else {
            return NAN;
        }
    }
};

// d_t phi = Pi, d_t Pi = Laplacian phi with half-width `order`, ghost width ngrow (room for dissipation)
template <int order, int ngrow_> struct ScalarWave {
    static constexpr int ncomp = 2;
    static constexpr int ngrow = ngrow_;
    struct Workspace {};

    template <int comp>
    AMREX_GPU_HOST_DEVICE amrex::Real operator()(const amrex::IntVect idx, amrex::Real,
                                                 const amrex::Array4<const amrex::Real> &arr,
                                                 const amrex::GeometryData &geom, Workspace &) const {
        using namespace treesap::tensor;
        using treesap::stencil::derivative;
        if constexpr (comp == 0) {
            return arr(idx, 1);
        } else {
            ScalarAlias<0> phi{{arr}};
            return contract([&]<int a>() { return derivative<order, a, a>(phi, {}, idx, geom); });
        }
    }
};

// nonlinear wave with a temporary kept in the workspace between components
template <int ngrow_> struct SelfDampedWave {
    static constexpr int ncomp = 2;
    static constexpr int ngrow = ngrow_;
    using PhiT = treesap::tensor::ScalarAlias<0>;
    using PiT = treesap::tensor::ScalarAlias<PhiT::next>;
    amrex::Real gamma;
    struct Workspace {
        PiT::Temporary pi;
    };

    template <int comp>
    AMREX_GPU_HOST_DEVICE amrex::Real operator()(const amrex::IntVect idx, amrex::Real,
                                                 const amrex::Array4<const amrex::Real> &arr,
                                                 const amrex::GeometryData &geom, Workspace &ws) const {
        using namespace treesap;
        using namespace tensor;
        using namespace stencil;
        PhiT phi{{arr}};
        PiT pi{{arr}};
        if constexpr (comp == 0) {
            ws.pi = pi.load(idx);
            return ws.pi({});
        } else {
            VectorAlias<>::Temporary g{};
            static_for<3>([&]<int a>() { g({a}) = derivative<ngrow, a>(phi, {}, idx, geom); });
            const amrex::Real g2 = contract([&]<int a>() { return g({a}) * g({a}); });
            const amrex::Real lap = contract([&]<int a>() { return derivative<ngrow, a, a>(phi, {}, idx, geom); });
            return lap - gamma * ws.pi({}) * g2;
        }
    }
};

// linearised ADM in vacuum with unit lapse and zero shift
template <int order, int ngrow_> struct VacuumADM {
    static constexpr int ncomp = 12;
    static constexpr int ngrow = ngrow_;
    struct Workspace {};

    template <int comp>
    AMREX_GPU_HOST_DEVICE amrex::Real operator()(const amrex::IntVect idx, amrex::Real,
                                                 const amrex::Array4<const amrex::Real> &arr,
                                                 const amrex::GeometryData &geom, Workspace &) const {
        using namespace treesap::tensor;
        using treesap::stencil::derivative;
        SymRank2Alias<0> h{{arr}};
        SymRank2Alias<h.next> K{{arr}};
        if constexpr (h.contains(comp)) {
            return -2.0 * K(idx, h.restore_absolute(comp));
        } else {
            constexpr auto ij = K.restore_absolute(comp);
            constexpr int i = ij[0], j = ij[1];
            auto dd_h = [&]<int p, int q>() {
                return contract([&]<int k>() { return derivative<order, k, p>(h, {k, q}, idx, geom); });
            };
            const amrex::Real s = symmetrise<i, j>(dd_h);
            const amrex::Real trace_part = contract([&]<int k>() { return derivative<order, i, j>(h, {k, k}, idx, geom); });
            const amrex::Real lap_part = contract([&]<int k>() { return derivative<order, k, k>(h, {i, j}, idx, geom); });
            return s - 0.5 * (trace_part + lap_part);
        }
    }
};

// (phi, Pi) wave plus a source term switched on in the upper half of the box in x: equal to ScalarWave at
// every cell with i < nx/2, so only a check spread over the domain tells them apart
template <int ngrow_> struct PatchedWave {
    static constexpr int ncomp = 2;
    static constexpr int ngrow = ngrow_;
    struct Workspace {};
    template <int comp>
    AMREX_GPU_HOST_DEVICE amrex::Real operator()(const amrex::IntVect idx, amrex::Real t,
                                                 const amrex::Array4<const amrex::Real> &arr,
                                                 const amrex::GeometryData &geom, Workspace &ws) const {
        typename ScalarWave<ngrow_, ngrow_>::Workspace w{};
        const amrex::Real base = ScalarWave<ngrow_, ngrow_>{}.template operator()<comp>(idx, t, arr, geom, w);
        const int half = geom.Domain().length(0) / 2;
        return (comp == 1 && idx[0] >= half) ? base + 0.25 * arr(idx, 0) : base;
    }
};

// A system the library carries no kernel for: third derivatives, an
// antisymmetric part, a Levi-Civita contraction (a curl) and a cyclic sum,
// position- and time-dependent coefficients. Components: phi (0), E_a (1-3).
//   d_t phi = -div E + 0.01 (D_xxx + D_yyy + D_zzz) phi + 0.001 D_xyz phi + 0.05 x (1 - t/2) phi
//   d_t E_a = -D_a phi + 0.2 (curl E)_a + 0.1 antisym_{a, a+1}(D_b E_c)
//             + 0.01 cyclic_sum_{a b c}(D_b E_c) + y * 0.01 E_a
template <int ngrow_> struct CurlWave {
    static constexpr int ncomp = 4;
    static constexpr int ngrow = ngrow_;
    struct Workspace {};
    template <int comp>
    AMREX_GPU_HOST_DEVICE amrex::Real operator()(const amrex::IntVect idx, amrex::Real t,
                                                 const amrex::Array4<const amrex::Real> &arr,
                                                 const amrex::GeometryData &geom, Workspace &) const {
        using namespace treesap;
        using namespace tensor;
        using namespace stencil;
        ScalarAlias<0> phi{{arr}};
        VectorAlias<phi.next> E{{arr}};
        const amrex::Real x = geom.ProbLo(0) + (idx[0] + 0.5) * geom.CellSize(0);
        const amrex::Real y = geom.ProbLo(1) + (idx[1] + 0.5) * geom.CellSize(1);
        if constexpr (phi.contains(comp)) {
            const amrex::Real div = contract([&]<int a>() { return derivative<ngrow, a>(E, {a}, idx, geom); });
            const amrex::Real d3 = contract([&]<int a>() { return derivative<ngrow, a, a, a>(phi, {}, idx, geom); });
            const amrex::Real dxyz = derivative<ngrow, 0, 1, 2>(phi, {}, idx, geom);
            return -div + 0.01 * d3 + 0.001 * dxyz + 0.05 * x * (1.0 - 0.5 * t) * phi(idx, {});
        } else {
            constexpr int a = E.restore_absolute(comp)[0];
            constexpr int a1 = (a + 1) % 3;
            auto dE = [&]<int b, int c>() { return derivative<ngrow, b>(E, {c}, idx, geom); };
            const amrex::Real curl = contract<2>([&]<int b, int c>() {
                return static_cast<amrex::Real>(levi_civita<a, b, c>) * derivative<ngrow, b>(E, {c}, idx, geom);
            });
            const amrex::Real anti = antisymmetrise<a, a1>(dE);
            const amrex::Real cyc = cyclic_permutation_sum<a, a1, (a + 2) % 3>([&]<int p, int q, int r>() {
                return derivative<ngrow, p>(E, {q}, idx, geom) * static_cast<amrex::Real>(r + 1);
            });
            return -derivative<ngrow, a>(phi, {}, idx, geom) + 0.2 * curl + 0.1 * anti + 0.01 * cyc +
                   y * 0.01 * E(idx, {a});
        }
    }
};

// not a registered system: d_t u = d_x u (advection)
template <int ngrow_> struct Advection {
    static constexpr int ncomp = 2;
    static constexpr int ngrow = ngrow_;
    struct Workspace {};
    template <int comp>
    AMREX_GPU_HOST_DEVICE amrex::Real operator()(const amrex::IntVect idx, amrex::Real,
                                                 const amrex::Array4<const amrex::Real> &arr,
                                                 const amrex::GeometryData &geom, Workspace &) const {
        return treesap::stencil::derivative<ngrow, 0>(arr, comp, idx, geom);
    }
};

// (phi, Pi) wave with an explicit time dependence
template <int ngrow_> struct ForcedWave {
    static constexpr int ncomp = 2;
    static constexpr int ngrow = ngrow_;
    struct Workspace {};
    template <int comp>
    amrex::Real operator()(const amrex::IntVect idx, amrex::Real t, const amrex::Array4<const amrex::Real> &arr,
                           const amrex::GeometryData &geom, Workspace &ws) const {
        const amrex::Real base = ScalarWave<ngrow_, ngrow_>{}.template operator()<comp>(idx, t, arr, geom, ws_cast(ws));
        return comp == 1 ? base + std::sin(t) : base;
    }
    static typename ScalarWave<ngrow_, ngrow_>::Workspace &ws_cast(Workspace &) {
        static typename ScalarWave<ngrow_, ngrow_>::Workspace w;
        return w;
    }
};

static_assert(treesap::RHSConcept<FirstOrderWave<1>> && treesap::RHSConcept<ScalarWave<2, 3>> &&
              treesap::RHSConcept<SelfDampedWave<2>> && treesap::RHSConcept<VacuumADM<2, 2>> &&
              treesap::RHSConcept<Advection<1>> && treesap::RHSConcept<ForcedWave<2>> &&
              treesap::RHSConcept<PatchedWave<2>> && treesap::RHSConcept<CurlWave<3>>);
