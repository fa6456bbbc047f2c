// user_rhs.hpp -- right-hand sides written the way a TreESaP user writes
// them (aliases, contract, derivative<>), with no reference to the library's
// registry. The drop-in integrators must recognise each of them and run the
// matching device kernel.
#pragma once

#include "pde-model/PDEModelConcepts.hpp"
#include "pde-model/tensor-alias/RankNAlias.hpp"
#include "pde-model/tensor-alias/SymRank2Alias.hpp"
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
    amrex::Real operator()(const amrex::IntVect idx, amrex::Real, const amrex::Array4<const amrex::Real> &arr,
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
    amrex::Real operator()(const amrex::IntVect idx, amrex::Real, const amrex::Array4<const amrex::Real> &arr,
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
    amrex::Real operator()(const amrex::IntVect idx, amrex::Real, const amrex::Array4<const amrex::Real> &arr,
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

// not a registered system: d_t u = d_x u (advection) -- must be rejected
template <int ngrow_> struct Advection {
    static constexpr int ncomp = 2;
    static constexpr int ngrow = ngrow_;
    struct Workspace {};
    template <int comp>
    amrex::Real operator()(const amrex::IntVect idx, amrex::Real, const amrex::Array4<const amrex::Real> &arr,
                           const amrex::GeometryData &geom, Workspace &) const {
        return treesap::stencil::derivative<ngrow, 0>(arr, comp, idx, geom);
    }
};

// (phi, Pi) wave with an explicit time dependence -- must be rejected
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
              treesap::RHSConcept<Advection<1>> && treesap::RHSConcept<ForcedWave<2>>);
