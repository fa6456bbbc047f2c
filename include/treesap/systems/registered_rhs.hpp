// registered_rhs.hpp -- the right-hand sides the device library carries
// kernels for, written as ordinary TreESaP RHS functors.
//
// Each functor's arithmetic (operation order included) is the one the
// library's stage kernels implement for the matching tsp_system id, so a
// functor here and the device kernel agree to the last bit. The integrators
// use them to recognise a user's RHS: a user functor that evaluates to the
// same values as one of these on a probe field is run with that system's
// kernel (see device/model_map.hpp).
//
//   PhiPiWaveRHS<n, ngrow>   (phi, Pi):     phi_t = Pi,  Pi_t = contract_a D_aa phi
//   PhiUWaveRHS<n, ngrow>    (phi, u_a):    phi_t = contract_a D_a u_a,  u_a,t = D_a phi
//                            (examples/01-wave-eqn/WaveEqnRHS.hpp:16-50)
//   NonLinearWaveRHS<n, ngrow> (phi, pi):   phi_t = pi,  pi_t = contract_a D_aa phi - gamma pi |grad phi|^2
//                            (examples/02-temporaries/NonLinearRHS.hpp:16-71)
//   LinearADMRHS<n, ngrow>   (h_ij, K_ij), two SymRank2 aliases at components 0 and 6:
//       h_ab,t = -2 K_ab
//       K_ab,t = symmetrise<a,b>(contract_m D_m D_a h_mb) - 0.5 (contract_m D_a D_b h_mm + contract_m D_m D_m h_ab)
// n is the derivative half-width, ngrow >= n the ghost width (room for KO).
#pragma once

#include "../pde-model/PDEModelConcepts.hpp"
#include "../pde-model/tensor-alias/RankNAlias.hpp"
#include "../pde-model/tensor-alias/SymRank2Alias.hpp"
#include "../util/static_for.hpp"
#include "../util/tensor_operations.hpp"
#include "../util/tensor_stencil.hpp"

#include <cmath>

namespace treesap::systems {

template <int n, int ngrow_ = n> struct PhiPiWaveRHS {
    static_assert(n >= 1 && n <= ngrow_);
    static constexpr int ncomp = 2;
    static constexpr int ngrow = ngrow_;
    static constexpr int stencil_n = n;
    struct Workspace {};

    template <int comp>
    AMREX_GPU_HOST_DEVICE amrex::Real operator()(const amrex::IntVect idx, amrex::Real, const amrex::Array4<const amrex::Real> &arr,
                           const amrex::GeometryData &geom, Workspace &) const {
        using namespace treesap::tensor;
        using treesap::stencil::derivative;
        ScalarAlias<0> phi{{arr}};
        ScalarAlias<phi.next> Pi{{arr}};
        if constexpr (phi.contains(comp)) {
            return Pi(idx, {});
        } else {
            return contract([&]<int a>() { return derivative<n, a, a>(phi, {}, idx, geom); });
        }
    }
};

template <int n, int ngrow_ = n> struct PhiUWaveRHS {
    static_assert(n >= 1 && n <= ngrow_);
    static constexpr int ncomp = 4;
    static constexpr int ngrow = ngrow_;
    static constexpr int stencil_n = n;
    struct Workspace {};

    template <int comp>
    AMREX_GPU_HOST_DEVICE amrex::Real operator()(const amrex::IntVect idx, amrex::Real, const amrex::Array4<const amrex::Real> &arr,
                           const amrex::GeometryData &geom, Workspace &) const {
        using namespace treesap::tensor;
        using treesap::stencil::derivative;
        ScalarAlias<0> phi{{arr}};
        VectorAlias<phi.next> u{{arr}};
        if constexpr (phi.contains(comp)) {
            return contract([&]<int a>() { return derivative<n, a>(u, {{a}}, idx, geom); });
        } else {
            constexpr int a = u.restore_absolute(comp)[0];
            return derivative<n, a>(phi, {}, idx, geom);
        }
    }
};

template <int n, int ngrow_ = n> struct NonLinearWaveRHS {
    static_assert(n >= 1 && n <= ngrow_);
    static constexpr int ncomp = 2;
    static constexpr int ngrow = ngrow_;
    static constexpr int stencil_n = n;
    using PhiAlias = treesap::tensor::ScalarAlias<0>;
    using PiAlias = treesap::tensor::ScalarAlias<PhiAlias::next>;
    struct Workspace {
        PiAlias::Temporary pi;
    };

    amrex::Real gamma = 0.0;

    template <int comp>
    AMREX_GPU_HOST_DEVICE amrex::Real operator()(const amrex::IntVect idx, amrex::Real, const amrex::Array4<const amrex::Real> &arr,
                           const amrex::GeometryData &geom, Workspace &ws) const {
        using namespace treesap::tensor;
        using treesap::stencil::derivative;
        PhiAlias phi{{arr}};
        PiAlias pi{{arr}};
        if constexpr (comp == 0) ws.pi = pi.load(idx);
        if constexpr (phi.contains(comp)) {
            return ws.pi({});
        } else {
            VectorAlias<>::Temporary grad{};
            treesap::static_for<3>([&]<int a>() { grad({{a}}) = derivative<n, a>(phi, {}, idx, geom); });
            const amrex::Real grad_sq = contract([&]<int a>() { return grad({{a}}) * grad({{a}}); });
            const amrex::Real lap = contract([&]<int a>() { return derivative<n, a, a>(phi, {}, idx, geom); });
            return lap - gamma * ws.pi({}) * grad_sq;
        }
    }
};

template <int n, int ngrow_ = n> struct LinearADMRHS {
    static_assert(n >= 1 && n <= ngrow_);
    static constexpr int ncomp = 12;
    static constexpr int ngrow = ngrow_;
    static constexpr int stencil_n = n;
    struct Workspace {};

    template <int comp>
    AMREX_GPU_HOST_DEVICE amrex::Real operator()(const amrex::IntVect idx, amrex::Real, const amrex::Array4<const amrex::Real> &arr,
                           const amrex::GeometryData &geom, Workspace &) const {
        using namespace treesap::tensor;
        using treesap::stencil::derivative;
        SymRank2Alias<0> h{{arr}};
        SymRank2Alias<h.next> K{{arr}};
        if constexpr (h.contains(comp)) {
            constexpr auto ab = h.restore_absolute(comp);
            return -2.0 * K(idx, ab);
        } else {
            constexpr auto ab = K.restore_absolute(comp);
            constexpr int a = ab[0], b = ab[1];
            // f<p,q> = contract_m D_m D_p h_{mq}
            auto f = [&]<int p, int q>() {
                return contract([&]<int m>() { return derivative<n, m, p>(h, {{m, q}}, idx, geom); });
            };
            const amrex::Real sym = symmetrise<a, b>(f);
            const amrex::Real tdd = contract([&]<int m>() { return derivative<n, a, b>(h, {{m, m}}, idx, geom); });
            const amrex::Real lap = contract([&]<int m>() { return derivative<n, m, m>(h, {{a, b}}, idx, geom); });
            return sym - 0.5 * (tdd + lap);
        }
    }
};

static_assert(RHSConcept<PhiPiWaveRHS<2>> && RHSConcept<PhiUWaveRHS<1>> && RHSConcept<NonLinearWaveRHS<1>> &&
              RHSConcept<LinearADMRHS<2>>);

} // namespace treesap::systems
