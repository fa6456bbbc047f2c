// PDEModelConcepts.hpp -- compile-time contracts of the plug-in API
// (pde-model/PDEModelConcepts.hpp:44-139).
//
// RHSConcept: trivially copyable; static ncomp > 0 and ngrow >= 0; a
//   default-constructible Workspace; and for every component c in [0, ncomp)
//   f.template operator()<c>(IntVect, Real t, const Array4<const Real>&,
//   const GeometryData&, Workspace&) returning something convertible to Real.
// BCConcept: trivially copyable; ncomp, ngrow; a void call operator
//   (IntVect, Array4<Real>, dcomp, numcomp, GeometryData, time, const BCRec*, bcomp, orig_comp).
// InitConcept: trivially copyable; ncomp; operator()<0>(x, y, z) -> Real.
// PDEModelConcept: ncomp, ngrow, rhs_t/init_t/bc_t satisfying the above and
//   the accessors get_rhs/get_init/get_bc/get_bc_recs/get_geometry/get_comp_names/dt_hint.
#pragma once

#include "../amrex_types.hpp"

#include <concepts>
#include <string>
#include <type_traits>
#include <utility>

namespace treesap {

namespace detail {
template <typename F, std::size_t... Is> constexpr bool every_component_callable(std::index_sequence<Is...>) {
    return (requires(const F &f, const amrex::IntVect idx, amrex::Real t, const amrex::Array4<const amrex::Real> &arr,
                     const amrex::GeometryData &geom, typename F::Workspace &ws) {
        { f.template operator()<static_cast<int>(Is)>(idx, t, arr, geom, ws) } -> std::convertible_to<amrex::Real>;
    } && ...);
}
} // namespace detail

template <typename F>
concept RHSConcept = std::is_trivially_copyable_v<F> && requires {
    typename std::integral_constant<int, F::ncomp>;
    typename std::integral_constant<int, F::ngrow>;
    requires F::ncomp > 0;
    requires F::ngrow >= 0;
    typename F::Workspace;
} && std::is_default_constructible_v<typename F::Workspace> &&
                     detail::every_component_callable<F>(std::make_index_sequence<F::ncomp>{});

template <typename F>
concept BCConcept = std::is_trivially_copyable_v<F> && requires {
    typename std::integral_constant<int, F::ncomp>;
    typename std::integral_constant<int, F::ngrow>;
    requires F::ncomp > 0;
    requires F::ngrow >= 0;
} && requires(const F &f, const amrex::IntVect &iv, const amrex::Array4<amrex::Real> &dest, int dcomp, int numcomp,
              const amrex::GeometryData &geom, amrex::Real time, const amrex::BCRec *bcr, int bcomp, int orig_comp) {
    { f(iv, dest, dcomp, numcomp, geom, time, bcr, bcomp, orig_comp) } -> std::same_as<void>;
};

template <typename F>
concept InitConcept = std::is_trivially_copyable_v<F> && requires {
    typename std::integral_constant<int, F::ncomp>;
    requires F::ncomp > 0;
} && requires(const F &f, amrex::Real x, amrex::Real y, amrex::Real z) {
    { f.template operator()<0>(x, y, z) } -> std::convertible_to<amrex::Real>;
};

template <typename M>
concept PDEModelConcept = requires {
    typename std::integral_constant<int, M::ncomp>;
    typename std::integral_constant<int, M::ngrow>;
    requires M::ncomp > 0;
    requires M::ngrow >= 0;
    typename M::rhs_t;
    typename M::init_t;
    typename M::bc_t;
} && requires(M m) {
    requires RHSConcept<typename M::rhs_t>;
    { m.get_rhs() } -> std::same_as<const typename M::rhs_t &>;
    requires InitConcept<typename M::init_t>;
    { m.get_init() } -> std::same_as<const typename M::init_t &>;
    requires BCConcept<typename M::bc_t>;
    { m.get_bc() } -> std::same_as<const typename M::bc_t &>;
    { m.get_bc_recs() } -> std::convertible_to<const amrex::Vector<amrex::BCRec> &>;
    { m.get_geometry() } -> std::same_as<const amrex::Geometry &>;
    { m.get_comp_names() } -> std::convertible_to<amrex::Vector<std::string>>;
    { m.dt_hint() } -> std::convertible_to<amrex::Real>;
};

} // namespace treesap
