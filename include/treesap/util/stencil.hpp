// stencil.hpp -- treesap::stencil::derivative on an Array4 component and the
// Kreiss-Oliger dissipation operator.
//
// derivative<n, Dirs...>(arr, comp, idx, geom): the centred derivative of
// half-width n along the listed directions (1 to 3 of them), as in
// util/stencil.hpp:42-52; the arithmetic is in stencil_impl.hpp.
//
// kreiss_oliger<r, dir>(arr, comp, idx, geom) (util/stencil.hpp:77-104):
//   ((-C(2r,r) u + sum_{s=1..r} -(-1)^s C(2r,r+s) (u[+s] + u[-s])) * 4^{-r}) * (1/dx)
// which is (-1)^{r+1} dx^{2r-1} / 4^r (D+ D-)^r; the integer coefficient of
// each term is formed in int64 before it multiplies the sum.
#pragma once

#include "../amrex_types.hpp"
#include "cexprmath.hpp"
#include "stencil_impl.hpp"

namespace treesap::stencil {

// Reference spelling: the array element type is the second template argument.
template <int n, typename ArrayType, int... Dirs>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real derivative(const amrex::Array4<ArrayType> &arr, int comp, const amrex::IntVect &idx,
                                          const amrex::GeometryData &geom) {
    static_assert(((Dirs >= 0 && Dirs <= 2) && ...), "directions must be 0, 1 or 2");
    static_assert(sizeof...(Dirs) >= 1 && sizeof...(Dirs) <= 3, "first, second and third derivatives only");
    return stencil_impl::derivative_impl<n, ArrayType, Dirs...>::eval(idx[0], idx[1], idx[2], comp, arr, geom);
}

// Shorter spelling with the element type deduced: derivative<n, Dirs...>(arr, comp, idx, geom).
template <int n, int... Dirs, typename ArrayType>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real derivative(const amrex::Array4<ArrayType> &arr, int comp, const amrex::IntVect &idx,
                                          const amrex::GeometryData &geom) {
    static_assert(((Dirs >= 0 && Dirs <= 2) && ...), "directions must be 0, 1 or 2");
    static_assert(sizeof...(Dirs) >= 1 && sizeof...(Dirs) <= 3, "first, second and third derivatives only");
    return stencil_impl::derivative_impl<n, ArrayType, Dirs...>::eval(idx[0], idx[1], idx[2], comp, arr, geom);
}

template <int order, int direction, typename ArrayType>
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real kreiss_oliger(const amrex::Array4<ArrayType> &arr, int comp, const amrex::IntVect &idx,
                                             const amrex::GeometryData &geom) {
    static_assert(order > 0, "Kreiss-Oliger order must be positive");
    using namespace cexprmath;
    constexpr int di = direction == 0, dj = direction == 1, dk = direction == 2;
    constexpr double norm = 1.0 / static_cast<double>(power(2, 2 * order));
    const amrex::Real grid_fac = 1.0 / geom.CellSize(direction);
    const int i = idx[0], j = idx[1], k = idx[2];
    amrex::Real sum = static_cast<double>(-binom(2 * order, order)) * arr(i, j, k, comp);
    for (int s = 1; s <= order; ++s) {
        const std::int64_t cs = -alternating_sign(s) * binom(2 * order, order + s);
        sum += static_cast<double>(cs) * (arr(i + s * di, j + s * dj, k + s * dk, comp) +
                                          arr(i - s * di, j - s * dj, k - s * dk, comp));
    }
    return sum * norm * grid_fac;
}

} // namespace treesap::stencil
