// tensor_stencil.hpp -- treesap::stencil::derivative on a component of a
// tensor alias (util/tensor_stencil.hpp:33-97):
//   derivative<n, Dirs...>(alias, {tensor indices}, idx, geom)
// is the derivative of the alias component the index tuple selects.
#pragma once

#include "../pde-model/tensor-alias/TensorAliasBase.hpp"
#include "stencil_impl.hpp"

namespace treesap::stencil {

template <int n, int d0, typename AliasType>
    requires(tensor::TensorAliasImplementation<AliasType>)
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real derivative(const AliasType &t, const tensor::IndexArray<AliasType::rank> &ti,
                                          const amrex::IntVect &idx, const amrex::GeometryData &geom) {
    using T = typename AliasType::Base::arr_t;
    return stencil_impl::derivative_impl<n, T, d0>::eval(idx[0], idx[1], idx[2], AliasType::Base::get_absolute(ti),
                                                         t.arr, geom);
}

template <int n, int d0, int d1, typename AliasType>
    requires(tensor::TensorAliasImplementation<AliasType>)
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real derivative(const AliasType &t, const tensor::IndexArray<AliasType::rank> &ti,
                                          const amrex::IntVect &idx, const amrex::GeometryData &geom) {
    using T = typename AliasType::Base::arr_t;
    return stencil_impl::derivative_impl<n, T, d0, d1>::eval(idx[0], idx[1], idx[2], AliasType::Base::get_absolute(ti),
                                                             t.arr, geom);
}

template <int n, int d0, int d1, int d2, typename AliasType>
    requires(tensor::TensorAliasImplementation<AliasType>)
AMREX_GPU_HOST_DEVICE AMREX_FORCE_INLINE amrex::Real derivative(const AliasType &t, const tensor::IndexArray<AliasType::rank> &ti,
                                          const amrex::IntVect &idx, const amrex::GeometryData &geom) {
    using T = typename AliasType::Base::arr_t;
    return stencil_impl::derivative_impl<n, T, d0, d1, d2>::eval(idx[0], idx[1], idx[2],
                                                                 AliasType::Base::get_absolute(ti), t.arr, geom);
}

} // namespace treesap::stencil
