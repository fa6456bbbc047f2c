// SymRank2Alias.hpp -- symmetric rank-2 tensor alias
// (pde-model/tensor-alias/SymRank2Alias.hpp:25-53): D(D+1)/2 components, (a, b)
// and (b, a) both at row(row+1)/2 + col with row = max(a, b), col = min(a, b),
// i.e. xx, yx, yy, zx, zy, zz in 3-D. This is the storage order of the
// 12-component (h_ij, K_ij) state.
#pragma once

#include "TensorAliasBase.hpp"

namespace treesap::tensor {

template <int comp_base_ = 0, int dim_ = 3, typename ArrayType = const amrex::Real>
struct SymRank2Alias : TensorAliasBase<SymRank2Alias<comp_base_, dim_, ArrayType>, comp_base_, (dim_ * (dim_ + 1)) / 2,
                                       2, dim_, ArrayType> {
    using Base =
        TensorAliasBase<SymRank2Alias<comp_base_, dim_, ArrayType>, comp_base_, (dim_ * (dim_ + 1)) / 2, 2, dim_, ArrayType>;
    using Base::dim;
    using Base::ncomp;
    using Base::next;
    using Base::rank;

    static constexpr int get_relative(const IndexArray<2> ti) {
        const int hi = ti[0] > ti[1] ? ti[0] : ti[1];
        const int lo = ti[0] > ti[1] ? ti[1] : ti[0];
        return hi * (hi + 1) / 2 + lo;
    }
    static constexpr IndexArray<2> restore_relative(int rel) {
        int row = 0;
        while ((row + 1) * (row + 2) / 2 <= rel) ++row;
        return {{row, rel - row * (row + 1) / 2}};
    }
};

static_assert(TensorAliasImplementation<SymRank2Alias<0>>);

} // namespace treesap::tensor
