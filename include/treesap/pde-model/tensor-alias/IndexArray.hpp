// IndexArray.hpp -- fixed-size tensor index tuple (pde-model/tensor-alias/IndexArray.hpp:21-29).
// A rank-0 array still holds one (unused) slot so that `{}` is valid for scalars.
#pragma once

#include "../../amrex_types.hpp"

namespace treesap::tensor {

template <int rank_> struct IndexArray {
    static_assert(rank_ >= 0, "tensor rank must be non-negative");
    int data[rank_ > 0 ? rank_ : 1]{};

    constexpr int &operator[](int i) { return data[i]; }
    constexpr const int &operator[](int i) const { return data[i]; }
};

} // namespace treesap::tensor
