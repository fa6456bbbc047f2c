// TensorAliasBase.hpp -- component aliasing of tensor fields stored as
// consecutive components of one Array4 (pde-model/tensor-alias/TensorAliasBase.hpp:136-287).
//
// An alias of rank R in D dimensions occupies ncomp components starting at
// comp_base; `next` is the first component after it, so aliases can be
// chained (VectorAlias<phi.next>). The derived class supplies the map between
// index tuples and relative components (get_relative / restore_relative); the
// concept below checks at compile time that the map is a bijection onto
// [0, ncomp) for every index tuple. An alias is an aggregate holding a
// reference to the array, initialised as `Alias a{{arr}}`.
#pragma once

#include "../../amrex_types.hpp"
#include "IndexArray.hpp"

#include <concepts>
#include <type_traits>

namespace treesap::tensor {

namespace detail {

// restore_relative must invert get_relative on [0, ncomp), and every one of
// the dim^rank index tuples must map into [0, ncomp).
template <typename A> constexpr bool alias_is_consistent() {
    for (int c = 0; c < A::ncomp; ++c)
        if (A::get_relative(A::restore_relative(c)) != c) return false;
    if constexpr (A::rank == 0) {
        return A::ncomp == 1 && A::get_relative(IndexArray<0>{}) == 0;
    } else {
        int total = 1;
        for (int r = 0; r < A::rank; ++r) total *= A::dim;
        for (int t = 0; t < total; ++t) {
            IndexArray<A::rank> idx{};
            int rest = t;
            for (int r = 0; r < A::rank; ++r) {
                idx[r] = rest % A::dim;
                rest /= A::dim;
            }
            const int c = A::get_relative(idx);
            if (c < 0 || c >= A::ncomp) return false;
        }
        return true;
    }
}

} // namespace detail

template <typename Derived>
concept TensorAliasImplementation = requires {
    typename Derived::Base;
    requires std::derived_from<Derived, typename Derived::Base>;
    typename std::integral_constant<int, Derived::ncomp>;
    requires Derived::ncomp > 0;
    typename std::integral_constant<int, Derived::rank>;
    requires Derived::rank >= 0;
    typename std::integral_constant<int, Derived::dim>;
    requires Derived::dim > 0;
    typename std::integral_constant<int, Derived::next>;
} && requires(int comp, IndexArray<Derived::rank> ti) {
    { Derived::restore_relative(comp) } -> std::same_as<IndexArray<Derived::rank>>;
    { Derived::get_relative(ti) } -> std::same_as<int>;
} && requires {
    typename std::integral_constant<int, Derived::get_relative(IndexArray<Derived::rank>{})>;
    typename std::integral_constant<int, Derived::restore_relative(0)[0]>;
} && detail::alias_is_consistent<Derived>();

template <typename Derived, int comp_base_, int ncomp_, int rank_, int dim_, typename ArrayType>
struct TensorAliasBase {
    static constexpr int comp_base = comp_base_;
    static constexpr int ncomp = ncomp_;
    static constexpr int rank = rank_;
    static constexpr int dim = dim_;
    static constexpr int next = comp_base_ + ncomp_;
    static_assert(ncomp > 0 && rank >= 0 && dim > 0);

    using arr_t = ArrayType;

    const amrex::Array4<ArrayType> &arr;

    // Per-cell register holding all components of the alias, indexed like the alias.
    struct Temporary {
        using value_type = std::remove_cv_t<ArrayType>;
        value_type data[ncomp]{};

        constexpr value_type &operator()(const IndexArray<rank> &ti) { return data[Derived::get_relative(ti)]; }
        constexpr const value_type &operator()(const IndexArray<rank> &ti) const {
            return data[Derived::get_relative(ti)];
        }
    };

    static constexpr bool contains(int absolute_comp) {
        return absolute_comp >= comp_base && absolute_comp < comp_base + ncomp;
    }
    static constexpr int get_absolute(const IndexArray<rank> &ti) { return comp_base + Derived::get_relative(ti); }
    static constexpr IndexArray<rank> restore_absolute(int absolute_comp) {
        return Derived::restore_relative(absolute_comp - comp_base);
    }

    AMREX_GPU_HOST_DEVICE ArrayType &operator()(int i, int j, int k, const IndexArray<rank> &ti) const {
        return arr(i, j, k, get_absolute(ti));
    }
    AMREX_GPU_HOST_DEVICE ArrayType &operator()(const amrex::IntVect &iv, const IndexArray<rank> &ti) const {
        return arr(iv, get_absolute(ti));
    }

    AMREX_GPU_HOST_DEVICE Temporary load(int i, int j, int k) const {
        Temporary t{};
        for (int c = 0; c < ncomp; ++c) t.data[c] = arr(i, j, k, comp_base + c);
        return t;
    }
    AMREX_GPU_HOST_DEVICE Temporary load(const amrex::IntVect &iv) const { return load(iv[0], iv[1], iv[2]); }
};

} // namespace treesap::tensor
