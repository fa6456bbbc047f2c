// tensor_operations.hpp -- index contraction and (anti)symmetrisation over
// compile-time tensor indices (util/tensor_operations.hpp:45-217).
//
// contract<rank, dim>(f) sums f.template operator()<i_1..i_rank>() over all
// index tuples as a left fold starting from a value-initialised result,
// outermost index first: result = {}; result += f<0,...>; result += f<1,...>; ...
// For rank > 1 each outer term is itself the contraction over the remaining
// indices. symmetrise<a,b>(f) = 0.5 * (f<a,b> + f<b,a>), and the rank-3 forms
// average over the six permutations (antisymmetrise with signs) or the three
// cyclic ones. These orders are what the results depend on bitwise.
#pragma once

#include "../amrex_types.hpp"

#include <utility>

namespace treesap::tensor {

namespace detail {
template <typename F, int... Is> auto contract_value(F &f, std::integer_sequence<int, Is...>)
    -> decltype(f.template operator()<Is...>());

// the contraction over the remaining indices with the outer one fixed at I
template <int I, typename F> struct fix_outer {
    F &f;
    template <int... Js> constexpr auto operator()() const { return f.template operator()<I, Js...>(); }
};
} // namespace detail

template <int rank, int dim, typename F> AMREX_FORCE_INLINE constexpr auto contract(F &&f) {
    static_assert(rank >= 1, "contraction needs at least one index");
    using R = decltype(detail::contract_value(f, std::make_integer_sequence<int, rank>{}));
    R result{};
    [&]<int... Is>(std::integer_sequence<int, Is...>) {
        if constexpr (rank == 1) {
            ((result += f.template operator()<Is>()), ...);
        } else {
            ((result += contract<rank - 1, dim>(detail::fix_outer<Is, std::remove_reference_t<F>>{f})), ...);
        }
    }(std::make_integer_sequence<int, dim>{});
    return result;
}

template <int rank, typename F> AMREX_FORCE_INLINE constexpr auto contract(F &&f) {
    return contract<rank, 3>(std::forward<F>(f));
}

template <typename F> AMREX_FORCE_INLINE constexpr auto contract(F &&f) { return contract<1, 3>(std::forward<F>(f)); }

template <int a, int b, typename F> AMREX_FORCE_INLINE constexpr auto symmetrise(F &&f) {
    return 0.5 * (f.template operator()<a, b>() + f.template operator()<b, a>());
}

template <int a, int b, typename F> AMREX_FORCE_INLINE constexpr auto antisymmetrise(F &&f) {
    if constexpr (a == b) return 0.0;
    else return 0.5 * (f.template operator()<a, b>() - f.template operator()<b, a>());
}

template <int a, int b, int c, typename F> AMREX_FORCE_INLINE constexpr auto symmetrise(F &&f) {
    return (1.0 / 6.0) * (f.template operator()<a, b, c>() + f.template operator()<b, c, a>() +
                          f.template operator()<c, a, b>() + f.template operator()<a, c, b>() +
                          f.template operator()<b, a, c>() + f.template operator()<c, b, a>());
}

template <int a, int b, int c, typename F> AMREX_FORCE_INLINE constexpr auto antisymmetrise(F &&f) {
    if constexpr (a == b || b == c || a == c) return 0.0;
    else
        return (1.0 / 6.0) * (f.template operator()<a, b, c>() + f.template operator()<b, c, a>() +
                              f.template operator()<c, a, b>() - f.template operator()<a, c, b>() -
                              f.template operator()<b, a, c>() - f.template operator()<c, b, a>());
}

template <int a, int b, int c, typename F> AMREX_FORCE_INLINE constexpr auto cyclic_permutation_sum(F &&f) {
    return (1.0 / 3.0) *
           (f.template operator()<a, b, c>() + f.template operator()<b, c, a>() + f.template operator()<c, a, b>());
}

} // namespace treesap::tensor
