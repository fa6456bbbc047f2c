// static_for.hpp -- compile-time loop over an integer range, calling
// f.template operator()<I>() for I = Start .. End-1 in increasing order.
// Same interface as the reference's util/static_for.hpp:38-68; the order of
// the calls fixes the order in which RHS components are evaluated.
#pragma once

#include "../amrex_types.hpp"

#include <utility>

namespace treesap {

namespace detail {
template <int Start, typename F, int... Is> constexpr void static_for_seq(F &f, std::integer_sequence<int, Is...>) {
    (f.template operator()<Start + Is>(), ...);
}
} // namespace detail

template <int Start, int End, typename F> AMREX_FORCE_INLINE constexpr void static_for(F &&f) {
    static_assert(Start <= End, "static_for: Start must not exceed End");
    detail::static_for_seq<Start>(f, std::make_integer_sequence<int, End - Start>{});
}

template <int End, typename F> AMREX_FORCE_INLINE constexpr void static_for(F &&f) {
    static_for<0, End>(std::forward<F>(f));
}

} // namespace treesap
