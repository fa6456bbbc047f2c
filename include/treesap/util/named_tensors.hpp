// named_tensors.hpp -- Kronecker delta and Levi-Civita symbol as compile-time
// constants (util/named_tensors.hpp:25-62).
#pragma once

namespace treesap::tensor {

template <int a, int b> constexpr int kronecker = (a == b) ? 1 : 0;

namespace detail {
// sign of the permutation (i_0..i_{d-1}) of (0..d-1); 0 if an index repeats or is out of range
template <int... I> constexpr int permutation_sign() {
    constexpr int d = sizeof...(I);
    int p[d > 0 ? d : 1] = {I...};
    for (int m = 0; m < d; ++m)
        if (p[m] < 0 || p[m] >= d) return 0;
    int inversions = 0;
    for (int m = 0; m < d; ++m)
        for (int q = m + 1; q < d; ++q) {
            if (p[m] == p[q]) return 0;
            if (p[m] > p[q]) ++inversions;
        }
    return inversions % 2 ? -1 : 1;
}
} // namespace detail

template <int... Indices> constexpr int levi_civita = detail::permutation_sign<Indices...>();

static_assert(levi_civita<0, 1, 2> == 1 && levi_civita<1, 0, 2> == -1 && levi_civita<2, 0, 1> == 1 &&
              levi_civita<0, 0, 2> == 0);

} // namespace treesap::tensor
