// cexprmath.hpp -- integer combinatorics used by the stencil weights and the
// Kreiss-Oliger coefficients, evaluated at compile time.
// Interface of util/cexprmath.hpp:32-105 (factorial, binom, power,
// alternating_sign, generalised_harmonic_number); the weights depend on the
// exact values and the left-to-right summation of the harmonic number.
#pragma once

#include "../amrex_types.hpp"

#include <cstdint>

namespace treesap::cexprmath {

constexpr std::int64_t factorial(int n) {
    std::int64_t f = 1;
    for (int m = 2; m <= n; ++m) f *= m;
    return f;
}

// C(n, k) by the multiplicative formula; exact in int64 for the sizes used (n <= 20).
constexpr std::int64_t binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    if (2 * k > n) k = n - k;
    std::int64_t c = 1;
    for (int m = 1; m <= k; ++m) c = c * (n - k + m) / m;
    return c;
}

constexpr std::int64_t power(std::int64_t base, int exp) {
    std::int64_t p = 1;
    for (int m = 0; m < exp; ++m) p *= base;
    return p;
}

constexpr int alternating_sign(int k) { return (k % 2 == 0) ? 1 : -1; }

// H_{n,m} = sum_{k=1..n} 1/k^m, summed in increasing k.
constexpr double generalised_harmonic_number(int n, int m) {
    double h = 0.0;
    for (int k = 1; k <= n; ++k) h += 1.0 / static_cast<double>(power(k, m));
    return h;
}

} // namespace treesap::cexprmath
