// stencil_impl.hpp -- centred finite-difference weights and the derivative
// evaluators behind treesap::stencil::derivative.
//
// Weights (closed forms of util/stencil_impl.hpp:56-140): for half-width n,
//   first:  w1(p) = (-1)^(p+1) (n!)^2 / (p (n-p)! (n+p)!),  odd in p;
//   second: w2(0) = -2 H_{n,2}; w2(p) = 2 (-1)^(p+1) (n!)^2 / (p^2 (n-p)! (n+p)!);
//   third:  w3(p) = (1 - p^2 H_{n,2}) * 6 (-1)^(p+1) (n!)^2 / (p^3 (n-p)! (n+p)!), odd in p.
// Each is formed as an int64 numerator and denominator and one double
// division, the centre weight of the second derivative as the harmonic sum;
// this fixes every weight to the last bit.
//
// Derivatives (util/stencil_impl.hpp:206-378), each a left-to-right sum over
// s = 1..n times a grid factor:
//   D_a u        = (1/dx_a) * sum_s w1(s) (u[+s e_a] - u[-s e_a])
//   D_a D_a u    = (1/(dx_a dx_a)) * (w2(0) u + sum_s w2(s) (u[+s e_a] + u[-s e_a]))
//   D_a D_b u    = (1/dx_a) * sum_s w1(s) (D_b u[+s e_a] - D_b u[-s e_a])      (a != b)
//   D_a D_a D_a u = (1/(dx_a^3)) * sum_s w3(s) (u[+s e_a] - u[-s e_a])
//   two equal directions: first derivative along the single one of the
//   unmixed second derivative along the repeated one; three distinct: first
//   derivative along the first of the mixed second derivative of the others.
#pragma once

#include "../amrex_types.hpp"
#include "cexprmath.hpp"

namespace treesap {

namespace stencil {
enum dir : int { x = 0, y = 1, z = 2 };
} // namespace stencil

namespace stencil_impl {

namespace weights {

template <int n> constexpr amrex::Real first(int offset) {
    static_assert(n >= 1, "first derivatives need a half-width of at least 1");
    using namespace cexprmath;
    if (offset == 0) return 0.0;
    const int p = offset < 0 ? -offset : offset;
    const std::int64_t num = alternating_sign(p + 1) * power(factorial(n), 2);
    const std::int64_t den = p * factorial(n - p) * factorial(n + p);
    const double w = static_cast<double>(num) / static_cast<double>(den);
    return offset > 0 ? w : -w;
}

template <int n> constexpr amrex::Real second(int offset) {
    static_assert(n >= 1, "second derivatives need a half-width of at least 1");
    using namespace cexprmath;
    if (offset == 0) return -2.0 * generalised_harmonic_number(n, 2);
    const int p = offset < 0 ? -offset : offset;
    const std::int64_t num = 2 * alternating_sign(p + 1) * power(factorial(n), 2);
    const std::int64_t den = p * p * factorial(n - p) * factorial(n + p);
    return static_cast<double>(num) / static_cast<double>(den);
}

template <int n> constexpr amrex::Real third(int offset) {
    static_assert(n >= 2, "third derivatives need a half-width of at least 2");
    using namespace cexprmath;
    if (offset == 0) return 0.0;
    const int p = offset < 0 ? -offset : offset;
    const std::int64_t num = 6 * alternating_sign(p + 1) * power(factorial(n), 2);
    const std::int64_t den = power(p, 3) * factorial(n - p) * factorial(n + p);
    const double w = (1.0 - static_cast<double>(p * p) * generalised_harmonic_number(n, 2)) *
                     static_cast<double>(num) / static_cast<double>(den);
    return offset > 0 ? w : -w;
}

} // namespace weights

// unit step along axis a
template <int a> constexpr int ex = a == 0 ? 1 : 0;
template <int a> constexpr int ey = a == 1 ? 1 : 0;
template <int a> constexpr int ez = a == 2 ? 1 : 0;

template <int n, typename T, int... Dirs> struct derivative_impl {
    static_assert(sizeof...(Dirs) >= 1 && sizeof...(Dirs) <= 3, "derivatives of order 1 to 3 only");
};

template <int n, typename T, int a> struct derivative_impl<n, T, a> {
    AMREX_GPU_HOST_DEVICE static amrex::Real eval(int i, int j, int k, int c, const amrex::Array4<T> &u, const amrex::GeometryData &g) {
        const amrex::Real gf = 1.0 / g.CellSize(a);
        amrex::Real acc = 0.0;
        for (int s = 1; s <= n; ++s)
            acc += weights::first<n>(s) *
                   (u(i + s * ex<a>, j + s * ey<a>, k + s * ez<a>, c) - u(i - s * ex<a>, j - s * ey<a>, k - s * ez<a>, c));
        return gf * acc;
    }
};

template <int n, typename T, int a, int b> struct derivative_impl<n, T, a, b> {
    AMREX_GPU_HOST_DEVICE static amrex::Real eval(int i, int j, int k, int c, const amrex::Array4<T> &u, const amrex::GeometryData &g) {
        if constexpr (a == b) {
            const amrex::Real gf = 1.0 / (g.CellSize(a) * g.CellSize(b));
            amrex::Real acc = weights::second<n>(0) * u(i, j, k, c);
            for (int s = 1; s <= n; ++s)
                acc += weights::second<n>(s) * (u(i + s * ex<a>, j + s * ey<a>, k + s * ez<a>, c) +
                                                u(i - s * ex<a>, j - s * ey<a>, k - s * ez<a>, c));
            return gf * acc;
        } else {
            const amrex::Real gf = 1.0 / g.CellSize(a);
            using inner = derivative_impl<n, T, b>;
            amrex::Real acc = 0.0;
            for (int s = 1; s <= n; ++s)
                acc += weights::first<n>(s) * (inner::eval(i + s * ex<a>, j + s * ey<a>, k + s * ez<a>, c, u, g) -
                                               inner::eval(i - s * ex<a>, j - s * ey<a>, k - s * ez<a>, c, u, g));
            return gf * acc;
        }
    }
};

template <int n, typename T, int a, int b, int d> struct derivative_impl<n, T, a, b, d> {
    AMREX_GPU_HOST_DEVICE static amrex::Real eval(int i, int j, int k, int c, const amrex::Array4<T> &u, const amrex::GeometryData &g) {
        if constexpr (a == b && b == d) {
            const amrex::Real gf = 1.0 / (g.CellSize(a) * g.CellSize(a) * g.CellSize(a));
            amrex::Real acc = 0.0;
            for (int s = 1; s <= n; ++s)
                acc += weights::third<n>(s) *
                       (u(i + s * ex<a>, j + s * ey<a>, k + s * ez<a>, c) - u(i - s * ex<a>, j - s * ey<a>, k - s * ez<a>, c));
            return gf * acc;
        } else {
            // outer axis o takes a first derivative of the inner second derivative (p, q)
            constexpr bool two_equal = (a == b || a == d || b == d);
            constexpr int o = !two_equal ? a : (a == b) ? d : (a == d) ? b : a;
            constexpr int p = !two_equal ? b : (a == b || a == d) ? a : b;
            constexpr int q = !two_equal ? d : p;
            using inner = derivative_impl<n, T, p, q>;
            const amrex::Real gf = 1.0 / g.CellSize(o);
            amrex::Real acc = 0.0;
            for (int s = 1; s <= n; ++s)
                acc += weights::first<n>(s) * (inner::eval(i + s * ex<o>, j + s * ey<o>, k + s * ez<o>, c, u, g) -
                                               inner::eval(i - s * ex<o>, j - s * ey<o>, k - s * ez<o>, c, u, g));
            return gf * acc;
        }
    }
};

} // namespace stencil_impl
} // namespace treesap
