"""Closed-form checks for the oracle on periodic grids. TEST INFRASTRUCTURE ONLY.

A single Fourier mode exp(i k.x) is an eigenvector of every constant-coefficient
finite-difference operator on a periodic grid, so for a linear system
u_t = L u the RK4 step of integrators/SingleLevelRK4Integrator.hpp:95-157 is
multiplication by the degree-4 Taylor polynomial G = sum_{m<=4} (dt L)^m / m!,
and N steps are G^N. This module computes the symbols from weights obtained
independently of the oracle: Fornberg's algorithm in exact rational arithmetic
(B. Fornberg, "Generation of finite difference formulas on arbitrarily spaced
grids", Math. Comp. 51 (1988) 699-706), and the Kreiss-Oliger operator as
(-1)^{r+1} dx^{2r-1} / 4^r (D+ D-)^r expanded by the binomial theorem.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np


def fornberg(m: int, offsets) -> list:
    """Exact weights c[j] with sum_j c[j] f(offsets[j]) = f^{(m)}(0) (unit spacing)."""
    x = [Fraction(o) for o in offsets]
    n = len(x)
    c = [[Fraction(0)] * (m + 1) for _ in range(n)]
    c[0][0] = Fraction(1)
    c1 = Fraction(1)
    c4 = x[0]
    for i in range(1, n):
        mn = min(i, m)
        c2 = Fraction(1)
        c5 = c4
        c4 = x[i]
        for j in range(i):
            c3 = x[i] - x[j]
            c2 *= c3
            if j == i - 1:
                for k in range(mn, 0, -1):
                    c[i][k] = c1 * (k * c[i - 1][k - 1] - c5 * c[i - 1][k]) / c2
                c[i][0] = -c1 * c5 * c[i - 1][0] / c2
            for k in range(mn, 0, -1):
                c[j][k] = (c4 * c[j][k] - k * c[j][k - 1]) / c3
            c[j][0] = c4 * c[j][0] / c3
        c1 = c2
    return [c[j][m] for j in range(n)]


def central_weights(order: int, n: int) -> dict:
    """Centred (2n+1)-point weights for derivative `order`, keyed by offset -n..n."""
    offs = list(range(-n, n + 1))
    w = fornberg(order, offs)
    return dict(zip(offs, w))


def ko_weights(r: int) -> dict:
    """(-1)^{r+1} / 4^r (D+D-)^r at unit spacing: coefficient of f(s)."""
    # (D+D-)^r f(0) = sum_s (-1)^{r-s} binom(2r, r+s) f(s)
    sign = (-1) ** (r + 1)
    return {s: Fraction(sign * (-1) ** (r - s) * math.comb(2 * r, r + s), 4 ** r) for s in range(-r, r + 1)}


def _axis_symbols(n_st: int, ko_r: int, theta: float, dx: float):
    w1 = central_weights(1, n_st)
    w2 = central_weights(2, n_st)
    d1 = sum(float(w) * complex(math.cos(s * theta), math.sin(s * theta)) for s, w in w1.items()) / dx
    d2 = sum(float(w) * math.cos(s * theta) for s, w in w2.items()) / (dx * dx)
    ko = 0.0
    if ko_r > 0:
        kw = ko_weights(ko_r)
        ko = sum(float(w) * math.cos(s * theta) for s, w in kw.items()) / dx
    return d1, d2, ko


def _sym2(a: int, b: int) -> int:
    row, col = max(a, b), min(a, b)
    return row * (row + 1) // 2 + col


def rhs_matrix(case, mode) -> np.ndarray:
    """Symbol L(k) of the decorated linear RHS for integer mode numbers `mode` (periodic box)."""
    get = (lambda k: case[k]) if isinstance(case, dict) else (lambda k: getattr(case, k))
    n = get("n")
    lo, hi = get("prob_lo"), get("prob_hi")
    dx = [(hi[a] - lo[a]) / n[a] for a in range(3)]
    syms = [_axis_symbols(get("stencil_n"), get("ko_order"), 2 * math.pi * mode[a] / n[a], dx[a]) for a in range(3)]
    d1 = [s[0] for s in syms]
    d2 = [s[1] for s in syms]
    kterm = get("ko_strength") * sum(s[2] for s in syms) if get("ko_order") > 0 else 0.0
    system = get("system")
    if system in ("phipi", "nonlinear"):
        # nonlinear with gamma = 0 reduces to the linear system
        L = np.zeros((2, 2), complex)
        L[0, 1] = 1.0
        L[1, 0] = d2[0] + d2[1] + d2[2]
    elif system == "phiu":
        L = np.zeros((4, 4), complex)
        for a in range(3):
            L[0, 1 + a] = d1[a]
            L[1 + a, 0] = d1[a]
    elif system == "tensor":
        L = np.zeros((12, 12), complex)

        def S(p, q):  # symbol of derivative<n, p, q>: unmixed D2 if p == q else nested D1 D1
            return d2[p] if p == q else d1[p] * d1[q]

        for a in range(3):
            for b in range(a + 1):
                h_ab, k_ab = _sym2(a, b), 6 + _sym2(a, b)
                L[h_ab, k_ab] += -2.0
                for m in range(3):
                    L[k_ab, _sym2(m, b)] += 0.5 * S(m, a)
                    L[k_ab, _sym2(m, a)] += 0.5 * S(m, b)
                    L[k_ab, _sym2(m, m)] += -0.5 * S(a, b)
                    L[k_ab, h_ab] += -0.5 * S(m, m)
    else:
        raise ValueError(system)
    return L + kterm * np.eye(L.shape[0])


def amplification(case, mode, scheme: str = "rk4") -> np.ndarray:
    get = (lambda k: case[k]) if isinstance(case, dict) else (lambda k: getattr(case, k))
    hL = get("dt") * rhs_matrix(case, mode)
    eye = np.eye(hL.shape[0], dtype=complex)
    if scheme == "euler":
        return eye + hL
    G = eye.copy()
    term = eye.copy()
    for m in range(1, 5):
        term = term @ hL / m
        G = G + term
    return G


def mode_field(case, mode, amps) -> np.ndarray:
    """Real state (ncomp, nz, ny, nx): Re(amps[c] exp(i k.x)) at cell centres."""
    get = (lambda k: case[k]) if isinstance(case, dict) else (lambda k: getattr(case, k))
    n = get("n")
    lo, hi = get("prob_lo"), get("prob_hi")
    phase = np.zeros((n[2], n[1], n[0]))
    for a, shape in ((0, (1, 1, n[0])), (1, (1, n[1], 1)), (2, (n[2], 1, 1))):
        dx = (hi[a] - lo[a]) / n[a]
        x = lo[a] + (np.arange(n[a]) + 0.5) * dx
        kx = 2 * math.pi * mode[a] / (hi[a] - lo[a])
        phase = phase + (kx * x).reshape(shape)
    e = np.exp(1j * phase)
    return np.stack([np.real(A * e) for A in amps])


def evolve(case, mode, amps, nsteps: int, scheme: str = "rk4") -> np.ndarray:
    G = amplification(case, mode, scheme)
    out = np.linalg.matrix_power(G, nsteps) @ np.asarray(amps, dtype=complex)
    return mode_field(case, mode, out)
