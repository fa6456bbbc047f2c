"""Pins the CPU oracle (oracle/oracle.c) against known answers, closed forms and
brute force. No GPU. Each check is chosen so that a plausible slip in the
oracle (a dropped term, a wrong sign or index, a transposed operand, a wrong
grid factor) fails at least one of them."""
import itertools
import json
import math
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from oracle import fourier
from oracle import oracle as orc

GOLDEN = Path(__file__).resolve().parent / "golden" / "central_differences.json"


def _case(**kw):
    base = dict(system="phipi", n=(12, 10, 8), stencil_n=2, ghost=2, ko_order=0, ko_strength=0.0,
                prob_lo=(0.0, 0.0, 0.0), prob_hi=(1.0, 1.0, 1.0), periodic=(1, 1, 1), bc="noop",
                abl_thickness=0, abl_max=0.0, abl_profile="quadratic", target=(0.0,) * 12, gamma=0.0, dt=0.01)
    base.update(kw)
    return base


# ---------------------------------------------------------------------------
# weights
# ---------------------------------------------------------------------------
def test_weights_match_textbook_table():
    g = json.loads(GOLDEN.read_text())
    for kind, idx in (("first", 0), ("second", 1), ("third", 2)):
        for n, vals in g[kind].items():
            n = int(n)
            w = orc.weights(n)[idx]
            want = [Fraction(v) for v in vals]
            if kind == "second":
                got = w[0:n + 1]
            else:
                got = w[1:n + 1]
            assert len(got) == len(want)
            for a, b in zip(got, want):
                assert a == pytest.approx(float(b), rel=2e-15, abs=0), (kind, n, got, want)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_weights_match_fornberg(n):
    w1, w2, w3 = orc.weights(n)
    f1 = fourier.central_weights(1, n)
    f2 = fourier.central_weights(2, n)
    for p in range(1, n + 1):
        # single rounding of an exact rational: correctly rounded
        assert w1[p] == float(f1[p])
        assert w2[p] == float(f2[p])
        assert orc.lib().or_w1(n, -p) == -w1[p]
    assert w2[0] == pytest.approx(float(f2[0]), rel=4e-16)
    if n >= 2:
        f3 = fourier.central_weights(3, n)
        for p in range(1, n + 1):
            assert w3[p] == pytest.approx(float(f3[p]), rel=1e-14)


def test_harmonic_and_binomial():
    L = orc.lib()
    for n in range(1, 8):
        assert L.or_harmonic(n, 2) == pytest.approx(sum(1.0 / k ** 2 for k in range(1, n + 1)), rel=1e-16)
        for k in range(0, n + 1):
            assert L.or_binom(n, k) == math.comb(n, k)


@pytest.mark.parametrize("r", [1, 2, 3, 4])
def test_ko_coefficients_and_symbol(r):
    g = json.loads(GOLDEN.read_text())["kreiss_oliger"][str(r)]
    coefs = orc.ko_coefs(r)
    assert coefs[0] == g["centre"]
    assert coefs[1:] == g["offsets"]
    exact = fourier.ko_weights(r)
    for s in range(r + 1):
        assert Fraction(coefs[s], 4 ** r) == exact[s]
    # symbol: norm * (c0 + 2 sum c_s cos(s t)) = -sin^{2r}(t/2)
    for t in np.linspace(0, math.pi, 17):
        sym = (coefs[0] + 2 * sum(coefs[s] * math.cos(s * t) for s in range(1, r + 1))) / 4 ** r
        assert sym == pytest.approx(-math.sin(t / 2) ** (2 * r), abs=1e-14)


# ---------------------------------------------------------------------------
# polynomial exactness of the point stencils (interior, no ghosts involved)
# ---------------------------------------------------------------------------
def _poly_case(n_st, dx=(0.1, 0.125, 0.2)):
    N = 4 * n_st + 3
    hi = tuple(N * d for d in dx)
    return _case(n=(N, N, N), stencil_n=n_st, ghost=n_st, prob_hi=hi, periodic=(0, 0, 0), bc="dirichlet")


def _centres(case):
    n = case["n"]
    return [(np.arange(n[a]) + 0.5) * (case["prob_hi"][a] / n[a]) for a in range(3)]


@pytest.mark.parametrize("n_st", [1, 2, 3, 4])
def test_first_and_second_derivative_polynomial_exactness(n_st):
    case = _poly_case(n_st)
    xs = _centres(case)
    N = case["n"][0]
    rng = np.random.default_rng(n_st)
    deg = 2 * n_st
    coef = rng.uniform(-1, 1, size=(deg + 2,))
    s = orc.OracleSolver(case)
    for axis in range(3):
        # p(x_axis) of degree 2n (D1 exact) and 2n+1 (D2 exact)
        for d, kind in ((deg, 1), (deg + 1, 2)):
            c = coef[: d + 1]
            field = np.zeros(s.shape)
            grid = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")[2 - axis]
            field[0] = np.polynomial.polynomial.polyval(grid, c)
            s.set_state(field)
            s.fill_ghosts(0)
            i = j = k = N // 2
            x0 = xs[axis][N // 2]
            dc = np.polynomial.polynomial.polyder(c, kind)
            want = np.polynomial.polynomial.polyval(x0, dc)
            got = s.point_derivative([axis] * kind, i, j, k)
            scale = max(1.0, abs(want), np.abs(np.polynomial.polynomial.polyval(x0, c)))
            assert got == pytest.approx(want, abs=1e-9 * scale), (axis, kind)
            # one degree more is NOT exact for D1 (truncation term present)
            if kind == 1:
                c2 = np.zeros(deg + 2)
                c2[deg + 1] = 1.0
                field[0] = np.polynomial.polynomial.polyval(grid, c2)
                s.set_state(field)
                got2 = s.point_derivative([axis], i, j, k)
                want2 = (deg + 1) * x0 ** deg
                assert abs(got2 - want2) > 1e-9 * max(1.0, abs(want2))


@pytest.mark.parametrize("n_st", [2, 3])
def test_mixed_and_third_derivatives_polynomial(n_st):
    case = _poly_case(n_st)
    xs = _centres(case)
    N = case["n"][0]
    s = orc.OracleSolver(case)
    Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    field = np.zeros(s.shape)
    # f = x^2 y^3 z + 3 x y - z^2 x^3
    field[0] = X ** 2 * Y ** 3 * Z + 3 * X * Y - Z ** 2 * X ** 3
    s.set_state(field)
    c = N // 2
    x, y, z = xs[0][c], xs[1][c], xs[2][c]
    want = {
        (0, 1): 2 * x * 3 * y ** 2 * z + 3,
        (1, 0): 2 * x * 3 * y ** 2 * z + 3,
        (0, 2): 2 * x * y ** 3 - 2 * z * 3 * x ** 2,
        (2, 0): 2 * x * y ** 3 - 2 * z * 3 * x ** 2,
        (1, 2): x ** 2 * 3 * y ** 2,
        (2, 1): x ** 2 * 3 * y ** 2,
        (0, 0, 1): 2 * 3 * y ** 2 * z,
        (0, 1, 2): 2 * x * 3 * y ** 2,
        (0, 0, 0): -6 * z ** 2,
        (1, 1, 1): 6 * x ** 2 * z,
    }
    for dirs, w in want.items():
        got = s.point_derivative(list(dirs), c, c, c)
        assert got == pytest.approx(w, rel=1e-8, abs=1e-8), dirs


# ---------------------------------------------------------------------------
# ABL sigma and ghost fill by brute force
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("profile", ["linear", "quadratic", "cubic"])
def test_abl_sigma_brute_force(profile):
    n = (11, 9, 10)
    T = 3
    case = _case(n=n, periodic=(0, 0, 0), bc="dirichlet", abl_thickness=T, abl_max=1.7, abl_profile=profile)
    s = orc.OracleSolver(case)
    bulk = [(a, b, c) for a in range(T, n[0] - T) for b in range(T, n[1] - T) for c in range(T, n[2] - T)]
    bulk = np.array(bulk)
    p = {"linear": lambda v: v, "quadratic": lambda v: v * v, "cubic": lambda v: v ** 3}[profile]
    for i, j, k in itertools.product(range(n[0]), range(n[1]), range(n[2])):
        d = np.sqrt(((bulk - np.array([i, j, k])) ** 2).sum(axis=1)).min()
        want = 0.0 if d == 0 else 1.7 * p(min(d / T, 1.0))
        assert s.abl_sigma(i, j, k) == pytest.approx(want, rel=1e-15, abs=0)


def _sigma_brute(n, T, smax, profile):
    """sigma(i,j,k) of ABLDecorator.hpp:115-132 by brute force: distance to the nearest bulk cell."""
    p = {"linear": lambda v: v, "quadratic": lambda v: v * v, "cubic": lambda v: v ** 3}[profile]
    lo = np.array([T, T, T])
    hi = np.array([n[0] - 1 - T, n[1] - 1 - T, n[2] - 1 - T])
    sig = np.zeros((n[2], n[1], n[0]))
    for k, j, i in itertools.product(range(n[2]), range(n[1]), range(n[0])):
        iv = np.array([i, j, k])
        d = float(np.sqrt(((iv - np.clip(iv, lo, hi)) ** 2).sum()))
        sig[k, j, i] = 0.0 if d == 0 else smax * p(min(d / T, 1.0))
    return sig


def _rk4_affine_step(s, t0, t1, phi, pi, dt):
    """One exact RK4 step of the spatially uniform ODE phi' = pi - s (phi - t0), pi' = -s (pi - t1).

    RK4 on a linear autonomous system is the degree-4 Taylor polynomial of the
    propagator; the affine forcing enters through the augmented matrix
    M = [[-s, 1, s t0], [0, -s, s t1], [0, 0, 0]] acting on (phi, pi, 1)."""
    M = np.array([[-s, 1.0, s * t0], [0.0, -s, s * t1], [0.0, 0.0, 0.0]]) * dt
    P = np.eye(3) + M + M @ M / 2 + M @ M @ M / 6 + M @ M @ M @ M / 24
    z = P @ np.array([phi, pi, 1.0])
    return z[0], z[1]


ABL_CLOSED_CASES = [
    # stencil half-width, KO order, box, layer thickness, max rate, profile, targets (phi, Pi)
    (1, 0, (22, 21, 23), 6, 2.0, "quadratic", (0.25, -0.375)),
    (2, 3, (50, 49, 51), 12, 1.5, "linear", (-0.5, 0.125)),
]


@pytest.mark.parametrize("n_st,ko,n,T,smax,profile,tgt", ABL_CLOSED_CASES)
def test_abl_damping_closed_form(n_st, ko, n, T, smax, profile, tgt):
    """ABLDecorator.hpp:102-108: F<c> - sigma(iv) (u_c - target[c]).

    A spatially constant (phi, Pi) = (a, p) on a periodic box stays constant
    wherever sigma is constant over the cells one RK4 step reads (four stages
    of axis stencils of half-width h = max(n, r): offsets d with
    sum_a ceil(|d_a| / h) <= 4); there the stencil terms vanish and the step
    is the RK4 polynomial of the damped affine ODE. Cells in the bulk
    (sigma = 0) and in the saturated corners (sigma = max) are checked against
    that closed form."""
    case = _case(n=n, stencil_n=n_st, ghost=max(n_st, ko), ko_order=ko, ko_strength=0.1 if ko else 0.0,
                 abl_thickness=T, abl_max=smax, abl_profile=profile, target=tgt + (0.0,) * 10, dt=0.25 / n[0])
    a, p = 0.75, 0.5
    s = orc.OracleSolver(case)
    st = np.zeros(s.shape)
    st[0] = a
    st[1] = p
    s.set_state(st)
    s.rk4_steps(1)
    got = s.get_state()
    sig = _sigma_brute(n, T, smax, profile)
    h = max(n_st, ko)
    rad = 4 * h
    reach = [(dx, dy, dz) for dz in range(-rad, rad + 1) for dy in range(-rad, rad + 1) for dx in range(-rad, rad + 1)
             if sum(-(-abs(v) // h) for v in (dx, dy, dz)) <= 4]
    reach = np.array(reach)
    checked = {0.0: 0, smax: 0}
    for k, j, i in itertools.product(range(n[2]), range(n[1]), range(n[0])):
        sv = sig[k, j, i]
        if sv not in checked:
            continue
        ii = (i + reach[:, 0]) % n[0]
        jj = (j + reach[:, 1]) % n[1]
        kk = (k + reach[:, 2]) % n[2]
        if not np.all(sig[kk, jj, ii] == sv):
            continue
        want = _rk4_affine_step(sv, tgt[0], tgt[1], a, p, case["dt"])
        assert got[0, k, j, i] == pytest.approx(want[0], rel=1e-13, abs=1e-14), (i, j, k, sv)
        assert got[1, k, j, i] == pytest.approx(want[1], rel=1e-13, abs=1e-14), (i, j, k, sv)
        checked[sv] += 1
    assert checked[0.0] > 0 and checked[smax] > 0, checked


def test_abl_damping_closed_form_is_sensitive():
    """The closed form separates the reference's damping from plausible slips:
    a flipped sign, swapped targets, or a damping of phi only."""
    n_st, ko, n, T, smax, profile, tgt = ABL_CLOSED_CASES[0]
    case = _case(n=n, stencil_n=n_st, ghost=n_st, abl_thickness=T, abl_max=smax, abl_profile=profile,
                 target=tgt + (0.0,) * 10, dt=0.25 / n[0])
    a, p = 0.75, 0.5
    s = orc.OracleSolver(case)
    st = np.zeros(s.shape)
    st[0] = a
    st[1] = p
    s.set_state(st)
    s.rk4_steps(1)
    corner = s.get_state()[:, 0, 0, 0]
    dt = case["dt"]
    right = _rk4_affine_step(smax, tgt[0], tgt[1], a, p, dt)
    assert corner[0] == pytest.approx(right[0], rel=1e-13) and corner[1] == pytest.approx(right[1], rel=1e-13)
    flipped = _rk4_affine_step(-smax, tgt[0], tgt[1], a, p, dt)
    swapped = _rk4_affine_step(smax, tgt[1], tgt[0], a, p, dt)
    M = np.array([[-smax, 1.0, smax * tgt[0]], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]) * dt
    phi_only = (np.eye(3) + M + M @ M / 2 + M @ M @ M / 6 + M @ M @ M @ M / 24) @ np.array([a, p, 1.0])
    for wrong in (flipped, swapped, phi_only[:2]):
        assert max(abs(corner[0] - wrong[0]), abs(corner[1] - wrong[1])) > 1e-4


@pytest.mark.parametrize("periodic,bc", [((1, 1, 1), "noop"), ((0, 0, 0), "dirichlet"), ((1, 0, 1), "dirichlet"),
                                         ((0, 1, 0), "noop")])
def test_ghost_fill_brute_force(periodic, bc):
    n = (7, 6, 5)
    G = 3
    case = _case(system="phiu", n=n, stencil_n=2, ghost=G, periodic=periodic, bc=bc,
                 target=tuple(0.25 * (c + 1) for c in range(12)))
    s = orc.OracleSolver(case)
    rng = np.random.default_rng(7)
    st = rng.standard_normal(s.shape)
    s.set_state(st)
    before = s.get_register(0)
    s.fill_ghosts(0)
    reg = s.get_register(0)
    # expected: ghosts outside along a non-periodic axis -> target (Dirichlet) or untouched (NoOp)
    # otherwise the periodic image of a valid cell
    for c in range(4):
        for kk, jj, ii in itertools.product(range(n[2] + 2 * G), range(n[1] + 2 * G), range(n[0] + 2 * G)):
            iv = (ii - G, jj - G, kk - G)
            inside = all(0 <= iv[a] < n[a] for a in range(3))
            if inside:
                assert reg[c, kk, jj, ii] == st[c, iv[2], iv[1], iv[0]]
                continue
            phys = any(not periodic[a] and not (0 <= iv[a] < n[a]) for a in range(3))
            if phys:
                want = 0.25 * (c + 1) if bc == "dirichlet" else before[c, kk, jj, ii]
            else:
                w = [iv[a] % n[a] for a in range(3)]
                want = st[c, w[2], w[1], w[0]]
            assert reg[c, kk, jj, ii] == want


# ---------------------------------------------------------------------------
# end to end: single Fourier mode closed form on periodic grids
# ---------------------------------------------------------------------------
SYSTEM_CASES = [
    ("phipi", 2, 0), ("phipi", 2, 3), ("phipi", 4, 4), ("phipi", 1, 1),
    ("phiu", 1, 0), ("phiu", 3, 3),
    ("nonlinear", 2, 2),
    ("tensor", 2, 0), ("tensor", 3, 3),
]


@pytest.mark.parametrize("system,n_st,ko", SYSTEM_CASES)
@pytest.mark.parametrize("scheme", ["rk4", "euler"])
def test_fourier_mode_closed_form(system, n_st, ko, scheme):
    n = (16, 12, 10)
    G = max(n_st, ko)
    case = _case(system=system, n=n, stencil_n=n_st, ghost=G, ko_order=ko, ko_strength=0.17 if ko else 0.0,
                 prob_hi=(1.0, 0.8, 1.25), dt=0.2 / 16)
    s = orc.OracleSolver(case)
    rng = np.random.default_rng(hash((system, n_st, ko)) % 2 ** 32)
    for mode in [(1, 0, 0), (2, -1, 1), (3, 2, -2)]:
        amps = rng.standard_normal(s.ncomp) + 1j * rng.standard_normal(s.ncomp)
        init = fourier.mode_field(case, mode, amps)
        s.set_state(init)
        nsteps = 6
        (s.rk4_steps if scheme == "rk4" else s.euler_steps)(nsteps)
        got = s.get_state()
        want = fourier.evolve(case, mode, amps, nsteps, scheme)
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert err < 1e-12, (mode, err)


def test_fourier_closed_form_is_sensitive():
    """A KO strength off by 1% is visible to the closed-form check (the check has teeth)."""
    case = _case(n=(16, 12, 10), stencil_n=2, ghost=3, ko_order=3, ko_strength=0.2, dt=0.2 / 16)
    s = orc.OracleSolver(case)
    amps = np.array([1.0 + 0.5j, -0.3j])
    s.set_state(fourier.mode_field(case, (3, 2, 1), amps))
    s.rk4_steps(6)
    wrong = dict(case, ko_strength=0.202)
    want_wrong = fourier.evolve(wrong, (3, 2, 1), amps, 6)
    err = np.linalg.norm(s.get_state() - want_wrong) / np.linalg.norm(want_wrong)
    assert err > 1e-6


def test_time_and_step_counter():
    case = _case(dt=0.1)
    s = orc.OracleSolver(case)
    s.set_state(np.zeros(s.shape))
    s.rk4_steps(3)
    t = 0.0
    for _ in range(3):
        t += 0.1
    assert s.steps == 3 and s.time == t


def test_nonlinear_term_matches_direct_evaluation():
    """NonLinearRHS: pi_t = lap(phi) - gamma * pi * |grad phi|^2 (examples/02-temporaries/NonLinearRHS.hpp:39-71);
    the gamma term is checked against numpy derivatives of a smooth periodic field."""
    n = (24, 24, 24)
    case = _case(system="nonlinear", n=n, stencil_n=4, ghost=4, gamma=0.7)
    s = orc.OracleSolver(case)
    lin = orc.OracleSolver(dict(case, system="phipi"))
    X = (np.arange(24) + 0.5) / 24
    Z, Y, Xg = np.meshgrid(X, X, X, indexing="ij")
    st = np.zeros(s.shape)
    st[0] = np.sin(2 * np.pi * Xg) * np.cos(2 * np.pi * Y) + 0.3 * np.sin(2 * np.pi * Z)
    st[1] = 0.5 + 0.1 * np.cos(2 * np.pi * (Xg + Y))
    s.set_state(st)
    lin.set_state(st)
    r = s.eval_rhs(0)
    rl = lin.eval_rhs(0)
    assert np.array_equal(r[0], st[1])
    gx = 2 * np.pi * np.cos(2 * np.pi * Xg) * np.cos(2 * np.pi * Y)
    gy = -2 * np.pi * np.sin(2 * np.pi * Xg) * np.sin(2 * np.pi * Y)
    gz = 0.3 * 2 * np.pi * np.cos(2 * np.pi * Z)
    want = rl[1] - 0.7 * st[1] * (gx ** 2 + gy ** 2 + gz ** 2)
    assert np.max(np.abs(r[1] - want)) < 1e-3 * np.max(np.abs(want))


# ---------------------------------------------------------------------------
# discrete adjoint of the oracle: dot-product identity <G u, v> = <u, G^T v>
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n_st,ko,periodic,abl", [(2, 0, (0, 0, 0), 3), (4, 4, (0, 0, 0), 4), (2, 3, (1, 1, 1), 0),
                                                  (3, 3, (1, 0, 1), 0), (1, 1, (0, 1, 0), 2)])
def test_oracle_adjoint_dot_product(n_st, ko, periodic, abl):
    n = (14, 11, 13)
    case = _case(n=n, stencil_n=n_st, ghost=max(n_st, ko), ko_order=ko, ko_strength=0.2 if ko else 0.0,
                 periodic=periodic, bc="dirichlet" if not all(periodic) else "noop", abl_thickness=abl,
                 abl_max=2.0 if abl else 0.0, dt=0.3 / 14)
    rng = np.random.default_rng(n_st * 10 + ko)
    fwd = orc.OracleSolver(case)
    adj = orc.OracleSolver(case)
    for _ in range(3):
        u = rng.standard_normal(fwd.shape)
        v = rng.standard_normal(fwd.shape)
        fwd.set_state(u)
        fwd.rk4_steps(1)
        gu = fwd.get_state()
        gtv = adj.adjoint_step(v)
        lhs = float(np.sum(gu * v))
        rhs = float(np.sum(u * gtv))
        assert lhs == pytest.approx(rhs, rel=1e-13, abs=1e-13 * np.sqrt(np.sum(gu * gu) * np.sum(v * v)))


# ---------------------------------------------------------------------------
# discrete adjoint of the NonLinear system: the oracle's G'(u)^T against an
# independent numpy tangent-linear model of the RK4 step (padded-array
# stencils), which is itself checked against centred differences of the
# oracle's forward step
# ---------------------------------------------------------------------------
def _np_nonlinear_step_and_tangent(case, u, v, sigma):
    """(G(u), G'(u) v) of one RK4 step of the decorated NonLinear system
    F_phi = pi + KO - s (phi - t0), F_pi = Lap phi - gamma pi |D1 phi|^2 + KO - s (pi - t1)
    (examples/02-temporaries/NonLinearRHS.hpp:52-71), in exact-arithmetic form."""
    n = case["stencil_n"]
    G = case["ghost"]
    r = case["ko_order"]
    gam = case["gamma"]
    dx = [(case["prob_hi"][a] - case["prob_lo"][a]) / case["n"][a] for a in range(3)]
    w1, w2, _ = orc.weights(n)
    kc = orc.ko_coefs(r) if r else []
    tgt = case["target"]
    dirichlet = case["bc"] == "dirichlet"

    def pad(f, value):
        P = f
        for a in range(3):  # x is numpy axis 2
            ax = 2 - a
            width = [(0, 0)] * 3
            width[ax] = (G, G)
            P = np.pad(P, width, mode="wrap") if case["periodic"][a] else np.pad(P, width, constant_values=value)
        return P

    def sh(P, a, q):
        ax = 2 - a
        sl = [slice(G, P.shape[d] - G) for d in range(3)]
        sl[ax] = slice(G + q, P.shape[ax] - G + q)
        return P[tuple(sl)]

    def D1(P, a):
        return sum(w1[q] * (sh(P, a, q) - sh(P, a, -q)) for q in range(1, n + 1)) / dx[a]

    def D2(P, a):
        return (w2[0] * sh(P, a, 0) + sum(w2[q] * (sh(P, a, q) + sh(P, a, -q)) for q in range(1, n + 1))) / dx[a] ** 2

    def KO(P):
        if not r:
            return 0.0
        tot = 0.0
        for a in range(3):
            res = kc[0] * sh(P, a, 0) + sum(kc[q] * (sh(P, a, q) + sh(P, a, -q)) for q in range(1, r + 1))
            tot = tot + res / 2.0 ** (2 * r) / dx[a]
        return case["ko_strength"] * tot

    def F(w):
        Pphi = pad(w[0], tgt[0] if dirichlet else 0.0)
        Ppi = pad(w[1], tgt[1] if dirichlet else 0.0)
        g = [D1(Pphi, a) for a in range(3)]
        gsq = g[0] ** 2 + g[1] ** 2 + g[2] ** 2
        fphi = w[1] + KO(Pphi) - sigma * (w[0] - tgt[0])
        fpi = sum(D2(Pphi, a) for a in range(3)) - gam * w[1] * gsq + KO(Ppi) - sigma * (w[1] - tgt[1])
        return np.stack([fphi, fpi]), g, gsq

    def J(w, dw):
        _, g, gsq = F(w)
        Pd = pad(dw[0], 0.0)
        Pdpi = pad(dw[1], 0.0)
        dg = [D1(Pd, a) for a in range(3)]
        dphi = dw[1] + KO(Pd) - sigma * dw[0]
        dpi = (sum(D2(Pd, a) for a in range(3)) - gam * (dw[1] * gsq + 2.0 * w[1] * (g[0] * dg[0] + g[1] * dg[1] +
                                                                                       g[2] * dg[2]))
               + KO(Pdpi) - sigma * dw[1])
        return np.stack([dphi, dpi])

    dt = case["dt"]
    a_next = [0.5 * dt, 0.5 * dt, dt]
    b = [dt / 6, dt / 3, dt / 3, dt / 6]
    a, da = u, v
    out, dout = u.copy(), v.copy()
    for q in range(4):
        k, _, _ = F(a)
        dk = J(a, da)
        out = out + b[q] * k
        dout = dout + b[q] * dk
        if q < 3:
            a, da = u + a_next[q] * k, v + a_next[q] * dk
    return out, dout


NL_ADJ_CASES = [  # stencil n, KO order, periodic, ABL thickness, gamma
    (2, 2, (1, 1, 1), 0, 0.3), (3, 3, (0, 0, 0), 3, 0.5), (1, 0, (1, 0, 1), 0, 0.2), (4, 4, (1, 1, 1), 0, 0.4),
    (2, 0, (0, 1, 0), 2, 0.6)]


def _nl_case(n_st, ko, periodic, abl, gamma, n=(13, 11, 12)):
    tgt = (0.05, -0.08) + (0.0,) * 10 if not all(periodic) else (0.0,) * 12
    return _case(system="nonlinear", n=n, stencil_n=n_st, ghost=max(n_st, ko), ko_order=ko,
                 ko_strength=0.15 if ko else 0.0, periodic=periodic, bc="dirichlet" if not all(periodic) else "noop",
                 abl_thickness=abl, abl_max=2.0 if abl else 0.0, target=tgt, gamma=gamma, dt=0.25 / 13,
                 prob_hi=(1.0, 0.9, 1.1))


@pytest.mark.parametrize("n_st,ko,periodic,abl,gamma", NL_ADJ_CASES)
def test_oracle_nonlinear_adjoint_against_tangent_linear_model(n_st, ko, periodic, abl, gamma):
    case = _nl_case(n_st, ko, periodic, abl, gamma)
    n = case["n"]
    sigma = _sigma_brute(n, abl, 2.0, "quadratic") if abl else np.zeros(n[::-1])
    rng = np.random.default_rng(100 + 10 * n_st + ko)
    s = orc.OracleSolver(case)
    u = 0.5 * rng.standard_normal(s.shape)
    v = rng.standard_normal(s.shape)
    lam = rng.standard_normal(s.shape)
    # the numpy step equals the oracle's forward step (to rounding) ...
    g_np, tlm = _np_nonlinear_step_and_tangent(case, u, v, sigma)
    s.set_state(u)
    s.rk4_steps(1)
    assert np.max(np.abs(s.get_state() - g_np)) < 1e-12 * np.max(np.abs(g_np))
    # ... its tangent equals centred differences of the oracle's forward step ...
    eps = 1e-5
    fd = []
    for sgn in (1.0, -1.0):
        t = orc.OracleSolver(case)
        t.set_state(u + sgn * eps * v)
        t.rk4_steps(1)
        fd.append(t.get_state())
    fd = (fd[0] - fd[1]) / (2 * eps)
    assert np.max(np.abs(fd - tlm)) < 1e-7 * np.max(np.abs(tlm))
    # ... and the oracle's adjoint is its transpose: <G'(u) v, lam> = <v, G'(u)^T lam>
    adj = orc.OracleSolver(case).adjoint_step_nl(lam, u)
    lhs, rhs = float(np.sum(tlm * lam)), float(np.sum(v * adj))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # the check has teeth: the linear (phi, Pi) transpose alone misses the nonlinear term
    lin = orc.OracleSolver(dict(case, system="phipi", gamma=0.0)).adjoint_step(lam)
    assert abs(float(np.sum(v * lin)) - lhs) > 1e-6 * abs(lhs)


def test_oracle_adjoint_is_not_the_forward():
    """The transpose differs from the forward step where the operator is not symmetric (the check has teeth)."""
    case = _case(n=(12, 10, 8), stencil_n=2, ghost=2, dt=0.03)
    s = orc.OracleSolver(case)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(s.shape)
    s.set_state(v)
    s.rk4_steps(1)
    assert np.max(np.abs(s.get_state() - orc.OracleSolver(case).adjoint_step(v))) > 1e-3
