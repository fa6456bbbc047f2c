"""Model descriptions and seeded synthetic initial data of the five workloads.

A ``Case`` is the runtime content of one TreESaP model: the registered RHS
(system), the geometry (pde-model/PDEModel.hpp:70-77 with amrex::Geometry),
the KO decorator (rhs-decorators/KODecorator.hpp:32-74), the absorbing layer
(pde-model/ABLPDEModel.hpp:47-50, rhs-decorators/ABLDecorator.hpp:58-77) and
the fixed time step ``dt_hint`` (PDEModel.hpp:122).

Initial data are generated on the host from a SplitMix64 stream (uniform =
top 53 bits x 2^-53, normal = Box-Muller), so every consumer of a case sees
identical bytes. Arrays are float64 of shape (ncomp, nz, ny, nx).

Workloads (sizes, stencils and decorators as the benchmark configurations
state them; unstated details are fixed here):
  wave64    (phi, Pi) wave 64^3, periodic unit cube, 4th-order stencils (n=2),
            KO order 3 (6th derivative) sigma=0.1, dt=dx/4, 100 steps, one
            Gaussian (width 0.1, centre U[0.3,0.7]^3)
  wave256   (phi, Pi) wave 256^3, non-periodic unit cube, 8th-order (n=4),
            KO order 4 sigma=0.2, 16-cell quadratic absorbing layer (max 2.0,
            Dirichlet target 0), dt=dx/4, 1000 steps, 8 seeded Gaussians
  tensor192 linearised ADM (h_ij, K_ij), 12 comps, 192^3, 6th-order (n=3), KO
            order 3 sigma=0.1, 12-cell absorbing layer, 500 steps, 16 Fourier
            modes per component
  ensemble  512 periodic (phi, Pi) instances, side from {32,48,64,96,128},
            n from {2,3,4}, KO order n, sigma U[0.05,0.3], 200 steps
  adjoint128 (phi, Pi) wave 128^3, n=2, no KO, 12-cell absorbing layer,
            500 steps, 16 sources, 64 receivers
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib


# ---------------------------------------------------------------------------
# model description
# ---------------------------------------------------------------------------
@dataclass
class Case:
    system: str = "phipi"
    n: tuple = (32, 32, 32)
    stencil_n: int = 2
    ghost: int = 2
    ko_order: int = 0
    ko_strength: float = 0.0
    prob_lo: tuple = (0.0, 0.0, 0.0)
    prob_hi: tuple = (1.0, 1.0, 1.0)
    periodic: tuple = (1, 1, 1)
    bc: str = "noop"
    abl_thickness: int = 0
    abl_max: float = 0.0
    abl_profile: str = "quadratic"
    target: tuple = field(default_factory=lambda: (0.0,) * _lib.TSP_MAX_COMP)
    gamma: float = 0.0
    dt: float = 0.0

    @property
    def ncomp(self) -> int:
        return _lib.NCOMP[self.system]

    @property
    def cells(self) -> int:
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])

    @property
    def shape(self) -> tuple:
        return (self.ncomp, int(self.n[2]), int(self.n[1]), int(self.n[0]))

    def dx(self) -> tuple:
        """Geometry::CellSize: (prob_hi - prob_lo) / n_cells."""
        return tuple((self.prob_hi[a] - self.prob_lo[a]) / self.n[a] for a in range(3))

    def to_dict(self) -> dict:
        return dataclasses.asdict(self)

    def replace(self, **kw) -> "Case":
        return dataclasses.replace(self, **kw)


def make_desc(case: Case) -> _lib.TspDesc:
    d = _lib.TspDesc()
    d.system = _lib.SYSTEMS[case.system]
    for a in range(3):
        d.n[a] = int(case.n[a])
        d.prob_lo[a] = float(case.prob_lo[a])
        d.prob_hi[a] = float(case.prob_hi[a])
        d.periodic[a] = int(case.periodic[a])
    d.stencil_n = int(case.stencil_n)
    d.ghost = int(case.ghost)
    d.ko_order = int(case.ko_order)
    d.bc = _lib.BCS[case.bc]
    d.ko_strength = float(case.ko_strength)
    d.abl_thickness = int(case.abl_thickness)
    d.abl_profile = _lib.PROFILES[case.abl_profile]
    d.abl_max = float(case.abl_max)
    tgt = list(case.target) + [0.0] * (_lib.TSP_MAX_COMP - len(case.target))
    for c in range(_lib.TSP_MAX_COMP):
        d.target[c] = float(tgt[c])
    d.gamma = float(case.gamma)
    d.dt = float(case.dt)
    return d


def case_from_desc(d: _lib.TspDesc) -> Case:
    """The Case a descriptor describes (inverse of make_desc)."""
    inv = lambda table, v: next(k for k, val in table.items() if val == v)  # noqa: E731
    return Case(
        system=inv(_lib.SYSTEMS, d.system),
        n=tuple(int(d.n[a]) for a in range(3)),
        stencil_n=int(d.stencil_n),
        ghost=int(d.ghost),
        ko_order=int(d.ko_order),
        ko_strength=float(d.ko_strength),
        prob_lo=tuple(float(d.prob_lo[a]) for a in range(3)),
        prob_hi=tuple(float(d.prob_hi[a]) for a in range(3)),
        periodic=tuple(int(d.periodic[a]) for a in range(3)),
        bc=inv(_lib.BCS, d.bc),
        abl_thickness=int(d.abl_thickness),
        abl_max=float(d.abl_max),
        abl_profile=inv(_lib.PROFILES, d.abl_profile),
        target=tuple(float(d.target[c]) for c in range(_lib.TSP_MAX_COMP)),
        gamma=float(d.gamma),
        dt=float(d.dt),
    )


# ---------------------------------------------------------------------------
# SplitMix64 stream
# ---------------------------------------------------------------------------
_M64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) with the standard constants."""

    def __init__(self, seed: int):
        self.state = seed & _M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        u = (self.next_u64() >> 11) * (1.0 / (1 << 53))
        return lo + (hi - lo) * u

    def normal(self, mean: float = 0.0, sd: float = 1.0) -> float:
        u1 = 1.0 - self.uniform()  # (0, 1]
        u2 = self.uniform()
        return mean + sd * math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def choice(self, seq):
        return seq[self.next_u64() % len(seq)]


def _seed(tag: str, seed: int) -> int:
    h = 0xCBF29CE484222325
    for ch in tag.encode():
        h = ((h ^ ch) * 0x100000001B3) & _M64
    return (h ^ (seed * 0x9E3779B97F4A7C15)) & _M64


# ---------------------------------------------------------------------------
# initial data
# ---------------------------------------------------------------------------
def cell_centres(case: Case):
    """x = lo + (i + 0.5) dx per axis (SingleLevelIntegrator.hpp:84-86), broadcastable (z, y, x)."""
    dx = case.dx()
    xs = [case.prob_lo[a] + (np.arange(case.n[a], dtype=np.float64) + 0.5) * dx[a] for a in range(3)]
    return xs[0][None, None, :], xs[1][None, :, None], xs[2][:, None, None]


def gaussians(case: Case, amps, widths, centres) -> np.ndarray:
    """phi = sum_m A_m exp(-|x - c_m|^2 / w_m^2) at cell centres (plain distance)."""
    X, Y, Z = cell_centres(case)
    phi = np.zeros((case.n[2], case.n[1], case.n[0]), dtype=np.float64)
    for A, w, c in zip(amps, widths, centres):
        r2 = (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2
        phi += A * np.exp(-r2 / (w * w))
    return phi


def fourier_field(case: Case, rng: SplitMix64, modes: int = 16, kmax: int = 8, sd: float = 0.01) -> np.ndarray:
    """sum_m a_m cos(2 pi k_m . x + psi_m), integer k with 0 < |k| <= kmax, a ~ N(0, sd)."""
    X, Y, Z = cell_centres(case)
    f = np.zeros((case.n[2], case.n[1], case.n[0]), dtype=np.float64)
    for _ in range(modes):
        while True:
            k = [int(rng.next_u64() % (2 * kmax + 1)) - kmax for _ in range(3)]
            kk = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
            if 0 < kk <= kmax * kmax:
                break
        a = rng.normal(0.0, sd)
        psi = rng.uniform(0.0, 2.0 * math.pi)
        f += a * np.cos(2.0 * math.pi * (k[0] * X + k[1] * Y + k[2] * Z) + psi)
    return f


# ---------------------------------------------------------------------------
# the workloads
# ---------------------------------------------------------------------------
def wave64(seed: int = 1, n: int = 64):
    """(phi, Pi) wave, periodic unit cube, n=2, KO r=3 sigma=0.1, dt=dx/4; 100 steps."""
    case = Case(system="phipi", n=(n, n, n), stencil_n=2, ghost=3, ko_order=3, ko_strength=0.1,
                periodic=(1, 1, 1), bc="noop", dt=0.25 * (1.0 / n))
    rng = SplitMix64(_seed("wave64", seed))
    c = [rng.uniform(0.3, 0.7) for _ in range(3)]
    state = np.zeros(case.shape)
    state[0] = gaussians(case, [1.0], [0.1], [c])
    return case, state, 100


def wave256(seed: int = 1, n: int = 256, abl: int = 16):
    """(phi, Pi) wave, Dirichlet-0 box with quadratic ABL, n=4, KO r=4 sigma=0.2; 1000 steps."""
    case = Case(system="phipi", n=(n, n, n), stencil_n=4, ghost=4, ko_order=4, ko_strength=0.2,
                periodic=(0, 0, 0), bc="dirichlet", abl_thickness=abl, abl_max=2.0, abl_profile="quadratic",
                dt=0.25 * (1.0 / n))
    rng = SplitMix64(_seed("wave256", seed))
    amps, widths, centres = [], [], []
    for _ in range(8):
        amps.append(rng.uniform(-1.0, 1.0))
        widths.append(rng.uniform(0.05, 0.15))
        centres.append([rng.uniform(0.3, 0.7) for _ in range(3)])
    state = np.zeros(case.shape)
    state[0] = gaussians(case, amps, widths, centres)
    return case, state, 1000


def tensor192(seed: int = 1, n: int = 192, abl: int = 12):
    """Linearised ADM (h_ij, K_ij), 12 comps, n^3, 6th-order stencils (n=3), KO order 3 sigma=0.1,
    12-cell quadratic absorbing layer (max 2.0, Dirichlet 0), dt=dx/4; every component a sum of 16
    Fourier modes (integer |k| <= 8, amplitudes N(0, 0.01)); 500 steps."""
    case = Case(system="tensor", n=(n, n, n), stencil_n=3, ghost=3, ko_order=3, ko_strength=0.1,
                periodic=(0, 0, 0), bc="dirichlet", abl_thickness=abl, abl_max=2.0, abl_profile="quadratic",
                dt=0.25 * (1.0 / n))
    rng = SplitMix64(_seed("tensor192", seed))
    state = np.empty(case.shape)
    for c in range(12):
        state[c] = fourier_field(case, rng, modes=16, kmax=8, sd=0.01)
    return case, state, 500


def fourier_state_device(case: Case, seed_tag: str, seed: int = 1, modes: int = 16, kmax: int = 8,
                         sd: float = 0.01, device: str = "cuda"):
    """The tensor workload's initial state built on the device with torch: the same SplitMix64 draws
    as `fourier_field`, evaluated with torch's cos (last-bit rounding may differ from numpy)."""
    import torch

    rng = SplitMix64(_seed(seed_tag, seed))
    dx = case.dx()
    xs = [case.prob_lo[a] + (torch.arange(case.n[a], dtype=torch.float64, device=device) + 0.5) * dx[a]
          for a in range(3)]
    X, Y, Z = xs[0].view(1, 1, -1), xs[1].view(1, -1, 1), xs[2].view(-1, 1, 1)
    st = torch.zeros(case.shape, dtype=torch.float64, device=device)
    for c in range(case.ncomp):
        for _ in range(modes):
            while True:
                k = [int(rng.next_u64() % (2 * kmax + 1)) - kmax for _ in range(3)]
                kk = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
                if 0 < kk <= kmax * kmax:
                    break
            amp = rng.normal(0.0, sd)
            psi = rng.uniform(0.0, 2.0 * math.pi)
            st[c] += amp * torch.cos(2.0 * math.pi * (k[0] * X + k[1] * Y + k[2] * Z) + psi)
    return st


def ensemble_params(seed: int = 1, count: int = 512, sides=(32, 48, 64, 96, 128)):
    """Cases and Gaussian centres of the ensemble: periodic (phi, Pi) instances with seeded side,
    stencil width n (KO order n, ghost width n), KO strength U[0.05, 0.3] and dt = dx/4."""
    rng = SplitMix64(_seed("ensemble", seed))
    out = []
    for _ in range(count):
        n = rng.choice(list(sides))
        sn = rng.choice([2, 3, 4])
        sigma = rng.uniform(0.05, 0.3)
        case = Case(system="phipi", n=(n, n, n), stencil_n=sn, ghost=sn, ko_order=sn, ko_strength=sigma,
                    periodic=(1, 1, 1), bc="noop", dt=0.25 * (1.0 / n))
        c = [rng.uniform(0.3, 0.7) for _ in range(3)]
        out.append((case, c))
    return out


def ensemble(seed: int = 1, count: int = 512, sides=(32, 48, 64, 96, 128)):
    """Ensemble instances with host initial data (phi = one Gaussian of width 0.1, Pi = 0); 200 steps each."""
    out = []
    for case, c in ensemble_params(seed, count, sides):
        state = np.zeros(case.shape)
        state[0] = gaussians(case, [1.0], [0.1], [c])
        out.append((case, state))
    return out, 200


def gaussian_state_device(case: Case, centre, width: float = 0.1, device: str = "cuda"):
    """The ensemble's initial state built on the device with torch (same formula as `gaussians`, whose
    last-bit rounding of exp may differ); used where host generation of ~1e8 cells would dominate."""
    import torch

    dx = case.dx()
    xs = [case.prob_lo[a] + (torch.arange(case.n[a], dtype=torch.float64, device=device) + 0.5) * dx[a]
          for a in range(3)]
    X = xs[0].view(1, 1, -1)
    Y = xs[1].view(1, -1, 1)
    Z = xs[2].view(-1, 1, 1)
    r2 = (X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2
    st = torch.zeros(case.shape, dtype=torch.float64, device=device)
    st[0] = torch.exp(-r2 / (width * width))
    return st


def small_case(system: str = "phipi", n=(24, 20, 16), stencil_n: int = 2, ko_order: int = 0,
               periodic=(1, 1, 1), abl: int = 0, seed: int = 3, gamma: float = 0.0):
    """Small smooth random case of any registered system (parity tests)."""
    ghost = max(stencil_n, ko_order)
    case = Case(system=system, n=tuple(n), stencil_n=stencil_n, ghost=ghost, ko_order=ko_order,
                ko_strength=0.15 if ko_order else 0.0, periodic=tuple(periodic),
                bc="noop" if all(periodic) else "dirichlet", abl_thickness=abl, abl_max=2.0 if abl else 0.0,
                abl_profile="quadratic", gamma=gamma, dt=0.25 * (1.0 / max(n)))
    rng = SplitMix64(_seed("small", seed))
    if not all(periodic):
        tgt = [rng.uniform(-0.1, 0.1) for _ in range(case.ncomp)]
        case = case.replace(target=tuple(tgt) + (0.0,) * (_lib.TSP_MAX_COMP - case.ncomp))
    state = np.zeros(case.shape)
    for c in range(case.ncomp):
        state[c] = fourier_field(case, rng, modes=6, kmax=3, sd=0.3)
    return case, state


def adjoint128(seed: int = 1, n: int = 128, batch: int = 16, nrec: int = 64, nsteps: int = 500, abl: int = 12):
    """Differentiable solve: (phi, Pi) wave n^3 (n=2 stencils, no KO), 12-cell quadratic absorbing layer
    (max 2.0, Dirichlet 0), dt = dx/4; `batch` sources (one Gaussian of width 0.1 each, centre
    U[0.3,0.7]^3), `nrec` receivers at seeded bulk cells, target traces N(0, 0.01) (numpy PCG64 stream).
    Returns (case, phi0 [batch, n, n, n], receivers [nrec, 3], target [batch, nsteps, nrec], nsteps)."""
    case = Case(system="phipi", n=(n, n, n), stencil_n=2, ghost=2, ko_order=0, ko_strength=0.0,
                periodic=(0, 0, 0), bc="dirichlet", abl_thickness=abl, abl_max=2.0, abl_profile="quadratic",
                dt=0.25 * (1.0 / n))
    rng = SplitMix64(_seed("adjoint128", seed))
    phi0 = np.empty((batch, n, n, n))
    for b in range(batch):
        c = [rng.uniform(0.3, 0.7) for _ in range(3)]
        phi0[b] = gaussians(case, [1.0], [0.1], [c])
    lo, hi = abl, n - 1 - abl
    receivers = np.array([[lo + rng.next_u64() % (hi - lo + 1) for _ in range(3)] for _ in range(nrec)],
                         dtype=np.int32)
    target = np.random.Generator(np.random.PCG64(_seed("adjoint128-target", seed))).normal(
        0.0, 0.01, size=(batch, nsteps, nrec))
    return case, phi0, receivers, target, nsteps
