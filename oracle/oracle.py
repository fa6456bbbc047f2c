"""ctypes front end of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point as the
checker and by bench.py's ``cpu_baseline`` leg. Product code never imports it.

The reference (TreESaP) cannot be built here: it is header-only C++ on top of
AMReX, which its CMake fetches from GitHub (CMakeLists.txt:19-45) and which is
absent from this image, so there is no ``oracle/_ref`` build. oracle.c
restates the reference arithmetic in C, citing file:line per function, and is
pinned by the checks in tests/test_oracle.py:
  - stencil weights against textbook centred-difference tables (Fornberg) and
    polynomial exactness of every derivative operator;
  - Kreiss-Oliger coefficients against the binomial closed form and the
    Fourier symbol -sin^{2r}(theta/2)/dx;
  - RK4 / Euler / KO / every registered RHS on periodic grids against the
    single-Fourier-mode closed form G(k)^N (oracle/fourier.py);
  - ABL sigma and the ghost fill against brute force on tiny grids;
  - the ABL damping term against the RK4 polynomial of the damped affine ODE
    (constant fields where sigma is constant over a step's reach);
  - bitwise invariances (thread count).
Parts the Fourier and brute-force checks cannot reach -- AMReX internals such
as FillPatch ordering on multi-box levels -- stay "parity unpinned".
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
MAX_COMP = 12
SYSTEMS = {"phipi": 0, "phiu": 1, "nonlinear": 2, "tensor": 3}
NCOMP = {"phipi": 2, "phiu": 4, "nonlinear": 2, "tensor": 12}
BCS = {"noop": 0, "dirichlet": 1}
PROFILES = {"linear": 0, "quadratic": 1, "cubic": 2}


class OrDesc(C.Structure):
    _fields_ = [
        ("system", C.c_int),
        ("n", C.c_int * 3),
        ("stencil_n", C.c_int),
        ("ghost", C.c_int),
        ("ko_order", C.c_int),
        ("ko_strength", C.c_double),
        ("prob_lo", C.c_double * 3),
        ("prob_hi", C.c_double * 3),
        ("periodic", C.c_int * 3),
        ("bc", C.c_int),
        ("abl_thickness", C.c_int),
        ("abl_max", C.c_double),
        ("abl_profile", C.c_int),
        ("target", C.c_double * MAX_COMP),
        ("gamma", C.c_double),
        ("dt", C.c_double),
    ]


_lib = None


def build() -> Path:
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = C.CDLL(str(LIB))
        VP = C.c_void_p
        DP = C.POINTER(C.c_double)
        h.or_create.restype = VP
        h.or_create.argtypes = [C.POINTER(OrDesc)]
        h.or_destroy.argtypes = [VP]
        for name in ("or_set_state", "or_get_state"):
            getattr(h, name).argtypes = [VP, DP]
        h.or_get_register.argtypes = [VP, C.c_int, DP]
        h.or_set_register.argtypes = [VP, C.c_int, DP]
        h.or_fill_ghosts.argtypes = [VP, C.c_int]
        h.or_eval_rhs.argtypes = [VP, C.c_int, DP]
        h.or_rk4_steps.argtypes = [VP, C.c_int]
        h.or_euler_steps.argtypes = [VP, C.c_int]
        h.or_time.restype = C.c_double
        h.or_time.argtypes = [VP]
        h.or_steps.restype = C.c_long
        h.or_steps.argtypes = [VP]
        h.or_point_derivative.restype = C.c_double
        h.or_point_derivative.argtypes = [VP] + [C.c_int] * 8
        h.or_abl_sigma.restype = C.c_double
        h.or_abl_sigma.argtypes = [VP, C.c_int, C.c_int, C.c_int]
        for name in ("or_w1", "or_w2", "or_w3"):
            getattr(h, name).restype = C.c_double
            getattr(h, name).argtypes = [C.c_int, C.c_int]
        h.or_harmonic.restype = C.c_double
        h.or_harmonic.argtypes = [C.c_int, C.c_int]
        h.or_binom.restype = C.c_int64
        h.or_binom.argtypes = [C.c_int, C.c_int]
        h.or_ko_coef.restype = C.c_int64
        h.or_ko_coef.argtypes = [C.c_int, C.c_int]
        h.or_adjoint_step.restype = C.c_int
        h.or_adjoint_step.argtypes = [VP, DP]
        h.or_adjoint_step_nl.restype = C.c_int
        h.or_adjoint_step_nl.argtypes = [VP, DP, DP]
        h.or_set_threads.argtypes = [C.c_int]
        h.or_max_threads.restype = C.c_int
        h.or_system_ncomp.restype = C.c_int
        h.or_system_ncomp.argtypes = [C.c_int]
        _lib = h
    return _lib


def _get(case, key):
    return case[key] if isinstance(case, dict) else getattr(case, key)


def make_desc(case) -> OrDesc:
    d = OrDesc()
    d.system = SYSTEMS[_get(case, "system")]
    for a in range(3):
        d.n[a] = int(_get(case, "n")[a])
        d.prob_lo[a] = float(_get(case, "prob_lo")[a])
        d.prob_hi[a] = float(_get(case, "prob_hi")[a])
        d.periodic[a] = int(_get(case, "periodic")[a])
    d.stencil_n = int(_get(case, "stencil_n"))
    d.ghost = int(_get(case, "ghost"))
    d.ko_order = int(_get(case, "ko_order"))
    d.ko_strength = float(_get(case, "ko_strength"))
    d.bc = BCS[_get(case, "bc")]
    d.abl_thickness = int(_get(case, "abl_thickness"))
    d.abl_max = float(_get(case, "abl_max"))
    d.abl_profile = PROFILES[_get(case, "abl_profile")]
    tgt = list(_get(case, "target")) + [0.0] * MAX_COMP
    for c in range(MAX_COMP):
        d.target[c] = float(tgt[c])
    d.gamma = float(_get(case, "gamma"))
    d.dt = float(_get(case, "dt"))
    return d


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleSolver:
    """CPU restatement of SingleLevelRK4Integrator / SingleLevelEulerIntegrator on one box."""

    def __init__(self, case):
        self.case = case
        self._desc = make_desc(case)
        self.ncomp = NCOMP[_get(case, "system")]
        n = _get(case, "n")
        self.shape = (self.ncomp, int(n[2]), int(n[1]), int(n[0]))
        G = int(_get(case, "ghost"))
        self.gshape = (self.ncomp, int(n[2]) + 2 * G, int(n[1]) + 2 * G, int(n[0]) + 2 * G)
        self._h = lib().or_create(C.byref(self._desc))
        if not self._h:
            raise ValueError("oracle rejected the case description")

    def close(self):
        if self._h:
            lib().or_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, state: np.ndarray):
        a = np.ascontiguousarray(state, dtype=np.float64)
        assert a.shape == self.shape, (a.shape, self.shape)
        lib().or_set_state(self._h, _dp(a))

    def get_state(self) -> np.ndarray:
        a = np.empty(self.shape)
        lib().or_get_state(self._h, _dp(a))
        return a

    def get_register(self, role: int) -> np.ndarray:
        a = np.empty(self.gshape)
        lib().or_get_register(self._h, role, _dp(a))
        return a

    def set_register(self, role: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.gshape
        lib().or_set_register(self._h, role, _dp(a))

    def fill_ghosts(self, role: int = 0):
        lib().or_fill_ghosts(self._h, role)

    def eval_rhs(self, role: int = 0) -> np.ndarray:
        out = np.empty(self.shape)
        lib().or_eval_rhs(self._h, role, _dp(out))
        return out

    def rk4_steps(self, n: int):
        lib().or_rk4_steps(self._h, int(n))

    def euler_steps(self, n: int):
        lib().or_euler_steps(self._h, int(n))

    def adjoint_step(self, lam: np.ndarray) -> np.ndarray:
        """G^T lam for one RK4 step of the linear (phi, Pi) system (registers are used as scratch)."""
        a = np.array(lam, dtype=np.float64, copy=True, order="C")
        assert a.shape == self.shape
        if lib().or_adjoint_step(self._h, _dp(a)) != 0:
            raise ValueError("adjoint step needs the (phi, Pi) system")
        return a

    def adjoint_step_nl(self, lam: np.ndarray, u: np.ndarray) -> np.ndarray:
        """G'(u)^T lam for one RK4 step of the NonLinear (phi, pi) system linearised at the state u
        (registers are used as scratch)."""
        a = np.array(lam, dtype=np.float64, copy=True, order="C")
        uu = np.ascontiguousarray(u, dtype=np.float64)
        assert a.shape == self.shape and uu.shape == self.shape
        if lib().or_adjoint_step_nl(self._h, _dp(a), _dp(uu)) != 0:
            raise ValueError("nonlinear adjoint step needs the NonLinear system")
        return a

    def point_derivative(self, dirs, i, j, k, c=0):
        d = list(dirs) + [0, 0, 0]
        return lib().or_point_derivative(self._h, len(dirs), d[0], d[1], d[2], i, j, k, c)

    def point_ko(self, order, direction, i, j, k, c=0):
        return lib().or_point_derivative(self._h, 0, order, direction, 0, i, j, k, c)

    def abl_sigma(self, i, j, k):
        return lib().or_abl_sigma(self._h, i, j, k)

    @property
    def time(self):
        return lib().or_time(self._h)

    @property
    def steps(self):
        return lib().or_steps(self._h)


def weights(n: int):
    """(w1[0..n], w2[0..n], w3[0..n]) as the reference forms them."""
    L = lib()
    w1 = [L.or_w1(n, p) for p in range(n + 1)]
    w2 = [L.or_w2(n, p) for p in range(n + 1)]
    w3 = [L.or_w3(n, p) if n > 1 else float("nan") for p in range(n + 1)]
    return w1, w2, w3


def ko_coefs(r: int):
    return [lib().or_ko_coef(r, s) for s in range(r + 1)]


def set_threads(n: int):
    """OpenMP threads of the stepping loops from now on."""
    os.environ["OMP_NUM_THREADS"] = str(n)
    lib().or_set_threads(int(n))


def max_threads() -> int:
    return lib().or_max_threads()
