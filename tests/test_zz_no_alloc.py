"""tsp_rk4_steps_noalloc: RK4 steps with the registers in the library's static
device scratch (no allocation in the call), checked through the C ABI against
the allocating path (tsp_set_state + tsp_rk4_steps + tsp_get_state on device
pointers), bitwise, on two calls in a row of which the second has a smaller
input, and its size limit."""
import ctypes as C
import sys

import numpy as np
import pytest

from treesap_amd import Solver, _lib
from treesap_amd.workloads import small_case

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _reference(case, state, nsteps):
    """The existing entry points on device pointers."""
    import torch

    s = Solver(case)
    x = torch.tensor(state, device="cuda")
    L = _lib.lib()
    torch.cuda.synchronize()
    _lib.check(L.tsp_set_state(s.handle, C.c_void_p(x.data_ptr()), 1), "tsp_set_state")
    _lib.check(L.tsp_rk4_steps(s.handle, int(nsteps)), "tsp_rk4_steps")
    y = torch.empty_like(x)
    _lib.check(L.tsp_get_state(s.handle, C.c_void_p(y.data_ptr()), 1), "tsp_get_state")
    out = y.cpu().numpy()
    s.close()
    return out


def _noalloc(s, state, nsteps):
    import torch

    x = torch.tensor(state, device="cuda")
    y = torch.full_like(x, float("nan"))
    torch.cuda.synchronize()
    _lib.check(_lib.lib().tsp_rk4_steps_noalloc(s.handle, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
                                                int(nsteps)), "tsp_rk4_steps_noalloc")
    _lib.check(_lib.lib().tsp_sync(s.handle), "tsp_sync")
    return y.cpu().numpy()


def test_noalloc_matches_allocating_path_twice_shrinking():
    pytest.importorskip("torch")
    big, st_big = small_case("phipi", n=(70, 36, 30), stencil_n=4, ko_order=4, periodic=(0, 0, 0), abl=6, seed=31)
    small, st_small = small_case("phipi", n=(21, 13, 17), stencil_n=2, ko_order=3, periodic=(1, 0, 1), seed=32)
    _say(f"[noalloc] big {big.n}, then small {small.n}")
    sb = Solver(big)
    ss = Solver(small)
    got_b = _noalloc(sb, st_big, 5)
    got_s = _noalloc(ss, st_small, 4)  # the scratch still holds the first call's larger registers
    want_b = _reference(big, st_big, 5)
    want_s = _reference(small, st_small, 4)
    assert np.array_equal(got_b, want_b)
    assert np.array_equal(got_s, want_s)
    # the solvers' own state and clock are untouched
    assert sb.steps == 0 and ss.steps == 0
    sb.close()
    ss.close()


def test_noalloc_with_stage_pairs_and_fill_mode():
    """The same stage kernels as the solver's configuration: stage pairs and the fill-pass mode."""
    pytest.importorskip("torch")
    case, st = small_case("phipi", n=(40, 33, 28), stencil_n=2, ko_order=2, periodic=(1, 1, 0), abl=0, seed=33)
    want = _reference(case, st, 3)
    for cfg in ({"stage_pairs": 1}, {"ghost_mode": 1}):
        _say(f"[noalloc] kernel {cfg}")
        s = Solver(case)
        s.kernel = cfg
        assert np.array_equal(_noalloc(s, st, 3), want)
        s.close()


def test_noalloc_tensor_system():
    """The tensor system's stage (its K components' KO sums go through the solver's own KO
    buffer, allocated with the solver) under the allocation-free entry point, RK4."""
    pytest.importorskip("torch")
    case, st = small_case("tensor", n=(21, 18, 16), stencil_n=2, ko_order=2, periodic=(0, 1, 0), abl=3, seed=35)
    _say(f"[noalloc] tensor {case.n}")
    s = Solver(case)
    got = _noalloc(s, st, 3)
    want = _reference(case, st, 3)
    assert np.isfinite(got).all() and np.array_equal(got, want)
    assert s.steps == 0
    s.close()


def test_noalloc_rejects_too_large_solvers():
    torch = pytest.importorskip("torch")
    case, _ = small_case("phipi", n=(120, 120, 120), stencil_n=2, ko_order=2, seed=34)
    s = Solver(case)
    x = torch.zeros(case.shape, dtype=torch.float64, device="cuda")
    st = _lib.lib().tsp_rk4_steps_noalloc(s.handle, C.c_void_p(x.data_ptr()), C.c_void_p(x.data_ptr()), 1)
    assert st == _lib.INVALID_ARG
    assert "scratch" in _lib.lib().tsp_last_error().decode()
    s.close()
