"""The benchmark workloads checked at their stated sizes (or, where the oracle
would take too long, at their stated parameters with size-independent
properties at full size):

  wave256 parameters (n=4, KO 4 sigma=0.2, 16-cell ABL, Dirichlet) on a 64^3
      box for the full 1000 RK4 steps, bitwise against the oracle;
  tensor192 (12 components, mixed derivatives) at 192^3: one RK4 step
      bitwise against the oracle, and linearity and mirror symmetry over 20
      graph-replayed steps;
  the 512-instance ensemble: every instance of the batched solve bitwise
      equal to its plain single-box solve after three steps;
  adjoint128 with one source for its full 500 steps and checkpoints every 50:
      the gradient against the tangent-linear directional derivative at 1e-12,
      and the checkpoint-recompute sweep bitwise equal to the stored sweep.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from treesap_amd import Batch, Solver
from treesap_amd.adjoint import Adjoint
from treesap_amd.workloads import adjoint128, ensemble_params, fourier_state_device, gaussian_state_device, wave256

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_wave256_parameters_1000_steps_match_oracle():
    """The wave256 model (n=4, KO order 4 sigma=0.2, quadratic ABL max 2.0, Dirichlet 0,
    dt = dx/4, eight Gaussians) on 64^3 for the workload's 1000 RK4 steps: bitwise."""
    _say("[fullsize] wave256 parameters, 64^3 x 1000 steps")
    case, state, nsteps = wave256(n=64)
    assert nsteps == 1000 and case.abl_thickness == 16
    s = Solver(case)
    s.set_state(state)
    s.graph_rk4_steps(nsteps)
    got = s.get_state()
    t = s.time
    s.close()
    o = orc.OracleSolver(case)
    o.set_state(state)
    o.rk4_steps(nsteps)
    want = o.get_state()
    assert np.isfinite(got).all() and np.abs(got).max() > 1e-6
    assert np.array_equal(got, want), f"max abs diff {np.max(np.abs(got - want))}"
    assert t == o.time


def _tensor_case_and_state():
    """The tensor192 case with its initial state built on the device (the workload's Fourier modes)
    and copied to the host, so the GPU solve and the oracle start from the same bytes."""
    from treesap_amd.workloads import Case

    case = Case(system="tensor", n=(192, 192, 192), stencil_n=3, ghost=3, ko_order=3, ko_strength=0.1,
                periodic=(0, 0, 0), bc="dirichlet", abl_thickness=12, abl_max=2.0, abl_profile="quadratic",
                dt=0.25 * (1.0 / 192))
    return case, fourier_state_device(case, "tensor192").cpu().numpy()


def test_tensor192_one_step_matches_oracle():
    """tensor192 at full size, one RK4 step, bitwise against the oracle."""
    _say("[fullsize] tensor192, 1 step")
    case, state = _tensor_case_and_state()
    assert case.n == (192, 192, 192) and case.ncomp == 12
    s = Solver(case)
    s.set_state(state)
    s.rk4_steps(1)
    got = s.get_state()
    s.close()
    o = orc.OracleSolver(case)
    o.set_state(state)
    o.rk4_steps(1)
    want = o.get_state()
    assert np.array_equal(got, want), f"max abs diff {np.max(np.abs(got - want))}"


# Sym2 components xx, yx, yy, zx, zy, zz of h (0-5) and K (6-11): sign under x -> -x
_X_PARITY = np.array([1, -1, 1, -1, 1, 1] * 2, dtype=np.float64)


def test_tensor192_full_size_properties():
    """20 graph-replayed steps at 192^3: exact linearity under scaling by 2 and the
    reflection x -> -x (tensor components with one x index change sign; every
    operator, the absorbing layer and the zero boundary values are reflection
    symmetric, and each stencil term maps to exactly +- its mirror)."""
    _say("[fullsize] tensor192 properties, 20 steps")
    case, state = _tensor_case_and_state()
    mirror = state[:, :, :, ::-1] * _X_PARITY[:, None, None, None]
    sym = state + mirror
    s = Solver(case)
    s.set_state(sym)
    s.graph_rk4_steps(20)
    ra = s.get_state()
    s.set_state(2.0 * sym)
    s.graph_rk4_steps(20)
    rb = s.get_state()
    s.close()
    assert np.isfinite(ra).all() and np.abs(ra).max() > 1e-4
    assert np.array_equal(rb, 2.0 * ra)
    assert np.array_equal(ra, ra[:, :, :, ::-1] * _X_PARITY[:, None, None, None])


def test_full_ensemble_batched_equals_plain():
    """All 512 ensemble instances (sides 32..128, n in {2,3,4}, KO order n): the batched
    graph-replayed solve equals each instance's plain single-box solve, bitwise."""
    torch = pytest.importorskip("torch")
    _say("[fullsize] ensemble 512 instances x 3 steps, batched vs plain")
    params = ensemble_params(count=512)
    b = Batch([c for c, _ in params])
    for i, (case, centre) in enumerate(params):
        b.set_state(i, gaussian_state_device(case, centre))
    b.graph_rk4_steps(3)
    bad = []
    for i, (case, centre) in enumerate(params):
        s = Solver(case)
        s.set_state(gaussian_state_device(case, centre))
        s.rk4_steps(3)
        got = s.get_state(out=torch.empty(case.shape, dtype=torch.float64, device="cuda"))
        want = b.get_state(i, out=torch.empty(case.shape, dtype=torch.float64, device="cuda"))
        if not torch.equal(got, want) or not bool(torch.isfinite(want).all()):
            bad.append(i)
        s.close()
    b.close()
    assert not bad, f"instances differing: {bad[:10]} ({len(bad)} of 512)"


def test_adjoint128_one_source_full_length():
    """adjoint128, one source, 500 steps, checkpoints every 50: the adjoint gradient
    equals the tangent-linear directional derivative (the forward solve of the
    direction, the system being linear with zero boundary values) to 1e-12, and
    the checkpoint-recompute sweep is bitwise the stored sweep."""
    _say("[fullsize] adjoint128, 1 source x 500 steps")
    case, phi0, rec, target, nsteps = adjoint128(batch=1)
    assert nsteps == 500 and case.n == (128, 128, 128)
    st = np.zeros(case.shape)
    st[0] = phi0[0]
    s = Solver(case)
    s.set_state(st)
    a = Adjoint(s, rec, nsteps, 50)
    a.forward()
    tr = a.trace()
    dtrace = 2.0 * (tr - target[0])
    g = a.backward(dtrace)
    g2 = a.backward(dtrace, recompute=True)
    assert np.array_equal(g, g2)
    rng = np.random.default_rng(11)
    v = np.zeros(case.shape)
    v[0] = rng.standard_normal(case.shape[1:])
    sv = Solver(case)
    sv.set_state(v)
    av = Adjoint(sv, rec, nsteps, nsteps)
    av.forward()
    tlm = float(np.sum(dtrace * av.trace()))
    adj = float(np.sum(g * v))
    _say(f"  <grad, v> = {adj:.17g}, TLM = {tlm:.17g}")
    assert adj == pytest.approx(tlm, rel=1e-12)
    a.close()
    av.close()
    s.close()
    sv.close()
