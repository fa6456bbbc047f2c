"""Discrete adjoint on the GPU: bitwise against the oracle's transposed RK4
step, the dot-product identity, tangent-linear and finite-difference
directional derivatives, checkpoint recomputation, and the torch.autograd
binding."""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from treesap_amd import Solver
from treesap_amd.adjoint import Adjoint, ReceiverProblem, receiver_loss
from treesap_amd.workloads import adjoint128, small_case

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


ADJ_CASES = [
    ((24, 20, 16), 2, 0, (0, 0, 0), 4),
    ((70, 21, 19), 4, 4, (0, 0, 0), 5),
    ((33, 18, 22), 3, 3, (1, 0, 1), 0),
    ((40, 24, 20), 2, 3, (1, 1, 1), 0),
]


def _case(n, sn, ko, per, abl, seed=5):
    case, _ = small_case("phipi", n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl, seed=seed)
    return case.replace(target=(0.0,) * 12)  # linear (homogeneous) boundary values


@pytest.mark.parametrize("n,sn,ko,per,abl", ADJ_CASES)
def test_transposed_step_matches_oracle_and_dot_product(n, sn, ko, per, abl):
    _say(f"[adjoint] apply n={n} stencil={sn} ko={ko} periodic={per} abl={abl}")
    case = _case(n, sn, ko, per, abl)
    s = Solver(case)
    a = Adjoint(s, [[1, 1, 1]], 1, 1)
    rng = np.random.default_rng(3)
    o = orc.OracleSolver(case)
    for _ in range(2):
        u = rng.standard_normal(case.shape)
        v = rng.standard_normal(case.shape)
        gtv = a.apply(v)
        assert np.array_equal(gtv, o.adjoint_step(v))
        s.set_state(u)
        s.rk4_steps(1)
        gu = s.get_state()
        lhs, rhs = float(np.sum(gu * v)), float(np.sum(u * gtv))
        assert lhs == pytest.approx(rhs, rel=1e-12)


@pytest.mark.parametrize("n,sn,ko,per,abl", [ADJ_CASES[1], ADJ_CASES[2]])
def test_pointer_access_transposed_step_matches_oracle(monkeypatch, n, sn, ko, per, abl):
    """The transposed stage through the kernel that registers of 4 GiB or
    more use (64-bit pointers), forced on small boxes."""
    _say(f"[adjoint] pointer access n={n} stencil={sn} ko={ko}")
    monkeypatch.setenv("TSP_POINTER_ACCESS", "1")
    case = _case(n, sn, ko, per, abl)
    s = Solver(case)
    a = Adjoint(s, [[1, 1, 1]], 1, 1)
    v = np.random.default_rng(4).standard_normal(case.shape)
    assert np.array_equal(a.apply(v), orc.OracleSolver(case).adjoint_step(v))


def _problem(n=(36, 30, 28), sn=2, ko=0, per=(0, 0, 0), abl=4, nsteps=9, K=4, nrec=7, seed=2):
    case = _case(n, sn, ko, per, abl, seed=seed)
    rng = np.random.default_rng(seed)
    rec = np.stack([rng.integers(0, n[a], size=nrec) for a in range(3)], axis=1).astype(np.int32)
    rec[1] = rec[0]  # two receivers on one cell
    _, st = small_case("phipi", n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl, seed=seed + 1)
    st[1] = 0.0
    target = rng.normal(0.0, 0.01, size=(nsteps, nrec))
    return case, st, rec, target, nsteps, K


def _oracle_trace_and_grad(case, st, rec, target, nsteps):
    o = orc.OracleSolver(case)
    o.set_state(st)
    trace = np.empty((nsteps, len(rec)))
    for n in range(nsteps):
        o.rk4_steps(1)
        phi = o.get_state()[0]
        trace[n] = [phi[k, j, i] for i, j, k in rec]
    dtrace = 2.0 * (trace - target)
    lam = np.zeros(o.shape)
    adj = orc.OracleSolver(case)
    for n in range(nsteps, 0, -1):
        for r, (i, j, k) in enumerate(rec):
            lam[0, k, j, i] += dtrace[n - 1, r]
        lam = adj.adjoint_step(lam)
    return trace, dtrace, lam


@pytest.mark.parametrize("ko,per,abl", [(0, (0, 0, 0), 4), (3, (1, 1, 1), 0), (2, (1, 0, 1), 3)])
def test_backward_matches_oracle_reverse_sweep(ko, per, abl):
    _say(f"[adjoint] backward ko={ko} periodic={per} abl={abl}")
    case, st, rec, target, nsteps, K = _problem(ko=ko, per=per, abl=abl)
    s = Solver(case)
    s.set_state(st)
    a = Adjoint(s, rec, nsteps, K)
    a.forward()
    trace = a.trace()
    otrace, odtrace, ograd = _oracle_trace_and_grad(case, st, rec, target, nsteps)
    assert np.array_equal(trace, otrace)
    dtrace = 2.0 * (trace - target)
    g = a.backward(dtrace)
    assert np.array_equal(g, ograd)
    # checkpoints are the forward states
    assert np.array_equal(a.checkpoint(0), st)
    o = orc.OracleSolver(case)
    o.set_state(st)
    o.rk4_steps(K)
    assert np.array_equal(a.checkpoint(1), o.get_state())
    assert np.array_equal(a.checkpoint(a.nslots - 1), s.get_state())


def test_recompute_mode_is_bitwise_equal_and_restores_the_solver():
    _say("[adjoint] recompute")
    case, st, rec, target, nsteps, K = _problem(nsteps=11, K=3)
    s = Solver(case)
    s.set_state(st)
    a = Adjoint(s, rec, nsteps, K)
    a.forward()
    final = s.get_state()
    t_final, n_final = s.time, s.steps
    dtrace = 2.0 * (a.trace() - target)
    g0 = a.backward(dtrace)
    g1 = a.backward(dtrace, recompute=True)
    assert np.array_equal(g0, g1)
    assert np.array_equal(s.get_state(), final)
    assert s.time == t_final and s.steps == n_final


def _loss(case, st, rec, target, nsteps):
    s = Solver(case)
    s.set_state(st)
    a = Adjoint(s, rec, nsteps, nsteps)
    a.forward()
    tr = a.trace()
    return float(np.sum((tr - target) ** 2)), tr


def test_tangent_linear_and_finite_differences():
    _say("[adjoint] TLM / FD")
    case, st, rec, target, nsteps, K = _problem(n=(40, 34, 30), abl=5, nsteps=15, K=5)
    s = Solver(case)
    s.set_state(st)
    a = Adjoint(s, rec, nsteps, K)
    a.forward()
    tr = a.trace()
    g = a.backward(2.0 * (tr - target))
    rng = np.random.default_rng(9)
    v = np.zeros(case.shape)
    v[0] = rng.standard_normal(case.shape[1:])
    # tangent linear: the system is linear with zero boundary values, so the TLM is the forward solve of v
    _, tr_v = _loss(case, v, rec, np.zeros_like(target), nsteps)
    tlm = float(np.sum(2.0 * (tr - target) * tr_v))
    adj = float(np.sum(g * v))
    assert adj == pytest.approx(tlm, rel=1e-12)
    eps = 1e-6
    lp, _ = _loss(case, st + eps * v, rec, target, nsteps)
    lm, _ = _loss(case, st - eps * v, rec, target, nsteps)
    fd = (lp - lm) / (2 * eps)
    assert adj == pytest.approx(fd, rel=1e-7)


def test_autograd_function_on_a_reduced_adjoint_workload():
    _say("[adjoint] torch.autograd")
    torch = pytest.importorskip("torch")
    case, phi0, rec, target, nsteps = adjoint128(n=40, batch=3, nrec=16, nsteps=30, abl=6)
    prob = ReceiverProblem(case, rec, nsteps, batch=3, checkpoint_every=10)
    x = torch.tensor(phi0, device="cuda", requires_grad=True)
    tgt = torch.tensor(target, device="cuda")
    loss = receiver_loss(x, tgt, prob)
    loss.backward()
    g = x.grad.detach().cpu().numpy()
    # same gradient through the low-level objects, source by source
    for b in range(3):
        st = np.zeros(case.shape)
        st[0] = phi0[b]
        s = Solver(case)
        s.set_state(st)
        a = Adjoint(s, rec, nsteps, 10)
        a.forward()
        gb = a.backward(2.0 * (a.trace() - target[b]))
        assert np.array_equal(gb[0], g[b])
    # gradcheck-style directional derivative against central differences
    v = torch.randn_like(x)
    eps = 1e-6
    with torch.no_grad():
        lp = receiver_loss(x + eps * v, tgt, prob)
        lm = receiver_loss(x - eps * v, tgt, prob)
    fd = float((lp - lm) / (2 * eps))
    assert float((x.grad * v).sum()) == pytest.approx(fd, rel=1e-7)
    prob.close()


@pytest.mark.parametrize("ko,per,abl", [(3, (1, 1, 1), 0), (2, (1, 0, 1), 3)])
def test_backward_with_stage_pairs_matches_oracle(ko, per, abl):
    """Forward sweep with the stage-pair kernels: the adjoint reads only the forward's state_old
    (receiver trace and checkpoints), so trace and gradient stay bitwise, in both backward modes."""
    _say(f"[adjoint] stage pairs ko={ko} periodic={per} abl={abl}")
    case, st, rec, target, nsteps, K = _problem(ko=ko, per=per, abl=abl)
    s = Solver(case)
    s.kernel = dict(ghost_mode=0, stage_pairs=1)
    s.set_state(st)
    a = Adjoint(s, rec, nsteps, K)
    a.forward()
    trace = a.trace()
    otrace, _, ograd = _oracle_trace_and_grad(case, st, rec, target, nsteps)
    assert np.array_equal(trace, otrace)
    dtrace = 2.0 * (trace - target)
    assert np.array_equal(a.backward(dtrace), ograd)
    assert np.array_equal(a.backward(dtrace, recompute=True), ograd)


@pytest.mark.parametrize("recompute", [False, True])
def test_batch_adjoint_equals_single_box_adjoints(recompute):
    """BatchAdjoint over four sources of one model (odd step count, checkpoint interval not
    dividing it, receivers sharing a cell): every instance's trace and gradient equal the
    single-box adjoint of that instance bitwise, stored and recomputed."""
    from treesap_amd import Batch
    from treesap_amd.adjoint import BatchAdjoint

    _say(f"[adjoint] batch adjoint recompute={recompute}")
    case, st0, rec, target, nsteps, K = _problem(n=(36, 30, 28), sn=2, ko=2, abl=4, nsteps=11, K=4)
    states = []
    for k in range(4):
        _, st = small_case("phipi", n=case.n, stencil_n=2, ko_order=2, periodic=case.periodic, abl=4, seed=40 + k)
        st[1] = 0.0
        states.append(st)
    b = Batch([case] * 4)
    for i, st in enumerate(states):
        b.set_state(i, st)
    ba = BatchAdjoint(b, rec, nsteps, K)
    ba.forward()
    tr = ba.trace()
    dtr = 2.0 * (tr - target[None])
    g = ba.backward(dtr, recompute=recompute).reshape(4, *case.shape)
    for i, st in enumerate(states):
        s = Solver(case)
        s.set_state(st)
        a = Adjoint(s, rec, nsteps, K)
        a.forward()
        assert np.array_equal(a.trace(), tr[i]), i
        gi = a.backward(dtr[i], recompute=recompute)
        assert np.array_equal(gi, g[i]), i
        a.close()
        s.close()
    # the batch is left at the end of the forward pass
    assert b.steps == nsteps
    ba.close()
    b.close()


def test_batch_adjoint_follows_a_changed_batch_configuration():
    """The batched adjoint keeps its own instance table (the adjoint registers), with each
    instance's z chunk: after the batch's kernel configuration changes from 5-plane z chunks
    to one chunk per box, a second forward and backward pass from the same states gives the
    same trace and gradient, bitwise."""
    from treesap_amd import Batch
    from treesap_amd.adjoint import BatchAdjoint

    _say("[adjoint] batch adjoint after a configuration change")
    case, _, rec, target, nsteps, K = _problem(n=(36, 30, 28), sn=2, ko=2, abl=4, nsteps=7, K=3)
    states = []
    for k in range(2):
        _, st = small_case("phipi", n=case.n, stencil_n=2, ko_order=2, periodic=case.periodic, abl=4, seed=60 + k)
        st[1] = 0.0
        states.append(st)
    b = Batch([case] * 2)
    b.kernel = {"tile_y": 8, "zchunk": 5, "grid_blocks": 0, "zalt": 1, "halo_ring": 1, "narrow_tiles": 1}
    ba = BatchAdjoint(b, rec, nsteps, K)
    runs = []
    for cfg in (None, {"tile_y": 16, "zchunk": 0, "grid_blocks": 9, "zalt": 2, "halo_ring": 0, "narrow_tiles": 0}):
        if cfg is not None:
            b.set_group_kernel(2, 2, cfg)
        for i, st in enumerate(states):
            b.set_state(i, st)
        ba.forward()
        tr = ba.trace()
        runs.append((tr, ba.backward(2.0 * (tr - target[None]), recompute=True)))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    ba.close()
    b.close()


def test_adjoint_rejects_bad_tensors():
    """Device tensors handed to the adjoint must be contiguous float64 of the exact shape."""
    torch = pytest.importorskip("torch")
    case, st, rec, target, nsteps, K = _problem()
    s = Solver(case)
    s.set_state(st)
    a = Adjoint(s, rec, nsteps, K)
    a.forward()
    good = torch.zeros((nsteps, len(rec)), dtype=torch.float64, device="cuda")
    for bad in (torch.zeros((nsteps, len(rec)), dtype=torch.float32, device="cuda"),
                torch.zeros((nsteps + 1, len(rec)), dtype=torch.float64, device="cuda"),
                torch.zeros((len(rec), nsteps), dtype=torch.float64, device="cuda").t()):
        with pytest.raises(ValueError):
            a.trace(out=bad)
    with pytest.raises(ValueError):
        a.backward(good, out=torch.zeros(case.shape, dtype=torch.float32, device="cuda"))
    a.trace(out=good)
    a.close()
    s.close()


# ---- NonLinear (phi, pi) system: G'(u)^T, linearised at the forward states ----

NL_CASES = [  # n, stencil, KO order, periodic, ABL, gamma
    ((26, 20, 18), 2, 2, (1, 1, 1), 0, 0.3),
    ((30, 22, 20), 3, 3, (0, 0, 0), 3, 0.5),
    ((21, 17, 19), 1, 0, (1, 0, 1), 0, 0.2),
    ((70, 21, 19), 4, 4, (0, 0, 0), 4, 0.4),
]


def _nl_case(n, sn, ko, per, abl, gamma, seed=8):
    case, st = small_case("nonlinear", n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl, seed=seed, gamma=gamma)
    return case, st


@pytest.mark.parametrize("n,sn,ko,per,abl,gamma", NL_CASES)
def test_nonlinear_transposed_step_matches_oracle(n, sn, ko, per, abl, gamma):
    """tsp_adjoint_apply on the NonLinear system is the transpose of the step's Jacobian at the
    solver's state, bitwise against the oracle's (itself pinned against a numpy tangent-linear
    model), for two states and two adjoint vectors."""
    _say(f"[adjoint] nonlinear apply n={n} stencil={sn} ko={ko} periodic={per} abl={abl}")
    case, st = _nl_case(n, sn, ko, per, abl, gamma)
    s = Solver(case)
    a = Adjoint(s, [[1, 1, 1]], 1, 1)
    rng = np.random.default_rng(9)
    o = orc.OracleSolver(case)
    for q in range(2):
        u = st if q == 0 else st + 0.3 * rng.standard_normal(case.shape)
        s.set_state(u)
        lam = rng.standard_normal(case.shape)
        got = a.apply(lam)
        assert np.array_equal(got, o.adjoint_step_nl(lam, u)), q
        assert np.array_equal(s.get_state(), u)  # the forward state is untouched


def _nl_problem(n=(30, 26, 22), sn=2, ko=2, per=(0, 0, 0), abl=4, gamma=0.4, nsteps=8, K=3, nrec=6, seed=12):
    case, st = _nl_case(n, sn, ko, per, abl, gamma, seed=seed)
    rng = np.random.default_rng(seed)
    rec = np.stack([rng.integers(0, n[a], size=nrec) for a in range(3)], axis=1).astype(np.int32)
    rec[1] = rec[0]
    target = rng.normal(0.0, 0.01, size=(nsteps, nrec))
    return case, st, rec, target, nsteps, K


def _oracle_trace_and_grad_nl(case, st, rec, target, nsteps):
    o = orc.OracleSolver(case)
    o.set_state(st)
    states, trace = [], np.empty((nsteps, len(rec)))
    for n in range(nsteps):
        states.append(o.get_state())
        o.rk4_steps(1)
        phi = o.get_state()[0]
        trace[n] = [phi[k, j, i] for i, j, k in rec]
    dtrace = 2.0 * (trace - target)
    lam = np.zeros(o.shape)
    adj = orc.OracleSolver(case)
    for n in range(nsteps, 0, -1):
        for r, (i, j, k) in enumerate(rec):
            lam[0, k, j, i] += dtrace[n - 1, r]
        lam = adj.adjoint_step_nl(lam, states[n - 1])
    return trace, dtrace, lam


@pytest.mark.parametrize("ko,per,abl,gamma", [(2, (0, 0, 0), 4, 0.4), (3, (1, 1, 1), 0, 0.6), (0, (1, 0, 1), 3, 0.3)])
def test_nonlinear_backward_matches_oracle_and_finite_differences(ko, per, abl, gamma):
    """The NonLinear backward pass (checkpoints every 3 of 8 steps, each segment's states
    recomputed and every reverse step linearised at its own state) equals the oracle's reverse
    sweep bitwise, leaves the solver at the end of the forward pass, and its gradient matches
    centred finite differences of the loss (eps = 1e-6) in a random direction."""
    _say(f"[adjoint] nonlinear backward ko={ko} periodic={per} abl={abl} gamma={gamma}")
    case, st, rec, target, nsteps, K = _nl_problem(ko=ko, per=per, abl=abl, gamma=gamma)
    s = Solver(case)
    s.set_state(st)
    a = Adjoint(s, rec, nsteps, K)
    a.forward()
    final, t_final = s.get_state(), s.time
    trace = a.trace()
    otrace, odtrace, ograd = _oracle_trace_and_grad_nl(case, st, rec, target, nsteps)
    assert np.array_equal(trace, otrace)
    g = a.backward(odtrace)
    assert np.array_equal(g, ograd)
    assert np.array_equal(a.backward(odtrace, recompute=True), ograd)
    assert np.array_equal(s.get_state(), final) and s.time == t_final

    def loss(u):
        t = Solver(case)
        t.set_state(u)
        tr = np.empty((nsteps, len(rec)))
        for n in range(nsteps):
            t.rk4_steps(1)
            phi = t.get_state()[0]
            tr[n] = [phi[k, j, i] for i, j, k in rec]
        return float(np.sum((tr - target) ** 2))

    v = np.random.default_rng(21).standard_normal(case.shape)
    eps = 1e-6
    fd = (loss(st + eps * v) - loss(st - eps * v)) / (2 * eps)
    assert float(np.sum(g * v)) == pytest.approx(fd, rel=1e-6)
