"""GPU parity: the HIP library against the CPU oracle on identical seeded inputs.

The bar is the reference's: <= 1e-12 relative L2 error per component after the
stated step count. Both sides evaluate the same expressions in the same order
without FMA contraction, so the tests also require bitwise equality where the
arithmetic is identical, which pins the kernels much harder than the tolerance.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from treesap_amd import Solver, _lib
from treesap_amd.workloads import small_case, wave64, wave256

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2_per_comp(got, want):
    out = []
    for c in range(got.shape[0]):
        den = np.linalg.norm(want[c])
        num = np.linalg.norm(got[c] - want[c])
        out.append(num / den if den > 0 else num)
    return np.array(out)


def run_both(case, state, nsteps, scheme="rk4", kernel=None, graph=False):
    s = Solver(case)
    if kernel:
        s.kernel = kernel
    s.set_state(state)
    if scheme == "rk4":
        (s.graph_rk4_steps if graph else s.rk4_steps)(nsteps)
    else:
        s.euler_steps(nsteps)
    got = s.get_state()
    t, n = s.time, s.steps
    s.close()
    o = orc.OracleSolver(case)
    o.set_state(state)
    (o.rk4_steps if scheme == "rk4" else o.euler_steps)(nsteps)
    want = o.get_state()
    assert n == o.steps and t == o.time
    return got, want


CASES = [
    # system, n, stencil, ko, periodic, abl
    ("phipi", (24, 20, 16), 2, 0, (1, 1, 1), 0),
    ("phipi", (24, 20, 16), 2, 3, (1, 1, 1), 0),
    ("phipi", (70, 21, 19), 4, 4, (0, 0, 0), 5),
    ("phipi", (33, 18, 22), 3, 4, (1, 0, 1), 0),
    ("phipi", (17, 9, 12), 1, 2, (0, 1, 0), 3),
    ("phiu", (24, 20, 16), 1, 0, (1, 1, 1), 0),
    ("phiu", (30, 14, 18), 3, 3, (0, 0, 0), 4),
    ("nonlinear", (24, 20, 16), 2, 2, (1, 1, 1), 0),
    ("nonlinear", (66, 12, 14), 4, 0, (0, 0, 0), 3),
    ("tensor", (20, 16, 12), 3, 3, (1, 1, 1), 0),
    ("tensor", (24, 18, 20), 2, 0, (0, 0, 0), 4),
    ("tensor", (17, 19, 15), 3, 3, (1, 0, 1), 3),
]


@pytest.mark.parametrize("system,n,sn,ko,per,abl", CASES)
@pytest.mark.parametrize("ghost_mode", [0, 1])
def test_rk4_matches_oracle(system, n, sn, ko, per, abl, ghost_mode):
    _say(f"[parity] rk4 {system} n={n} stencil={sn} ko={ko} periodic={per} abl={abl} ghost_mode={ghost_mode}")
    case, state = small_case(system, n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl,
                             gamma=0.05 if system == "nonlinear" else 0.0)
    got, want = run_both(case, state, 7, kernel={"ghost_mode": ghost_mode})
    err = rel_l2_per_comp(got, want)
    assert np.all(err <= REL_TOL), err
    assert np.array_equal(got, want), f"max abs diff {np.max(np.abs(got - want))}"


@pytest.mark.parametrize("system,n,sn,ko,per,abl", CASES[:4] + CASES[5:7])
def test_euler_matches_oracle(system, n, sn, ko, per, abl):
    _say(f"[parity] euler {system} n={n}")
    case, state = small_case(system, n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl)
    case = case.replace(dt=0.05 * case.dt)
    got, want = run_both(case, state, 5, scheme="euler")
    assert np.all(rel_l2_per_comp(got, want) <= REL_TOL)
    assert np.array_equal(got, want)


TILE_CFGS = [
    {"ghost_mode": 0, "tile_y": 2, "zchunk": 3, "xcd_remap": 0, "zalt": 1, "halo_ring": 0},
    {"ghost_mode": 0, "tile_y": 16, "zchunk": 7, "xcd_remap": 1, "zalt": 1, "halo_ring": 1},
    {"ghost_mode": 0, "tile_y": 16, "zchunk": 9, "xcd_remap": 0, "zalt": 0, "halo_ring": 0},
    {"ghost_mode": 1, "tile_y": 4, "zchunk": 5, "xcd_remap": 1, "zalt": 1, "halo_ring": 0},
    {"ghost_mode": 1, "tile_y": 8, "zchunk": 1, "xcd_remap": 0, "zalt": 0, "halo_ring": 1},
    {"ghost_mode": 0, "tile_y": 8, "zchunk": 2, "xcd_remap": 1, "zalt": 1, "halo_ring": 1},
    {"ghost_mode": 0, "tile_y": 32, "zchunk": 6, "xcd_remap": 1, "zalt": 1, "halo_ring": 1},
    {"ghost_mode": 1, "tile_y": 32, "zchunk": 23, "xcd_remap": 0, "zalt": 0, "halo_ring": 0},
    {"ghost_mode": 0, "tile_y": 16, "zchunk": 7, "xcd_remap": 1, "zalt": 2, "halo_ring": 1},
    {"ghost_mode": 1, "tile_y": 8, "zchunk": 4, "xcd_remap": 0, "zalt": 2, "halo_ring": 0},
]


@pytest.mark.parametrize("system,n,stencil_n,ko,periodic,abl", [CASES[2], CASES[3], CASES[6], CASES[8]])
def test_pointer_access_kernel_matches_oracle(monkeypatch, system, n, stencil_n, ko, periodic, abl):
    """The stage kernel that registers of 4 GiB or more use (64-bit pointers,
    64 x 8 tiles), forced on small boxes."""
    _say(f"[parity] pointer access {system} n={n} stencil={stencil_n} ko={ko}")
    monkeypatch.setenv("TSP_POINTER_ACCESS", "1")
    case, state = small_case(system, n=n, stencil_n=stencil_n, ko_order=ko, periodic=periodic, abl=abl,
                             gamma=0.05 if system == "nonlinear" else 0.0)
    got, want = run_both(case, state, 3)
    assert np.array_equal(got, want)


def test_every_tile_shape_is_bitwise_identical():
    _say("[parity] tile shapes")
    case, state = small_case("phipi", n=(130, 27, 23), stencil_n=4, ko_order=4, periodic=(0, 1, 0), abl=4)
    outs = [run_both(case, state, 4, kernel=cfg)[0] for cfg in TILE_CFGS]
    _, want = run_both(case, state, 4)
    for o in outs:
        assert np.array_equal(o, want)


@pytest.mark.parametrize("stencil_n, ko_order", [(3, 3), (2, 0), (4, 4)])
def test_tensor_block_shapes_are_bitwise_identical(stencil_n, ko_order):
    """The tensor kernel's block shapes (tile_y 8: 16 x 8 x 2, 4: 8 x 8 x 4, 16: 16 x 8 x 4) and
    z-slab depths of its block order (zchunk) all equal the oracle bitwise, on a box ragged
    against every shape; tsp_autotune picks among them and leaves the state as it was."""
    _say(f"[parity] tensor block shapes n={stencil_n} ko={ko_order}")
    case, state = small_case("tensor", n=(21, 19, 13), stencil_n=stencil_n, ko_order=ko_order,
                             periodic=(1, 0, 0), abl=3, seed=31)
    want = None
    for ty in (8, 4, 16):
        for zc in (1, 2, 3):
            got, want = run_both(case, state, 2, kernel={"ghost_mode": 1, "tile_y": ty, "zchunk": zc})
            assert np.array_equal(got, want), (ty, zc)
    s = Solver(case)
    s.set_state(state)
    s.rk4_steps(1)
    before = s.get_state()
    cfg = s.autotune(1)
    assert cfg["tile_y"] in (4, 8, 16) and cfg["zchunk"] in (1, 2, 4) and s.steps == 1
    assert np.array_equal(s.get_state(), before)
    s.rk4_steps(1)
    assert np.array_equal(s.get_state(), want)
    s.close()


@pytest.mark.parametrize("nsteps", [1, 2, 5, 6])
def test_graph_replay_equals_eager(nsteps):
    _say(f"[parity] graph replay nsteps={nsteps}")
    case, state = small_case("phipi", n=(40, 24, 20), stencil_n=2, ko_order=3, periodic=(1, 1, 1))
    g, want = run_both(case, state, nsteps, graph=True)
    assert np.array_equal(g, want)
    # graph replays after an odd eager step (other parity) and after re-initialisation
    s = Solver(case)
    s.set_state(state)
    s.rk4_steps(1)
    s.graph_rk4_steps(nsteps)
    s.set_state(state)
    s.graph_rk4_steps(nsteps + 1)
    got = s.get_state()
    o = orc.OracleSolver(case)
    o.set_state(state)
    o.rk4_steps(nsteps + 1)
    assert np.array_equal(got, o.get_state())


def test_graph_prepare_and_mixed_schemes():
    """Graphs are keyed by the register roles they were captured with: prepared
    graphs of both parities, odd remainders, and Euler steps (which rotate the
    roles differently) between graph replays all equal the oracle."""
    _say("[parity] graph prepare, Euler between RK4 graph replays")
    case, state = small_case("phipi", n=(40, 24, 20), stencil_n=2, ko_order=3, periodic=(1, 0, 1), abl=3)
    s = Solver(case)
    s.set_state(state)
    s.graph_prepare()
    s.graph_rk4_steps(3)
    s.euler_steps(1)
    s.graph_rk4_steps(4)
    s.graph_prepare()
    s.euler_steps(1)
    s.graph_rk4_steps(5)
    got = s.get_state()
    o = orc.OracleSolver(case)
    o.set_state(state)
    o.rk4_steps(3)
    o.euler_steps(1)
    o.rk4_steps(4)
    o.euler_steps(1)
    o.rk4_steps(5)
    assert np.array_equal(got, o.get_state())
    assert s.steps == o.steps == 14 and s.time == o.time


@pytest.mark.parametrize("per,bc", [((1, 1, 1), "noop"), ((0, 0, 0), "dirichlet"), ((1, 0, 1), "dirichlet")])
def test_ghost_fill_matches_oracle(per, bc):
    _say(f"[parity] ghost fill periodic={per} bc={bc}")
    case, state = small_case("phiu", n=(21, 13, 11), stencil_n=3, ko_order=4, periodic=per)
    case = case.replace(bc=bc, target=tuple(0.5 + c for c in range(12)))
    s = Solver(case)
    s.set_state(state)
    s.fill_ghosts(_lib.ROLE_OLD)
    s.sync()
    got = s.get_register(_lib.ROLE_OLD)
    o = orc.OracleSolver(case)
    o.set_state(state)
    o.fill_ghosts(0)
    assert np.array_equal(got, o.get_register(0))


def test_wave64_full_workload_matches_oracle():
    """64^3 periodic (phi, Pi), n=2, KO r=3 sigma=0.1, 100 RK4 steps."""
    _say("[parity] wave64 100 steps")
    case, state, nsteps = wave64()
    got, want = run_both(case, state, nsteps, graph=True)
    err = rel_l2_per_comp(got, want)
    _say(f"  rel L2 per comp {err}")
    assert np.all(err <= REL_TOL)
    assert np.array_equal(got, want)


def test_wave256_full_size_one_step_matches_oracle():
    """256^3, n=4, KO r=4, 16-cell ABL, Dirichlet: one RK4 step at full size against the oracle."""
    _say("[parity] wave256 full size, 1 step")
    case, state, _ = wave256()
    got, want = run_both(case, state, 1)
    assert np.all(rel_l2_per_comp(got, want) <= REL_TOL)
    assert np.array_equal(got, want)


def test_tensor_workload_at_64_matches_oracle():
    """The linearised ADM workload (12 comps, n=3, KO r=3, 12-cell ABL, 16 Fourier modes per
    component) at 64^3 for 3 RK4 steps (the oracle needs about a minute per step at 192^3)."""
    _say("[parity] tensor workload at 64^3, 3 steps")
    from treesap_amd.workloads import tensor192

    case, state, _ = tensor192(n=64)
    got, want = run_both(case, state, 3)
    assert np.all(rel_l2_per_comp(got, want) <= REL_TOL)
    assert np.array_equal(got, want)


def test_wave256_full_size_properties():
    """Size-independent properties at 256^3 over 40 steps: linearity under exact scaling
    by 2 and mirror symmetry of reflected data (all operators are reflection symmetric)."""
    _say("[parity] wave256 properties")
    case, state, _ = wave256()
    sym = state + state[:, :, :, ::-1]  # mirror-symmetric in x
    a = Solver(case)
    a.set_state(sym)
    a.graph_rk4_steps(40)
    ra = a.get_state()
    a.set_state(2.0 * sym)
    a.graph_rk4_steps(40)
    rb = a.get_state()
    a.close()
    assert np.array_equal(rb, 2.0 * ra)
    assert np.array_equal(ra, ra[:, :, :, ::-1])
    assert np.isfinite(ra).all() and np.abs(ra).max() > 1e-3


def test_device_pointers_in_and_out():
    """set/get state through device memory (torch tensors) equals the host path."""
    torch = pytest.importorskip("torch")
    case, state = small_case("phipi", n=(32, 16, 8), stencil_n=2, ko_order=2)
    s = Solver(case)
    s.set_state(torch.from_numpy(state).cuda())
    s.rk4_steps(3)
    out = torch.empty(case.shape, dtype=torch.float64, device="cuda")
    s.get_state(out)
    host = s.get_state()
    assert np.array_equal(out.cpu().numpy(), host)


def test_autotune_restores_state():
    case, state = small_case("phipi", n=(64, 32, 32), stencil_n=4, ko_order=4, periodic=(0, 0, 0), abl=4)
    s = Solver(case)
    s.set_state(state)
    s.rk4_steps(2)
    before = s.get_state()
    cfg = s.autotune(1)
    assert cfg["tile_y"] in (2, 4, 8, 16, 32)
    assert s.steps == 2
    assert np.array_equal(s.get_state(), before)
    s.rk4_steps(3)
    o = orc.OracleSolver(case)
    o.set_state(state)
    o.rk4_steps(5)
    assert np.array_equal(s.get_state(), o.get_state())


# (stencil width, KO order) pairs every fused-kernel system supports
_FUSED_PAIRS = [(1, 0), (1, 1), (1, 2), (2, 0), (2, 2), (2, 3), (3, 0), (3, 3), (3, 4), (4, 0), (4, 4)]


def _random_case(rng):
    system = ["phipi", "phiu", "nonlinear"][int(rng.integers(3))]
    sn, ko = _FUSED_PAIRS[int(rng.integers(len(_FUSED_PAIRS)))]
    H = max(sn, ko)
    n = (int(rng.integers(max(H, 5), 75)), int(rng.integers(max(H, 3), 41)), int(rng.integers(max(H, 3), 31)))
    per = tuple(int(v) for v in rng.integers(0, 2, 3))
    abl = int(rng.integers(1, min(n) // 2)) if min(n) >= 6 and rng.random() < 0.5 else 0
    cfg = {"ghost_mode": int(rng.integers(2)), "tile_y": int(rng.choice([2, 4, 8, 16, 32])),
           "zchunk": int(rng.integers(1, n[2] + 1)), "xcd_remap": int(rng.integers(2)), "zalt": int(rng.integers(3)),
           "halo_ring": int(rng.integers(2))}
    return system, n, sn, ko, per, abl, cfg


# (stencil width, KO order) pairs the tensor kernel supports
_TENSOR_PAIRS = [(1, 0), (1, 1), (2, 0), (2, 2), (3, 0), (3, 3), (4, 0), (4, 4)]


@pytest.mark.parametrize("seed", range(12))
def test_random_tensor_cases_match_oracle(seed):
    """Seeded random tensor (h, K) cases: box shapes ragged against every block shape, stencil/KO
    pairs, boundary mixes, absorbing layers, block shape and z-slab depth of the block order, RK4 or
    Euler; bitwise against the oracle."""
    rng = np.random.default_rng(3000 + seed)
    sn, ko = _TENSOR_PAIRS[int(rng.integers(len(_TENSOR_PAIRS)))]
    H = max(sn, ko)
    n = (int(rng.integers(max(H, 3), 41)), int(rng.integers(max(H, 3), 29)), int(rng.integers(max(H, 3), 23)))
    per = tuple(int(v) for v in rng.integers(0, 2, 3))
    abl = int(rng.integers(1, min(n) // 2)) if min(n) >= 6 and rng.random() < 0.5 else 0
    cfg = {"ghost_mode": 1, "tile_y": int(rng.choice([4, 8, 16])), "zchunk": int(rng.integers(1, 6))}
    scheme = "euler" if rng.random() < 0.25 else "rk4"
    _say(f"[parity] random tensor seed={seed} n={n} stencil={sn} ko={ko} periodic={per} abl={abl} cfg={cfg} {scheme}")
    case, state = small_case("tensor", n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl, seed=90 + seed)
    got, want = run_both(case, state, 2, scheme=scheme, kernel=cfg)
    assert np.array_equal(got, want), f"max abs diff {np.max(np.abs(got - want))}"


@pytest.mark.parametrize("seed", range(24))
def test_random_shapes_and_configs_match_oracle(seed):
    """Seeded random systems, box shapes (ragged against every tile shape, z chunks down to one plane,
    shorter than the halo ring's lookahead), boundary mixes and kernel configurations, bitwise against
    the oracle over three RK4 steps."""
    rng = np.random.default_rng(1000 + seed)
    system, n, sn, ko, per, abl, cfg = _random_case(rng)
    _say(f"[parity] random seed={seed} {system} n={n} stencil={sn} ko={ko} periodic={per} abl={abl} cfg={cfg}")
    case, state = small_case(system, n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl, seed=50 + seed,
                             gamma=0.05 if system == "nonlinear" else 0.0)
    got, want = run_both(case, state, 3, kernel=cfg)
    assert np.array_equal(got, want), f"max abs diff {np.max(np.abs(got - want))}"
