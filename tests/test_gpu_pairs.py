"""GPU parity of the stage-pair mode (RK4 stages (0,1) and (2,3) each in one launch).

The pair kernel evaluates every cell with the same expression as the single-stage
kernel, so the state after any number of steps must equal the oracle's bit for bit,
for every box shape against the 32 x 32 extended tiles, boundary mix (periodic
wrap of the first stage's ring, ghost values of the intermediate register along
non-periodic faces), absorbing layer and z chunk.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from treesap_amd import Solver, _lib
from treesap_amd.workloads import small_case, wave256

pytestmark = pytest.mark.gpu

# (stencil width, KO order) pairs with a pair kernel: KO dissipation on
PAIRS = [(1, 1), (1, 2), (2, 2), (2, 3), (3, 3), (3, 4), (4, 4)]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _solve(case, state, nsteps, kernel, graph=False):
    s = Solver(case)
    s.kernel = kernel
    assert s.kernel["stage_pairs"] == kernel.get("stage_pairs", 0)
    s.set_state(state)
    (s.graph_rk4_steps if graph else s.rk4_steps)(nsteps)
    got, t, n = s.get_state(), s.time, s.steps
    s.close()
    return got, t, n


def _oracle(case, state, nsteps):
    o = orc.OracleSolver(case)
    o.set_state(state)
    o.rk4_steps(nsteps)
    return o.get_state(), o.time, o.steps


FIXED = [
    # n, stencil, ko, periodic, abl, zchunk, xcd_remap
    ((24, 20, 16), 2, 3, (1, 1, 1), 0, 16, 1),
    ((64, 64, 64), 2, 3, (1, 1, 1), 0, 32, 0),
    ((70, 21, 19), 4, 4, (0, 0, 0), 5, 7, 1),
    ((49, 50, 23), 4, 4, (0, 0, 0), 6, 23, 0),
    ((33, 18, 22), 3, 4, (1, 0, 1), 0, 5, 1),
    ((17, 9, 12), 1, 2, (0, 1, 0), 3, 1, 1),
    ((40, 26, 30), 3, 3, (1, 1, 0), 4, 64, 0),
    ((8, 8, 8), 4, 4, (1, 1, 1), 0, 8, 1),
    ((25, 24, 9), 1, 1, (0, 0, 1), 2, 3, 0),
]


@pytest.mark.parametrize("n,sn,ko,per,abl,zc,xr", FIXED)
def test_pairs_match_oracle(n, sn, ko, per, abl, zc, xr):
    _say(f"[pairs] n={n} stencil={sn} ko={ko} periodic={per} abl={abl} zchunk={zc} xcd_remap={xr}")
    case, state = small_case("phipi", n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl, seed=7)
    got, t, k = _solve(case, state, 3, dict(ghost_mode=0, zchunk=zc, xcd_remap=xr, stage_pairs=1))
    want, to, ko_ = _oracle(case, state, 3)
    assert (k, t) == (ko_, to)
    assert np.array_equal(got, want), f"max abs diff {np.max(np.abs(got - want))}"


@pytest.mark.parametrize("seed", range(16))
def test_random_pair_cases_match_oracle(seed):
    """Seeded random stencils, boxes (periodic axes of at least 2H cells), boundary mixes, absorbing
    layers and z chunks, bitwise against the oracle and against the single-stage kernels."""
    rng = np.random.default_rng(4000 + seed)
    sn, ko = PAIRS[int(rng.integers(len(PAIRS)))]
    H = max(sn, ko)
    per = tuple(int(v) for v in rng.integers(0, 2, 3))
    n = tuple(int(rng.integers(2 * H if p else H, hi)) for p, hi in zip(per, (80, 50, 34)))
    abl = int(rng.integers(1, min(n) // 2)) if min(n) >= 6 and rng.random() < 0.5 else 0
    cfg = dict(ghost_mode=0, zchunk=int(rng.integers(1, n[2] + 1)), xcd_remap=int(rng.integers(2)), stage_pairs=1)
    _say(f"[pairs] random seed={seed} n={n} stencil={sn} ko={ko} periodic={per} abl={abl} cfg={cfg}")
    case, state = small_case("phipi", n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl, seed=90 + seed)
    got, _, _ = _solve(case, state, 4, cfg)
    single, _, _ = _solve(case, state, 4, dict(ghost_mode=0, stage_pairs=0))
    want, _, _ = _oracle(case, state, 4)
    assert np.array_equal(single, want)
    assert np.array_equal(got, want), f"max abs diff {np.max(np.abs(got - want))}"


@pytest.mark.parametrize("nsteps", [2, 5])
def test_pair_graph_replay_equals_oracle(nsteps):
    case, state = small_case("phipi", n=(40, 30, 20), stencil_n=4, ko_order=4, periodic=(0, 0, 0), abl=4, seed=11)
    got, _, _ = _solve(case, state, nsteps, dict(ghost_mode=0, zchunk=8, stage_pairs=1), graph=True)
    want, _, _ = _oracle(case, state, nsteps)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("system,sn,ko,per,n,kernel,needle", [
    ("phiu", 2, 2, (1, 1, 1), (24, 20, 16), {}, "(phi, Pi)"),
    ("phipi", 2, 0, (1, 1, 1), (24, 20, 16), {}, "KO dissipation"),
    ("phipi", 2, 3, (1, 1, 1), (24, 20, 16), {"ghost_mode": 1}, "ghost_mode 0"),
    ("phipi", 4, 4, (1, 1, 1), (24, 7, 16), {}, "2H cells"),
])
def test_pairs_rejected_where_they_do_not_apply(system, sn, ko, per, n, kernel, needle):
    case, _ = small_case(system, n=n, stencil_n=sn, ko_order=ko, periodic=per)
    s = Solver(case)
    try:
        with pytest.raises(_lib.TspError, match="stage_pairs") as ei:
            s.kernel = dict(kernel, stage_pairs=1)
        assert needle in str(ei.value)
        assert s.kernel["stage_pairs"] == 0
    finally:
        s.close()


def test_wave256_full_size_one_step_with_pairs_matches_oracle():
    case, state, _ = wave256()
    _say("[pairs] wave256 full size, one step")
    got, _, _ = _solve(case, state, 1, dict(ghost_mode=0, zchunk=64, xcd_remap=1, stage_pairs=1))
    want, _, _ = _oracle(case, state, 1)
    assert np.array_equal(got, want), f"max abs diff {np.max(np.abs(got - want))}"


def test_tune_cache_round_trip_with_pairs(monkeypatch, tmp_path):
    """The tuning cache stores all seven configuration fields; a stored stage-pair configuration is
    taken as is (last matching line wins) and stays bitwise, and lines of the six-field format are
    ignored."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("TSP_TUNE_CACHE", str(cache))
    case, state = small_case("phipi", n=(48, 40, 24), stencil_n=2, ko_order=2, periodic=(1, 1, 1), seed=5)
    s = Solver(case)
    s.set_state(state)
    first = s.autotune(1)
    s.close()
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and len(lines[0].split()) == 9
    key = lines[0].split()[0]
    assert [int(v) for v in lines[0].split()[1:8]] == [first[k] for k in
                                                        ("ghost_mode", "tile_y", "zchunk", "xcd_remap", "zalt",
                                                         "halo_ring", "stage_pairs")]
    with cache.open("a") as f:
        f.write(f"{key} 0 8 12 0 0 0 1 0.5\n")   # a stage-pair configuration
        f.write(f"{key} 0 16 64 1 1 1 0.4\n")    # six-field format: ignored
    s = Solver(case)
    s.set_state(state)
    cfg = s.autotune(1)
    assert cfg == dict(ghost_mode=0, tile_y=8, zchunk=12, xcd_remap=0, zalt=0, halo_ring=0, stage_pairs=1)
    s.rk4_steps(3)
    got = s.get_state()
    s.close()
    want, _, _ = _oracle(case, state, 3)
    assert np.array_equal(got, want)


def test_euler_steps_unaffected_by_pair_mode():
    """Euler steps run the single-stage kernel whatever stage_pairs says."""
    case, state = small_case("phipi", n=(30, 20, 18), stencil_n=2, ko_order=3, periodic=(1, 0, 1), abl=3, seed=13)
    s = Solver(case)
    s.kernel = dict(ghost_mode=0, stage_pairs=1)
    s.set_state(state)
    s.euler_steps(3)
    got = s.get_state()
    s.close()
    o = orc.OracleSolver(case)
    o.set_state(state)
    o.euler_steps(3)
    assert np.array_equal(got, o.get_state())
