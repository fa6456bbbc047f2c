"""Batched mode: every instance of a batch must equal (bitwise) its solve on its
own, and the oracle, for every batch configuration and for graph replay."""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from treesap_amd import Batch, Solver
from treesap_amd.workloads import ensemble, small_case

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mixed_cases():
    specs = [
        ((70, 21, 19), 4, 4, (0, 0, 0), 5),
        ((24, 20, 16), 2, 3, (1, 1, 1), 0),
        ((33, 18, 22), 3, 3, (1, 0, 1), 0),
        ((17, 9, 12), 1, 1, (0, 1, 0), 3),
        ((40, 24, 20), 2, 2, (1, 1, 1), 0),
        ((32, 32, 32), 4, 0, (1, 1, 1), 0),
        ((48, 16, 8), 2, 0, (0, 0, 0), 2),
        ((64, 20, 30), 3, 0, (1, 1, 1), 0),
        ((96, 40, 12), 2, 2, (1, 1, 1), 0),
    ]
    out = []
    for k, (n, sn, ko, per, abl) in enumerate(specs):
        case, state = small_case("phipi", n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl, seed=11 + k)
        case = case.replace(dt=case.dt * (1.0 + 0.1 * k))  # every instance has its own time step
        out.append((case, state))
    return out


BATCH_CFGS = [
    {"tile_y": 16, "zchunk": 0, "grid_blocks": 0, "zalt": 1, "halo_ring": 1, "narrow_tiles": 1},
    {"tile_y": 8, "zchunk": 5, "grid_blocks": 7, "zalt": 0, "halo_ring": 1, "narrow_tiles": 1},
    {"tile_y": 16, "zchunk": 64, "grid_blocks": 300, "zalt": 1, "halo_ring": 0, "narrow_tiles": 1},
    {"tile_y": 8, "zchunk": 0, "grid_blocks": 0, "zalt": 1, "halo_ring": 0, "narrow_tiles": 0},
    {"tile_y": 16, "zchunk": 3, "grid_blocks": 0, "zalt": 1, "halo_ring": 1, "narrow_tiles": 0},
    {"tile_y": 8, "zchunk": 4, "grid_blocks": 0, "zalt": 2, "halo_ring": 1, "narrow_tiles": 1},
]


@pytest.mark.parametrize("cfg", BATCH_CFGS)
def test_batch_equals_single_solves_and_oracle(cfg):
    _say(f"[batch] cfg {cfg}")
    cases = mixed_cases()
    b = Batch([c for c, _ in cases])
    b.kernel = cfg
    for i, (_, st) in enumerate(cases):
        b.set_state(i, st)
    b.rk4_steps(5)
    for i, (case, st) in enumerate(cases):
        got = b.get_state(i)
        s = Solver(case)
        s.set_state(st)
        s.rk4_steps(5)
        assert np.array_equal(got, s.get_state()), i
        o = orc.OracleSolver(case)
        o.set_state(st)
        o.rk4_steps(5)
        assert np.array_equal(got, o.get_state()), i
        assert b.time(i) == o.time
    assert b.steps == 5


def test_batch_graph_equals_eager():
    _say("[batch] graph")
    cases = mixed_cases()
    a = Batch([c for c, _ in cases])
    g = Batch([c for c, _ in cases])
    for i, (_, st) in enumerate(cases):
        a.set_state(i, st)
        g.set_state(i, st)
    a.rk4_steps(7)
    g.graph_rk4_steps(3)
    g.graph_rk4_steps(4)
    for i in range(len(cases)):
        assert np.array_equal(a.get_state(i), g.get_state(i))


def test_batch_graph_after_config_switch():
    """A captured graph holds the work list of its configuration: graph steps
    under A, a switch to B (work list rebuilt), eager steps, back to A (rebuilt
    again), graph steps: every instance equals plain eager stepping."""
    _say("[batch] graph after config switch")
    cases = mixed_cases()
    A = dict(BATCH_CFGS[0])
    B = dict(BATCH_CFGS[1])
    a = Batch([c for c, _ in cases])
    g = Batch([c for c, _ in cases])
    for i, (_, st) in enumerate(cases):
        a.set_state(i, st)
        g.set_state(i, st)
    a.rk4_steps(9)
    g.kernel = A
    g.graph_rk4_steps(4)
    g.kernel = B
    g.rk4_steps(1)
    g.kernel = A
    g.graph_rk4_steps(4)
    for i in range(len(cases)):
        assert np.array_equal(a.get_state(i), g.get_state(i)), i


def test_small_ensemble_workload():
    """A 12-instance slice of the ensemble workload (sides 32..64, n in {2,3,4}) for 20 steps."""
    _say("[batch] ensemble slice")
    inst, _ = ensemble(count=12, sides=(32, 48, 64))
    b = Batch([c for c, _ in inst])
    for i, (_, st) in enumerate(inst):
        b.set_state(i, st)
    b.graph_rk4_steps(20)
    for i, (case, st) in enumerate(inst[:4]):
        o = orc.OracleSolver(case)
        o.set_state(st)
        o.rk4_steps(20)
        assert np.array_equal(b.get_state(i), o.get_state()), i


# (stencil width, KO order) pairs of the batched kernel
_BATCH_PAIRS = [(1, 0), (1, 1), (2, 0), (2, 2), (2, 3), (3, 0), (3, 3), (4, 0), (4, 4)]


@pytest.mark.parametrize("seed", range(6))
def test_random_batches_match_oracle(seed):
    """Seeded random batches: 5-9 (phi, Pi) instances of random size, stencil/KO pair, boundaries and time
    step under a random batch configuration (tile height, z chunk, grid, march order, halo ring, narrow
    tiles); every instance bitwise against the oracle after three RK4 steps."""
    rng = np.random.default_rng(2000 + seed)
    cases = []
    for k in range(int(rng.integers(5, 10))):
        sn, ko = _BATCH_PAIRS[int(rng.integers(len(_BATCH_PAIRS)))]
        H = max(sn, ko)
        n = (int(rng.integers(max(H, 5), 110)), int(rng.integers(max(H, 3), 40)), int(rng.integers(max(H, 3), 24)))
        per = tuple(int(v) for v in rng.integers(0, 2, 3))
        abl = int(rng.integers(1, min(n) // 2)) if min(n) >= 6 and rng.random() < 0.4 else 0
        case, state = small_case("phipi", n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl, seed=300 + 10 * seed + k)
        cases.append((case.replace(dt=case.dt * (1.0 + 0.05 * k)), state))
    # tile height, halo ring and narrow tiles cover their combinations over the seeds; the rest is random
    cfg = {"tile_y": [8, 16][seed % 2], "zchunk": int(rng.choice([0, 1, 5, 64])),
           "grid_blocks": int(rng.choice([0, 7, 300])), "zalt": int(rng.integers(3)),
           "halo_ring": [1, 1, 0, 1, 0, 1][seed], "narrow_tiles": [1, 1, 1, 0, 0, 1][seed]}
    _say(f"[batch] random seed={seed} {len(cases)} instances {[c.n for c, _ in cases]} cfg={cfg}")
    b = Batch([c for c, _ in cases])
    b.kernel = cfg
    for i, (_, st) in enumerate(cases):
        b.set_state(i, st)
    b.rk4_steps(3)
    for i, (case, st) in enumerate(cases):
        o = orc.OracleSolver(case)
        o.set_state(st)
        o.rk4_steps(3)
        assert np.array_equal(b.get_state(i), o.get_state()), (i, case.n)


@pytest.mark.parametrize("seed", range(6))
def test_random_group_configurations_match_oracle(seed):
    """Seeded random batches in which every (stencil width, KO order) group runs its own random
    kernel configuration; every instance bitwise against the oracle after one eager and two
    graph-replayed RK4 steps."""
    rng = np.random.default_rng(2500 + seed)
    cases = []
    for k in range(int(rng.integers(5, 10))):
        sn, ko = _BATCH_PAIRS[int(rng.integers(len(_BATCH_PAIRS)))]
        H = max(sn, ko)
        n = (int(rng.integers(max(H, 5), 110)), int(rng.integers(max(H, 3), 40)), int(rng.integers(max(H, 3), 24)))
        per = tuple(int(v) for v in rng.integers(0, 2, 3))
        abl = int(rng.integers(1, min(n) // 2)) if min(n) >= 6 and rng.random() < 0.4 else 0
        case, state = small_case("phipi", n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl, seed=500 + 10 * seed + k)
        cases.append((case.replace(dt=case.dt * (1.0 + 0.05 * k)), state))
    b = Batch([c for c, _ in cases])
    groups = sorted({(c.stencil_n, c.ko_order) for c, _ in cases})
    cfgs = {}
    for sn, ko in groups:
        cfgs[(sn, ko)] = {"tile_y": int(rng.choice([8, 16])), "zchunk": int(rng.choice([0, 1, 5, 64])),
                          "grid_blocks": int(rng.choice([0, 7, 300])), "zalt": int(rng.integers(3)),
                          "halo_ring": int(rng.integers(2)), "narrow_tiles": int(rng.integers(2))}
        b.set_group_kernel(sn, ko, cfgs[(sn, ko)])
    _say(f"[batch] random group configurations seed={seed} {[c.n for c, _ in cases]} {cfgs}")
    for i, (_, st) in enumerate(cases):
        b.set_state(i, st)
    b.rk4_steps(1)
    b.graph_rk4_steps(2)
    for i, (case, st) in enumerate(cases):
        o = orc.OracleSolver(case)
        o.set_state(st)
        o.rk4_steps(3)
        assert np.array_equal(b.get_state(i), o.get_state()), (i, case.n, case.stencil_n, case.ko_order)
    b.close()


def test_batch_autotune_restores_state_and_keeps_results():
    """tsp_batch_autotune times its candidates on the batch itself: afterwards every instance's
    state, the step count and the clock are what they were, and stepping with the chosen
    configuration equals the oracle."""
    _say("[batch] autotune")
    cases = mixed_cases()
    b = Batch([c for c, _ in cases])
    for i, (_, st) in enumerate(cases):
        b.set_state(i, st)
    b.rk4_steps(1)
    before = [b.get_state(i) for i in range(len(cases))]
    t_before = [b.time(i) for i in range(len(cases))]
    cfg = b.autotune(2)
    assert cfg == b.kernel and cfg["tile_y"] in (8, 16)
    assert b.steps == 1 and [b.time(i) for i in range(len(cases))] == t_before
    for i in range(len(cases)):
        assert np.array_equal(b.get_state(i), before[i]), i
    b.graph_rk4_steps(2)
    for i, (case, st) in enumerate(cases):
        o = orc.OracleSolver(case)
        o.set_state(st)
        o.rk4_steps(3)
        assert np.array_equal(b.get_state(i), o.get_state()), i


def test_group_kernels_run_their_own_configurations():
    """Every (system, stencil width, KO order) group runs its own kernel configuration
    (tsp_batch_set_group_kernel): with other tile heights, z chunks, grids, rings and narrow
    tiles per group, every instance stays bitwise equal to the oracle, eager and graph-replayed;
    tsp_batch_set_kernel gives every group one configuration again; pairs without batched
    kernels and invalid configurations are refused."""
    from treesap_amd._lib import TspError

    _say("[batch] per-group configurations")
    cases = mixed_cases()
    per_group = {
        (4, 4): {"tile_y": 8, "zchunk": 7, "grid_blocks": 0, "zalt": 2, "halo_ring": 1, "narrow_tiles": 1},
        (2, 3): {"tile_y": 16, "zchunk": 64, "grid_blocks": 5, "zalt": 0, "halo_ring": 0, "narrow_tiles": 0},
        (3, 3): {"tile_y": 8, "zchunk": 3, "grid_blocks": 0, "zalt": 1, "halo_ring": 0, "narrow_tiles": 1},
        (2, 2): {"tile_y": 16, "zchunk": 0, "grid_blocks": 64, "zalt": 2, "halo_ring": 1, "narrow_tiles": 1},
    }
    b = Batch([c for c, _ in cases])
    default = b.kernel
    for (sn, ko), cfg in per_group.items():
        b.set_group_kernel(sn, ko, cfg)
    for (sn, ko), cfg in per_group.items():
        assert b.group_kernel(sn, ko) == cfg
    assert b.group_kernel(1, 1) == default and b.kernel == default
    for i, (_, st) in enumerate(cases):
        b.set_state(i, st)
    b.rk4_steps(3)
    b.graph_rk4_steps(2)
    for i, (case, st) in enumerate(cases):
        o = orc.OracleSolver(case)
        o.set_state(st)
        o.rk4_steps(5)
        assert np.array_equal(b.get_state(i), o.get_state()), (i, case.stencil_n, case.ko_order)
    b.kernel = BATCH_CFGS[3]
    assert all(b.group_kernel(sn, ko) == BATCH_CFGS[3] for sn, ko in list(per_group) + [(1, 1)])
    with pytest.raises(TspError):
        b.group_kernel(3, 1)  # no batched kernels for KO order 1 with n = 3
    with pytest.raises(TspError):
        b.set_group_kernel(2, 2, {"tile_y": 4})
    assert b.group_kernel(2, 2) == BATCH_CFGS[3]
    b.close()


def test_batch_tune_cache_holds_one_line_per_group(monkeypatch, tmp_path):
    """tsp_batch_autotune stores one cache line per (system, stencil width, KO order) group; a
    second batch of the same shape takes every group's configuration from the cache (an edited
    line wins), and its instances still equal the oracle."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("TSP_TUNE_CACHE", str(cache))
    _say("[batch] tune cache per group")
    cases = mixed_cases()[:5]  # groups (4,4), (2,3), (3,3), (1,1), (2,2)
    b = Batch([c for c, _ in cases])
    for i, (_, st) in enumerate(cases):
        b.set_state(i, st)
    b.autotune(2)
    tuned = {(c.stencil_n, c.ko_order): b.group_kernel(c.stencil_n, c.ko_order) for c, _ in cases}
    b.close()
    lines = cache.read_text().splitlines()
    assert len(lines) == len(tuned) and all(len(ln.split()) == 8 for ln in lines)
    # the (2, 2) group's line, edited to another configuration
    groups = [ln.split()[0] for ln in lines]
    key22 = [k for k in groups if k.endswith("|g3")]  # group 3: (phi, Pi) with (stencil_n, ko_order) = (2, 2)
    assert len(key22) == 1
    with cache.open("a") as f:
        f.write(f"{key22[0]} 8 3 17 2 0 0 0.1\n")
    b = Batch([c for c, _ in cases])
    for i, (_, st) in enumerate(cases):
        b.set_state(i, st)
    b.autotune(2)
    for (sn, ko), cfg in tuned.items():
        want = cfg if (sn, ko) != (2, 2) else dict(tile_y=8, zchunk=3, grid_blocks=17, zalt=2, halo_ring=0,
                                                    narrow_tiles=0)
        assert b.group_kernel(sn, ko) == want, (sn, ko)
    b.graph_rk4_steps(2)
    for i, (case, st) in enumerate(cases):
        o = orc.OracleSolver(case)
        o.set_state(st)
        o.rk4_steps(2)
        assert np.array_equal(b.get_state(i), o.get_state()), i
    b.close()


def test_mixed_system_batch_matches_oracle():
    """(phi, Pi), (phi, u) and NonLinear instances in one batch: one launch per stage and
    (system, stencil width, KO order) group; every instance bitwise equal to its single-box
    solve and to the oracle, eager and graph-replayed, RK4 and Euler, under two batch
    configurations."""
    _say("[batch] mixed systems")
    specs = [("phiu", (40, 22, 18), 2, 2, (0, 0, 0), 4, 0.0), ("nonlinear", (33, 20, 16), 2, 3, (1, 1, 1), 0, 0.3),
             ("phipi", (24, 20, 16), 3, 3, (1, 0, 1), 0, 0.0), ("nonlinear", (70, 12, 14), 4, 4, (0, 0, 0), 3, 0.1),
             ("phiu", (20, 18, 24), 1, 1, (1, 1, 1), 0, 0.0), ("phiu", (64, 16, 10), 4, 0, (0, 1, 0), 0, 0.0),
             ("nonlinear", (17, 9, 12), 1, 0, (0, 1, 0), 0, 0.2)]
    cases = []
    for k, (sysname, n, sn, ko, per, abl, gamma) in enumerate(specs):
        case, state = small_case(sysname, n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl, seed=70 + k,
                                 gamma=gamma)
        cases.append((case.replace(dt=case.dt * (1.0 + 0.07 * k)), state))
    for cfg, graph in ((BATCH_CFGS[0], False), (BATCH_CFGS[1], True)):
        b = Batch([c for c, _ in cases])
        b.kernel = cfg
        for i, (_, st) in enumerate(cases):
            b.set_state(i, st)
        if graph:
            b.graph_rk4_steps(4)
        else:
            b.rk4_steps(4)
        b.euler_steps(1)
        for i, (case, st) in enumerate(cases):
            o = orc.OracleSolver(case)
            o.set_state(st)
            o.rk4_steps(4)
            o.euler_steps(1)
            assert np.array_equal(b.get_state(i), o.get_state()), (i, case.system, cfg)
            s = Solver(case)
            s.set_state(st)
            s.rk4_steps(4)
            s.euler_steps(1)
            assert np.array_equal(b.get_state(i), s.get_state()), (i, case.system)
        b.close()
