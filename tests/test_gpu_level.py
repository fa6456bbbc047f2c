"""Multi-box levels on one GPU (tsp_level_*): the domain cut into boxes of at most
`block` cells per axis, ghosts exchanged between the boxes before every stage.
Results must not depend on the block size: every configuration is bitwise equal
to the single-box solver and to the oracle."""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from treesap_amd import Level, Solver
from treesap_amd.workloads import small_case

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


LEVEL_CASES = [
    # (n, stencil, KO order, periodic, ABL, blocks)
    ((40, 36, 30), 4, 4, (0, 0, 0), 6, [(16, 16, 16), (13, 7, 9), (4, 4, 4), (64, 64, 8)]),
    ((33, 20, 24), 2, 3, (1, 0, 1), 0, [(8, 12, 5), (11, 20, 24), (3, 3, 3)]),
    ((24, 24, 24), 3, 3, (1, 1, 1), 0, [(12, 12, 12), (7, 24, 10)]),
]


@pytest.mark.parametrize("n,sn,ko,per,abl,blocks", LEVEL_CASES)
def test_level_matches_single_box_and_oracle(n, sn, ko, per, abl, blocks):
    case, state = small_case("phipi", n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl, seed=21)
    o = orc.OracleSolver(case)
    o.set_state(state)
    o.rk4_steps(5)
    want = o.get_state()
    s = Solver(case)
    s.set_state(state)
    s.rk4_steps(5)
    assert np.array_equal(s.get_state(), want)
    for blk in blocks:
        _say(f"[level] n={n} stencil={sn} ko={ko} periodic={per} abl={abl} block={blk}")
        lv = Level(case, blk)
        # the boxes tile the domain, at most `blk` cells per axis
        covered = np.zeros(n[::-1], dtype=np.int32)
        for lo0, lo1, lo2, hi0, hi1, hi2 in lv.boxes:
            assert all(hi - lo + 1 <= b for lo, hi, b in zip((lo0, lo1, lo2), (hi0, hi1, hi2), blk))
            covered[lo2:hi2 + 1, lo1:hi1 + 1, lo0:hi0 + 1] += 1
        assert (covered == 1).all()
        lv.set_state(state)
        assert np.array_equal(lv.get_state(), state)
        lv.rk4_steps(2)
        lv.graph_rk4_steps(3)
        got = lv.get_state()
        assert np.array_equal(got, want), f"block {blk}: max abs diff {np.max(np.abs(got - want))}"
        assert lv.steps == 5 and lv.time == o.time
        lv.close()


def test_level_euler_and_kernel_configs():
    """Euler steps and every batch configuration on a level with an absorbing layer."""
    case, state = small_case("phipi", n=(37, 29, 26), stencil_n=4, ko_order=4, periodic=(0, 1, 0), abl=5, seed=5)
    case = case.replace(dt=0.1 * case.dt)
    o = orc.OracleSolver(case)
    o.set_state(state)
    o.euler_steps(3)
    want_e = o.get_state()
    o.set_state(state)
    o.rk4_steps(3)
    want_r = o.get_state()
    for cfg in ({"tile_y": 8, "zchunk": 5, "grid_blocks": 7, "zalt": 0, "halo_ring": 1, "narrow_tiles": 1},
                {"tile_y": 16, "zchunk": 0, "grid_blocks": 0, "zalt": 1, "halo_ring": 0, "narrow_tiles": 0}):
        _say(f"[level] euler and rk4 with {cfg}")
        lv = Level(case, (20, 11, 14))
        lv.kernel = cfg
        lv.set_state(state)
        lv.euler_steps(3)
        assert np.array_equal(lv.get_state(), want_e)
        lv.set_state(state)
        lv.rk4_steps(3)
        assert np.array_equal(lv.get_state(), want_r)
        lv.close()


@pytest.mark.parametrize("system,gamma", [("phiu", 0.0), ("nonlinear", 0.3)])
def test_level_of_other_systems(system, gamma):
    """(phi, u) and NonLinear levels: every block size bitwise equal to the oracle's single box
    (Dirichlet faces with an absorbing layer, and a mixed-periodic box)."""
    for n, sn, ko, per, abl, blocks in (((30, 26, 22), 2, 2, (0, 0, 0), 4, [(11, 9, 22), (6, 6, 6)]),
                                        ((24, 20, 18), 3, 3, (1, 0, 1), 0, [(10, 7, 18), (24, 20, 5)])):
        case, state = small_case(system, n=n, stencil_n=sn, ko_order=ko, periodic=per, abl=abl, seed=33, gamma=gamma)
        o = orc.OracleSolver(case)
        o.set_state(state)
        o.rk4_steps(4)
        want = o.get_state()
        for blk in blocks:
            _say(f"[level] {system} n={n} block={blk}")
            lv = Level(case, blk)
            lv.set_state(state)
            lv.rk4_steps(1)
            lv.graph_rk4_steps(3)
            got = lv.get_state()
            assert np.array_equal(got, want), f"{system} block {blk}: max abs diff {np.max(np.abs(got - want))}"
            lv.close()
