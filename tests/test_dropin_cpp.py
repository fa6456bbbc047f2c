"""C++ drop-in API (include/treesap/): host checks, RHS arithmetic against the
oracle, and whole TreESaP-style programs on the GPU against the oracle.

The programs under tests/cpp are built by tests/cpp/build.py (part of the
build entry point); a missing binary is built on first use.
"""
from __future__ import annotations

import ctypes as C
import math
import re
import subprocess
import sys
import zlib
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "tests" / "cpp"))

import build as cpp_build  # noqa: E402  (tests/cpp/build.py)

from oracle import oracle as orc  # noqa: E402
from treesap_amd import _lib  # noqa: E402
from treesap_amd.workloads import Case, case_from_desc  # noqa: E402


def _run(args, **kw):
    res = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=300, **kw)
    if res.returncode != 0:
        raise AssertionError(f"{args[0]} exited {res.returncode}\nstdout:\n{res.stdout}\nstderr:\n{res.stderr}")
    return res


def test_host_selftest(tmp_path):
    res = _run([cpp_build.binary("dropin_host"), "selftest", tmp_path])
    assert "passed" in res.stdout


# FunctionInit expressions against Python's evaluation of the same arithmetic.
# The reference hands these strings to AMReX's parser, which is not in the
# reference tree; what is checked is the grammar's mathematical meaning:
# precedence, left-associative + - * /, right-associative power that binds
# tighter than unary minus ("-x^2" is -(x^2), "2^-1" is 0.5), the functions
# and pi. Python's float operators and math functions call the same libm, so
# the values agree to the bit.
EXPR_CASES = [
    ("x + y * z", lambda x, y, z: x + y * z),
    ("x - y - z", lambda x, y, z: x - y - z),
    ("x / y / z", lambda x, y, z: x / y / z),
    ("x * y / z * x", lambda x, y, z: x * y / z * x),
    ("(x + y) * (z - x)", lambda x, y, z: (x + y) * (z - x)),
    ("-x^2", lambda x, y, z: -(x ** 2)),
    ("-x**2 + y", lambda x, y, z: -(x ** 2) + y),
    ("2^3^2", lambda x, y, z: 2.0 ** (3.0 ** 2.0)),
    ("2^-1 + x^-2", lambda x, y, z: 2.0 ** -1.0 + x ** -2.0),
    ("--x", lambda x, y, z: -(-x)),
    ("+x - -y", lambda x, y, z: x - (-y)),
    ("x*-y", lambda x, y, z: x * (-y)),
    ("1e-3 * x + 2.5E2", lambda x, y, z: 1e-3 * x + 2.5e2),
    ("exp(-(x^2 + y^2 + z^2)/0.1^2)", lambda x, y, z: math.exp(-(x ** 2 + y ** 2 + z ** 2) / 0.1 ** 2)),
    ("sqrt(x*x + y*y)", lambda x, y, z: math.sqrt(x * x + y * y)),
    ("sin(pi*x) * cos(2*pi*y)", lambda x, y, z: math.sin(math.pi * x) * math.cos(2 * math.pi * y)),
    ("tan(x) + atan(y) + atan2(y, x)", lambda x, y, z: math.tan(x) + math.atan(y) + math.atan2(y, x)),
    ("asin(z/2) + acos(z/2)", lambda x, y, z: math.asin(z / 2) + math.acos(z / 2)),
    ("sinh(x) - cosh(y) * tanh(z)", lambda x, y, z: math.sinh(x) - math.cosh(y) * math.tanh(z)),
    ("log(abs(x) + 1) + log10(fabs(y) + 2)", lambda x, y, z: math.log(abs(x) + 1) + math.log10(abs(y) + 2)),
    ("floor(10*x) + ceil(10*y)", lambda x, y, z: float(math.floor(10 * x) + math.ceil(10 * y))),
    ("pow(x, 3) + min(x, y) + max(y, z)", lambda x, y, z: x ** 3.0 + min(x, y) + max(y, z)),
    ("heaviside(x) - heaviside(-x)", lambda x, y, z: (1.0 if x > 0 else 0.0) - (1.0 if -x > 0 else 0.0)),
    ("0.5*(1 + tanh((sqrt(x^2+y^2) - 0.3)/0.05))",
     lambda x, y, z: 0.5 * (1 + math.tanh((math.sqrt(x ** 2 + y ** 2) - 0.3) / 0.05))),
    (" x ^ 2 * 3 ", lambda x, y, z: x ** 2 * 3),
]
EXPR_ERRORS = ["x + (y", "w * 2", "foo(x)", "sin(x, y)", "pow(x)", "x +", "3 4"]


# The device program codes (tsp_expr_code), restated from their documented
# meaning in include/treesap_hip.h with Python's math: the check of what the
# drop-in FunctionInit emits and of what tsp_init_expr computes.
_UNARY = {"neg": lambda v: -v, "sqrt": math.sqrt, "exp": math.exp, "log": math.log, "log10": math.log10,
          "sin": math.sin, "cos": math.cos, "tan": math.tan, "asin": math.asin, "acos": math.acos,
          "atan": math.atan, "sinh": math.sinh, "cosh": math.cosh, "tanh": math.tanh, "abs": abs,
          "floor": lambda v: float(math.floor(v)), "ceil": lambda v: float(math.ceil(v)),
          "heaviside": lambda v: 1.0 if v > 0 else 0.0}
_BINARY = {"add": lambda a, b: a + b, "sub": lambda a, b: a - b, "mul": lambda a, b: a * b,
           "div": lambda a, b: a / b, "pow": math.pow, "atan2": math.atan2, "min": min, "max": max}
_CODE_NAME = {v: k for k, v in _lib.EXPR_CODES.items()}
# codes whose device results round as the host's (IEEE operations, exact functions)
_EXACT_CODES = {_lib.EXPR_CODES[k] for k in ("num", "x", "y", "z", "add", "sub", "mul", "div", "neg", "sqrt",
                                             "abs", "floor", "ceil", "heaviside", "min", "max")}


def run_postfix(prog, x, y, z):
    st = []
    for code, value in prog:
        name = _CODE_NAME[code]
        if name == "num":
            st.append(value)
        elif name in ("x", "y", "z"):
            st.append({"x": x, "y": y, "z": z}[name])
        elif name in _BINARY:
            b = st.pop()
            st.append(_BINARY[name](st.pop(), b))
        else:
            st.append(_UNARY[name](st.pop()))
    assert len(st) == 1
    return st[0]


def _parse_program(text):
    f = text.split()
    return [(int(f[i]), float(f[i + 1])) for i in range(0, len(f), 2)]


def _expr_programs(tmp_path, texts, points):
    lines = [f"{x!r} {y!r} {z!r}\t{t}" for t in texts for x, y, z in points]
    f = tmp_path / "exprs.txt"
    f.write_text("\n".join(lines) + "\n")
    out = _run([cpp_build.binary("dropin_host"), "expr", f]).stdout.splitlines()
    assert len(out) == len(lines)
    return lines, out


def test_function_init_expressions_match_python(tmp_path):
    """The host evaluation of every expression, and the postfix program it
    compiles to for the device (run by the documented meaning of its codes),
    equal Python's evaluation of the same arithmetic bit for bit."""
    rng = np.random.default_rng(7)
    pts = rng.uniform(-0.9, 0.9, size=(6, 3)).tolist()
    lines, out = _expr_programs(tmp_path, [t for t, _ in EXPR_CASES] + EXPR_ERRORS, pts)
    nwant = len(EXPR_CASES) * len(pts)
    want = [fn(x, y, z) for _, fn in EXPR_CASES for x, y, z in pts]
    for o, w, line in zip(out[:nwant], want, lines):
        value, prog = o.split("\t")
        assert float(value) == w, (line, value, w)
        x, y, z = (float(v) for v in line.split("\t")[0].split())
        assert run_postfix(_parse_program(prog), x, y, z) == w, (line, prog)
    assert all(v.startswith("error ") for v in out[nwant:]), out[nwant:]
    # the programs put operands before operations: "x - y" is x y sub, not y x sub
    prog = _parse_program(out[[t for t, _ in EXPR_CASES].index("x - y - z") * len(pts)].split("\t")[1])
    assert [_CODE_NAME[c] for c, _ in prog] == ["x", "y", "sub", "z", "sub"]


@pytest.mark.gpu
def test_device_function_init_matches_python(tmp_path):
    """tsp_init_expr runs the postfix programs of every expression at the cell
    centres x = lo + (i + 0.5) dx of a ragged box whose first cell has index
    (3, -2, 5) (SingleLevelIntegrator.hpp:75-93): bit for bit where the program
    uses only IEEE operations and exact functions, within 1e-13 relative
    (device libm) elsewhere; malformed programs are rejected before launch."""
    from treesap_amd import Solver
    from treesap_amd.workloads import small_case
    texts = [t for t, _ in EXPR_CASES] + ["x", "1.5", "y*z - x/3", "sqrt(x*x + y*y + z*z)"]
    _, out = _expr_programs(tmp_path, texts, [(0.1, 0.2, 0.3)])
    progs = [_parse_program(o.split("\t")[1]) for o in out]
    n, lo, hi, index_lo = (9, 7, 6), (-0.93, -0.81, -0.72), (0.87, 0.74, 0.79), (3, -2, 5)  # no centre at 0
    dx = [(hi[a] - lo[a]) / n[a] for a in range(3)]
    xs = [[lo[a] + (index_lo[a] + i + 0.5) * dx[a] for i in range(n[a])] for a in range(3)]
    case, _ = small_case("phipi", n=n, stencil_n=1, ko_order=0, periodic=(1, 1, 1), seed=3)
    case = case.replace(prob_lo=lo, prob_hi=hi)
    s = Solver(case)
    lib = _lib.lib()
    lo3 = (C.c_int32 * 3)(*index_lo)
    for t, prog in zip(texts, progs):
        # component 0: the expression; component 1: its negation
        ops_py = prog + prog + [(_lib.EXPR_CODES["neg"], 0.0)]
        ops = (_lib.TspExprOp * len(ops_py))(*[_lib.TspExprOp(c, 0, v) for c, v in ops_py])
        start = (C.c_int32 * 3)(0, len(prog), len(ops_py))
        _lib.check(lib.tsp_init_expr(s.handle, ops, start, lo3), "tsp_init_expr")
        got = s.get_state()
        want = np.empty(case.shape)
        for k, z in enumerate(xs[2]):
            for j, y in enumerate(xs[1]):
                want[0, k, j] = [run_postfix(prog, x, y, z) for x in xs[0]]
        want[1] = -want[0]
        if all(c in _EXACT_CODES for c, _ in prog):
            assert np.array_equal(got, want, equal_nan=True), t
        else:
            np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14 * np.nanmax(np.abs(want)), err_msg=t)
    assert s.time == 0.0 and s.steps == 0
    bad = {"underflow": [(_lib.EXPR_CODES["add"], 0.0)],
           "two values left": [(_lib.EXPR_CODES["x"], 0.0), (_lib.EXPR_CODES["y"], 0.0)],
           "unknown code": [(99, 0.0)],
           "too deep": [(_lib.EXPR_CODES["x"], 0.0)] * (_lib.TSP_EXPR_MAX_STACK + 1)}
    for why, prog in bad.items():
        ops = (_lib.TspExprOp * len(prog))(*[_lib.TspExprOp(c, 0, v) for c, v in prog])
        start = (C.c_int32 * 3)(0, len(prog), len(prog) + 0)
        assert lib.tsp_init_expr(s.handle, ops, start, lo3) != 0, why


# decorated registered right-hand sides, same ids as eval_case() in dropin_host.cpp
EVAL_CASES = {
    "phipi_n2_g3_ko3_abl4": dict(system="phipi", stencil_n=2, ghost=3, ko_order=3, ko_strength=0.1,
                                 abl_thickness=4, abl_max=2.0),
    "phipi_n4_g4_ko4_abl5": dict(system="phipi", stencil_n=4, ghost=4, ko_order=4, ko_strength=0.2,
                                 abl_thickness=5, abl_max=2.0),
    "phiu_n1_g1_ko1": dict(system="phiu", stencil_n=1, ghost=1, ko_order=1, ko_strength=0.05),
    "phiu_n3_g3_abl3": dict(system="phiu", stencil_n=3, ghost=3, abl_thickness=3, abl_max=1.5),
    "nonlinear_n2_g2_ko2": dict(system="nonlinear", stencil_n=2, ghost=2, ko_order=2, ko_strength=0.1, gamma=0.3),
    "tensor_n2_g2_ko2_abl3": dict(system="tensor", stencil_n=2, ghost=2, ko_order=2, ko_strength=0.1,
                                  abl_thickness=3, abl_max=1.0),
    "tensor_n3_g3": dict(system="tensor", stencil_n=3, ghost=3),
}


@pytest.mark.parametrize("case_id", sorted(EVAL_CASES))
@pytest.mark.parametrize("geometry", ["periodic", "dirichlet", "mixed_noop"])
def test_registered_rhs_matches_oracle(tmp_path, case_id, geometry):
    """The C++ functors (stencils, aliases, contractions, KO and ABL decorators)
    evaluate bit for bit what the oracle evaluates -- the same arithmetic the
    device kernels are tested against."""
    n = (13, 11, 12)
    periodic, bc = {"periodic": ((1, 1, 1), "noop"), "dirichlet": ((0, 0, 0), "dirichlet"),
                    "mixed_noop": ((1, 0, 1), "noop")}[geometry]
    kw = dict(EVAL_CASES[case_id])
    if kw.get("abl_thickness", 0) and geometry == "periodic":
        pytest.skip("absorbing layers are used on non-periodic boxes")
    case = Case(n=n, periodic=periodic, bc=bc, dt=0.01, **kw)
    rng = np.random.default_rng(zlib.crc32(f"{case_id}/{geometry}".encode()))
    state = rng.uniform(-1.0, 1.0, size=case.shape)
    o = orc.OracleSolver(case)
    if bc == "noop":
        # NoOp leaves ghost memory as it is: give the oracle's register zero ghosts, like the program's
        reg = np.zeros(o.gshape)
        G = case.ghost
        reg[:, G:G + n[2], G:G + n[1], G:G + n[0]] = state
        o.set_register(0, reg)
    else:
        o.set_state(state)
    want = o.eval_rhs(0)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    state.astype(np.float64).tofile(fin)
    _run([cpp_build.binary("dropin_host"), "eval", case_id, *n, *periodic, 1 if bc == "dirichlet" else 0, fin, fout])
    got = np.fromfile(fout, dtype=np.float64).reshape(case.shape)
    assert np.array_equal(got, want), f"max abs diff {np.max(np.abs(got - want))}"


# ---- whole programs on the GPU ------------------------------------------------------

SCENARIOS = ["first_order", "abl_ko", "nonlinear_euler", "tensor", "level_euler"]
# boxes of the level each program's block size gives ((phi, Pi), (phi, u) and NonLinear models are
# split; the tensor system runs one box)
LEVEL_BOXES = {"first_order": 8, "abl_ko": 8, "nonlinear_euler": 6, "tensor": 1, "level_euler": 6}


def _read_desc(path: Path) -> _lib.TspDesc:
    raw = path.read_bytes()
    d = _lib.TspDesc()
    assert len(raw) == C.sizeof(d)
    C.memmove(C.addressof(d), raw, len(raw))
    return d


@pytest.fixture(scope="module")
def gpu_runs(tmp_path_factory):
    out = tmp_path_factory.mktemp("dropin")
    print(f"[dropin_gpu] scenarios {SCENARIOS} -> {out}", flush=True)
    _run([cpp_build.binary("dropin_gpu"), out, *SCENARIOS])
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("scenario", SCENARIOS)
def test_dropin_program_matches_oracle(gpu_runs, scenario):
    d = gpu_runs / scenario
    case = case_from_desc(_read_desc(d / "desc.bin"))
    kind, steps, time, boxes = (d / "meta.txt").read_text().split()
    steps = int(steps)
    assert int(boxes) == LEVEL_BOXES[scenario]
    init = np.fromfile(d / "init.bin", dtype=np.float64).reshape(case.shape)
    final = np.fromfile(d / "final.bin", dtype=np.float64).reshape(case.shape)
    o = orc.OracleSolver(case)
    o.set_state(init)
    (o.rk4_steps if kind == "rk4" else o.euler_steps)(steps)
    want = o.get_state()
    assert np.array_equal(final, want), f"{scenario}: max abs diff {np.max(np.abs(final - want))}"
    t = 0.0
    for _ in range(steps):
        t += case.dt
    assert float(time) == t
    assert np.isfinite(final).all() and np.abs(final).max() > 0


# The initial data of each program (tests/cpp/dropin_gpu.cpp), restated with
# Python's math: (box lo, box hi, cells per axis, one function per component).
def _tensor_init(c):
    a, b, e = c + 1, (c % 3) + 1, (c % 4) + 1
    return lambda x, y, z: 0.01 * math.sin(2 * math.pi * (a * x + b * y)) * math.cos(2 * math.pi * e * z)


INIT_EXPRS = {
    "first_order": ((-2.0,) * 3, (2.0,) * 3, (32, 32, 32),
                    [lambda x, y, z: math.exp(-(x ** 2 + y ** 2 + z ** 2) / 0.5 ** 2)] + [lambda x, y, z: 0.0] * 3),
    "abl_ko": ((0.0,) * 3, (1.0,) * 3, (48, 48, 48),
               [lambda x, y, z: 0.8 * math.exp(-((x - 0.45) ** 2 + (y - 0.55) ** 2 + (z - 0.5) ** 2) / 0.1 ** 2)
                - 0.5 * math.exp(-((x - 0.6) ** 2 + (y - 0.4) ** 2 + (z - 0.55) ** 2) / 0.07 ** 2),
                lambda x, y, z: 0.0]),
    "nonlinear_euler": ((0.0,) * 3, (1.0, 0.9, 0.8), (40, 36, 32),
                        [lambda x, y, z: math.sin(2 * math.pi * x) * math.cos(2 * math.pi * y / 0.9)
                         + 0.3 * math.sin(2 * math.pi * z / 0.8),
                         lambda x, y, z: 0.5 * math.cos(2 * math.pi * (x + y))]),
    "tensor": ((0.0,) * 3, (1.0,) * 3, (24, 24, 24), [_tensor_init(c) for c in range(12)]),
    "level_euler": ((0.0,) * 3, (1.0, 0.9, 0.8), (20, 18, 16),
                    [lambda x, y, z: math.sin(2 * math.pi * x) * math.cos(2 * math.pi * y / 0.9)
                     + 0.3 * math.sin(2 * math.pi * z / 0.8),
                     lambda x, y, z: 0.5 * math.cos(2 * math.pi * (x + y))]),
}


@pytest.mark.gpu
@pytest.mark.parametrize("scenario", SCENARIOS)
def test_dropin_initial_state_matches_cell_centres(gpu_runs, scenario):
    """initialise() evaluates the FunctionInit expressions at the cell centres
    x = lo + (i + 0.5) * dx with dx = (hi - lo) / n (SingleLevelIntegrator.hpp:75-93,
    Geometry::CellSize), on the device: the state equals Python's evaluation of
    the same expressions there to 1e-13 (device libm). A half-cell shift, a dx
    from n - 1 or a swapped axis fails it."""
    d = gpu_runs / scenario
    case = case_from_desc(_read_desc(d / "desc.bin"))
    lo, hi, n, fns = INIT_EXPRS[scenario]
    assert tuple(case.n) == n and tuple(case.prob_lo) == lo and tuple(case.prob_hi) == hi
    init = np.fromfile(d / "init.bin", dtype=np.float64).reshape(case.shape)
    dx = [(hi[a] - lo[a]) / n[a] for a in range(3)]
    xs = [[lo[a] + (i + 0.5) * dx[a] for i in range(n[a])] for a in range(3)]
    want = np.empty(case.shape)
    for c, fn in enumerate(fns):
        for k, z in enumerate(xs[2]):
            for j, y in enumerate(xs[1]):
                want[c, k, j] = [fn(x, y, z) for x in xs[0]]
    # FunctionInit runs on the device: exp, sin and cos may differ from the host's libm in the last bits
    np.testing.assert_allclose(init, want, rtol=1e-13, atol=1e-14 * np.max(np.abs(want)))
    # the check has teeth: a half-cell shift of x changes the state (on a row through the middle,
    # where the data are not below the tolerance)
    shifted = [lo[0] + i * dx[0] for i in range(n[0])]
    jm, km = n[1] // 2, n[2] // 2
    assert not np.allclose(init[0, km, jm], [fns[0](x, xs[1][jm], xs[2][km]) for x in shifted],
                           rtol=1e-13, atol=1e-14 * np.max(np.abs(want)))


_FAB_HEAD = re.compile(rb"^FAB \(\(8, \(64 11 52 0 1 12 0 1023\)\),\(8, \(8 7 6 5 4 3 2 1\)\)\)"
                       rb"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \(0,0,0\)\) (\d+)\n")
_BOX_TEXT = re.compile(r"^\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \(0,0,0\)\)$")


def read_plotfile(path: Path):
    """Reads a single-level plotfile by AMReX's documented layout (HyperCLaw-V1.1
    Header, Level_0/Cell_H FabArray header, FABs at the offsets Cell_H gives) and
    checks its internal consistency; returns (names, time, step, boxes, dense state
    a[c, k, j, i] over the domain)."""
    hdr = (path / "Header").read_text().split("\n")
    assert hdr[0] == "HyperCLaw-V1.1"
    nc = int(hdr[1])
    names = hdr[2:2 + nc]
    q = 2 + nc
    assert hdr[q] == "3" and hdr[q + 2] == "0"  # dimensions, finest level
    time = float(hdr[q + 1])
    prob_lo = [float(v) for v in hdr[q + 3].split()]
    prob_hi = [float(v) for v in hdr[q + 4].split()]
    assert hdr[q + 5] == ""  # no refinement ratios on one level
    dom = [int(v) for v in _BOX_TEXT.match(hdr[q + 6].strip()).groups()]
    step = int(hdr[q + 7].split()[0])
    dx = [float(v) for v in hdr[q + 8].split()]
    assert hdr[q + 9] == "0" and hdr[q + 10] == "0"  # Cartesian, no boundary data
    lev, nbox, ltime = hdr[q + 11].split()
    assert lev == "0" and float(ltime) == time and int(hdr[q + 12]) == step
    nbox = int(nbox)
    extents = [[float(v) for v in hdr[q + 13 + 3 * b + d].split()] for b in range(nbox) for d in range(3)]
    assert hdr[q + 13 + 3 * nbox] == "Level_0/Cell"
    n = [dom[3 + d] - dom[d] + 1 for d in range(3)]
    ch = (path / "Level_0" / "Cell_H").read_text().split("\n")
    assert ch[0:4] == ["1", "1", str(nc), "0"]  # version, NFiles, components, no ghosts
    assert ch[4] == f"({nbox} 0"
    boxes = [[int(v) for v in _BOX_TEXT.match(ch[5 + b]).groups()] for b in range(nbox)]
    assert ch[5 + nbox] == ")" and int(ch[6 + nbox]) == nbox
    fod = [ch[7 + nbox + b].split() for b in range(nbox)]
    assert all(f[0] == "FabOnDisk:" and f[1] == "Cell_D_00000" for f in fod) and ch[7 + 2 * nbox] == ""
    r = 8 + 2 * nbox
    assert ch[r] == f"{nbox},{nc}"
    mins = [[float(v) for v in ch[r + 1 + b].rstrip(",").split(",")] for b in range(nbox)]
    assert ch[r + 1 + nbox] == "" and ch[r + 2 + nbox] == f"{nbox},{nc}"
    maxs = [[float(v) for v in ch[r + 3 + nbox + b].rstrip(",").split(",")] for b in range(nbox)]
    raw = (path / "Level_0" / "Cell_D_00000").read_bytes()
    dense = np.full((nc, n[2], n[1], n[0]), np.nan)
    covered = np.zeros((n[2], n[1], n[0]), dtype=int)
    pos = 0
    for b, box in enumerate(boxes):
        off = int(fod[b][2])
        assert off == pos, "FABs follow each other in box order"
        m = _FAB_HEAD.match(raw[off:off + 200])
        assert m and [int(v) for v in m.groups()[:6]] == box and int(m.group(7)) == nc
        bn = [box[3 + d] - box[d] + 1 for d in range(3)]
        cnt = nc * bn[0] * bn[1] * bn[2]
        data = np.frombuffer(raw, dtype="<f8", count=cnt, offset=off + m.end()).reshape(nc, bn[2], bn[1], bn[0])
        pos = off + m.end() + 8 * cnt
        sl = tuple(slice(box[d] - dom[d], box[3 + d] - dom[d] + 1) for d in (2, 1, 0))
        dense[(slice(None),) + sl] = data
        covered[sl] += 1
        assert [float(v) for v in data.reshape(nc, -1).min(axis=1)] == mins[b]
        assert [float(v) for v in data.reshape(nc, -1).max(axis=1)] == maxs[b]
        for d in range(3):  # physical extent of the box in the Header
            assert extents[3 * b + d] == [prob_lo[d] + box[d] * dx[d], prob_lo[d] + (box[3 + d] + 1) * dx[d]]
    assert pos == len(raw) and np.all(covered == 1), "the boxes tile the domain exactly once"
    assert [dx[d] == (prob_hi[d] - prob_lo[d]) / n[d] for d in range(3)] == [True] * 3
    return names, time, step, boxes, dense


@pytest.mark.gpu
def test_dropin_plotfiles(gpu_runs):
    """configureExporter(path, k) writes the state before every k-th step: on a
    level of 2 x 2 x 2 boxes (the first-order wave, exports at steps 0, 4, 8 of 12)
    and on one of 3 x 2 x 1 boxes (block size 7 x 9 x 64 on 20 x 18 x 16 cells,
    exports at 0 and 5 of 6), one FAB per box. Every file is read back by the documented layout and
    its payload equals the oracle's state at that step bit for bit."""
    for scenario, names_want, every, steps, kind in (("first_order", ["phi", "u_0", "u_1", "u_2"], 4, 12, "rk4"),
                                                     ("level_euler", ["phi", "Pi"], 5, 6, "euler")):
        d = gpu_runs / scenario
        case = case_from_desc(_read_desc(d / "desc.bin"))
        init = np.fromfile(d / "init.bin", dtype=np.float64).reshape(case.shape)
        names = sorted(p.name for p in d.iterdir() if p.name.startswith("plt"))
        assert names == [f"plt{s:05d}" for s in range(0, steps, every)], names
        o = orc.OracleSolver(case)
        o.set_state(init)
        for s in range(0, steps, every):
            if s:
                o.rk4_steps(every) if kind == "rk4" else o.euler_steps(every)
            got_names, t, step, boxes, dense = read_plotfile(d / f"plt{s:05d}")
            assert got_names == names_want and step == s and t == o.time
            assert np.array_equal(dense, o.get_state()), (scenario, s)
        if scenario == "level_euler":
            assert len(boxes) == 6
            assert sorted({(b[0], b[3]) for b in boxes}) == [(0, 6), (7, 13), (14, 19)]
            assert sorted({(b[1], b[4]) for b in boxes}) == [(0, 8), (9, 17)]


_TIMER_LINE = re.compile(r"^\[Timer: (.+?)\] Run: ([0-9.]+) (µs|ms|s) \| Total: ([0-9.]+) (µs|ms|s) \| "
                         r"Avg: ([0-9.]+) (µs|ms|s) \| StdDev: ([0-9.]+) (µs|ms|s) \((\d+) runs\)$")
_OVERVIEW_LINE = re.compile(r"^\[Timer: (.{35,}?)\s*\] Executions: +(\d+) \| Total: +([0-9.]+) (µs|ms|s) \| "
                            r"Avg: +([0-9.]+) (µs|ms|s) \| StdDev: +([0-9.]+) (µs|ms|s)$")
_UNIT_US = {"µs": 1.0, "ms": 1e3, "s": 1e6}


@pytest.mark.gpu
def test_scoped_timer_times_the_device(tmp_path):
    """ScopedTimer (util/ScopedTimer.hpp:34-149) prints the reference's lines, and its
    time is the device stream's: a scope that enqueues 12 steps behind 120 queued ones
    reports about 12/132 of the host time around both, where a host clock (which
    would wait for the queue) reports all of it."""
    res = _run([cpp_build.binary("dropin_gpu"), tmp_path, "timer"])
    lines = res.stdout.splitlines()
    host_us = float(next(l for l in lines if l.startswith("host_us")).split()[1])
    runs = {}
    for l in lines:
        m = _TIMER_LINE.match(l)
        if m:
            runs.setdefault(m.group(1), []).append(float(m.group(2)) * _UNIT_US[m.group(3)])
    scope = runs["scope of 12 steps"]
    assert len(scope) == 1
    frac = scope[0] / host_us
    print(f"scope {scope[0]:.1f} us of host {host_us:.1f} us ({frac:.3f})", flush=True)
    assert 0.02 < frac < 0.35, frac
    assert len(runs["Integrator::step() [STEP]"]) == 3 and len(runs["Integrator::step() [PRE]"]) == 3
    assert len(runs["Integrator::advance()"]) == 1
    # one RK4 step's device time against the 12-step scope's
    assert 0.3 < min(runs["Integrator::step() [STEP]"]) * 12 / scope[0] < 3.0
    start = lines.index("--- --- --- ScopedTimer Profiling Overview --- --- ---")
    over = [_OVERVIEW_LINE.match(l) for l in lines[start + 1:] if l.startswith("[Timer:")]
    assert all(over) and len(over) == len(runs)
    totals = [float(m.group(3)) * _UNIT_US[m.group(4)] for m in over]
    assert totals == sorted(totals, reverse=True)


# ---- functors the library does not carry: the generic kernel (tests/cpp/dropin_generic.cpp) ----

GENERIC_SCENARIOS = ["first_order", "abl_tensor", "nonlinear_euler", "curl_rk4", "curl_euler", "patched"]
_Z = (0.0,) * 12
# scenarios whose functors restate a system the oracle carries (same arithmetic)
GENERIC_ORACLE = {
    "first_order": Case(system="phiu", n=(24, 20, 22), stencil_n=2, ghost=2, ko_order=2, ko_strength=0.1,
                        periodic=(0, 0, 0), bc="dirichlet", target=(0.1, -0.2, 0.3, 0.05) + _Z[4:], dt=0.25 / 24),
    "abl_tensor": Case(system="tensor", n=(20, 18, 16), prob_hi=(1.0, 0.9, 0.8), stencil_n=2, ghost=2, ko_order=2,
                       ko_strength=0.1, periodic=(0, 0, 0), bc="dirichlet", abl_thickness=3, abl_max=1.5,
                       abl_profile="quadratic", dt=0.2 / 20),
    "nonlinear_euler": Case(system="nonlinear", n=(26, 22, 18), prob_hi=(1.0, 0.9, 0.8), stencil_n=2, ghost=2,
                            ko_order=2, ko_strength=0.1, periodic=(1, 1, 1), bc="noop", gamma=0.05, dt=0.004),
}


@pytest.fixture(scope="module")
def generic_runs(tmp_path_factory):
    out = tmp_path_factory.mktemp("generic")
    print(f"[dropin_generic] scenarios {GENERIC_SCENARIOS} -> {out}", flush=True)
    _run([cpp_build.binary("dropin_generic"), out, *GENERIC_SCENARIOS])
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("scenario", GENERIC_SCENARIOS)
def test_generic_kernel_program(generic_runs, scenario):
    """Unmapped right-hand sides run on the stage kernel compiled from the functor
    (descriptor system TSP_SYS_CUSTOM). Restatements of the oracle's systems match
    the oracle bitwise; functors with third derivatives, a curl, antisymmetric and
    cyclic parts, position and time dependence, or a source confined to half the
    box match the same functor stepped on the host (tests/cpp/host_rk4.hpp) bitwise."""
    d = generic_runs / scenario
    kind, steps, time, system = (d / "meta.txt").read_text().split()
    assert int(system) == 4, "the generic kernel did not run"
    steps = int(steps)
    final = np.fromfile(d / "final.bin", dtype=np.float64)
    init = np.fromfile(d / "init.bin", dtype=np.float64)
    if scenario in GENERIC_ORACLE:
        case = GENERIC_ORACLE[scenario]
        o = orc.OracleSolver(case)
        o.set_state(init.reshape(case.shape))
        (o.rk4_steps if kind == "rk4" else o.euler_steps)(steps)
        want = o.get_state().ravel()
        t = 0.0
        for _ in range(steps):
            t += case.dt
        assert float(time) == t
    else:
        want = np.fromfile(d / "host.bin", dtype=np.float64)
    assert np.isfinite(final).all() and not np.array_equal(final, init)
    assert np.array_equal(final, want), f"{scenario}: max abs diff {np.max(np.abs(final - want))}"


# ---- boundary functors the library does not carry (tests/cpp/dropin_bc.cpp) ----

BC_SCENARIOS = {"library_rk4": 0, "library_euler": 0, "tensor": 3, "generic_shifted": 4}  # -> descriptor system


@pytest.fixture(scope="module")
def bc_runs(tmp_path_factory):
    out = tmp_path_factory.mktemp("custom_bc")
    print(f"[dropin_bc] scenarios {list(BC_SCENARIOS)} -> {out}", flush=True)
    _run([cpp_build.binary("dropin_bc"), out, *BC_SCENARIOS])
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("scenario", list(BC_SCENARIOS))
def test_custom_boundary_program(bc_runs, scenario):
    """A boundary functor other than NoOpBC and DirichletBC (time- and position-dependent
    values, copies of the nearest valid cell scaled by the model's BCRec) runs on the kernel
    compiled from it after each fill's periodic step (descriptor bc TSP_BC_CUSTOM), under the
    library's (phi, Pi) and tensor kernels (RK4 with autotuning and advance(), Euler) and the
    generic kernel on a domain whose first cell is not (0, 0, 0) with an absorbing layer. The
    result matches the same functors stepped on the host (tests/cpp/host_rk4.hpp) bitwise, and
    differs from the NoOp boundary's. Two further steps through tsp_rk4_steps_noalloc under the
    same boundary callback match the host too; graph replay and ghost_mode 0 are refused."""
    d = bc_runs / scenario
    kind, steps, time, system, bc = (d / "meta.txt").read_text().split()
    assert int(bc) == 2, "the custom boundary path did not run"
    assert int(system) == BC_SCENARIOS[scenario]
    final = np.fromfile(d / "final.bin", dtype=np.float64)
    init = np.fromfile(d / "init.bin", dtype=np.float64)
    want = np.fromfile(d / "host.bin", dtype=np.float64)
    noop = np.fromfile(d / "noop.bin", dtype=np.float64)
    assert np.isfinite(final).all() and not np.array_equal(final, init)
    assert not np.array_equal(want, noop), "the boundary functor does not affect this scenario"
    assert np.array_equal(final, want), f"{scenario}: max abs diff {np.max(np.abs(final - want))}"
    if scenario == "library_rk4":
        na = np.fromfile(d / "noalloc.bin", dtype=np.float64)
        na_host = np.fromfile(d / "noalloc_host.bin", dtype=np.float64)
        assert not np.array_equal(na, final)
        assert np.array_equal(na, na_host), f"noalloc: max abs diff {np.max(np.abs(na - na_host))}"
        graph_status, mode_status = map(int, (d / "refusals.txt").read_text().split())
        assert graph_status == _lib.UNSUPPORTED and mode_status == _lib.INVALID_ARG
