"""C-ABI checks that need no GPU: the library loads, exports every function the
header declares, the ctypes mirrors of the structs match the C layout, and
argument validation (which runs before any HIP call) returns the documented
status codes."""
import ctypes as C
import subprocess
import textwrap
from pathlib import Path

import pytest

from treesap_amd import _lib
from treesap_amd.workloads import Case, make_desc

ROOT = Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    # and the binding covers every declared symbol
    assert set(declared) == set(_lib.PROTOTYPES), set(declared) ^ set(_lib.PROTOTYPES)


def test_struct_layout_matches_c(tmp_path):
    fields = [f for f, _ in _lib.TspDesc._fields_]
    prog = textwrap.dedent("""
        #include <stdio.h>
        #include <stddef.h>
        #include "treesap_hip.h"
        int main(void) {
            printf("desc %zu\\n", sizeof(tsp_desc));
            printf("cfg %zu\\n", sizeof(tsp_kernel_cfg));
            printf("bcfg %zu\\n", sizeof(tsp_batch_cfg));
            printf("expr %zu\\n", sizeof(tsp_expr_op));
            printf("expr_value %zu\\n", offsetof(tsp_expr_op, value));
            printf("heaviside %d\\n", (int)TSP_EXPR_HEAVISIDE);
            printf("maxstack %d\\n", (int)TSP_EXPR_MAX_STACK);
    """)
    for f in fields:
        prog += f'    printf("{f} %zu\\n", offsetof(tsp_desc, {f}));\n'
    prog += "    return 0;\n}\n"
    src = tmp_path / "layout.c"
    src.write_text(prog)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    assert int(out["desc"]) == C.sizeof(_lib.TspDesc)
    assert int(out["cfg"]) == C.sizeof(_lib.TspKernelCfg)
    assert int(out["bcfg"]) == C.sizeof(_lib.TspBatchCfg)
    assert int(out["expr"]) == C.sizeof(_lib.TspExprOp) and int(out["expr_value"]) == _lib.TspExprOp.value.offset
    assert int(out["heaviside"]) == _lib.EXPR_CODES["heaviside"] and int(out["maxstack"]) == _lib.TSP_EXPR_MAX_STACK
    for f in fields:
        assert int(out[f]) == getattr(_lib.TspDesc, f).offset, f


def test_version_and_defaults():
    L = _lib.lib()
    assert L.tsp_version().decode() == "0.1.0"
    d = _lib.TspDesc()
    assert L.tsp_desc_default(C.byref(d), 0, 32) == _lib.OK
    assert list(d.n) == [32, 32, 32] and list(d.periodic) == [1, 1, 1]
    assert d.dt == 0.25 / 32
    assert L.tsp_desc_default(C.byref(d), 9, 32) == _lib.INVALID_ARG
    assert L.tsp_rk4_bytes_per_cell(C.byref(d)) == 128 * 2


def _create_status(case: Case):
    d = make_desc(case)
    h = C.c_void_p()
    st = _lib.lib().tsp_create(C.byref(d), None, C.byref(h))
    msg = _lib.lib().tsp_last_error().decode()
    assert not h.value or st == _lib.OK
    return st, msg


GOOD = Case(system="phipi", n=(16, 16, 16), stencil_n=2, ghost=2, dt=0.01)


@pytest.mark.parametrize("bad,status,needle", [
    (dict(system="nope"), None, None),
    (dict(n=(0, 16, 16)), _lib.INVALID_ARG, "n[0]"),
    (dict(stencil_n=0), _lib.INVALID_ARG, "stencil_n"),
    (dict(stencil_n=5, ghost=5), _lib.INVALID_ARG, "stencil_n"),
    (dict(ghost=1), _lib.INVALID_ARG, "ghost"),
    (dict(ko_order=3, ghost=2), _lib.INVALID_ARG, "ghost"),
    (dict(ko_order=1, ghost=2), _lib.UNSUPPORTED, "no stage kernel"),
    (dict(periodic=(0, 1, 2)), _lib.INVALID_ARG, "periodic"),
    (dict(n=(1, 16, 16)), _lib.INVALID_ARG, "periodic axis"),
    (dict(periodic=(0, 0, 0), bc="dirichlet", abl_thickness=8), _lib.INVALID_ARG, "bulk"),
    (dict(prob_hi=(0.0, 1.0, 1.0)), _lib.INVALID_ARG, "prob_hi"),
    (dict(dt=float("nan")), _lib.INVALID_ARG, "dt"),
    (dict(system="tensor", ko_order=3, ghost=3), _lib.UNSUPPORTED, "ko_order in {0, n}"),
])
def test_create_validation(bad, status, needle):
    if status is None:
        with pytest.raises(KeyError):
            make_desc(GOOD.replace(**bad))
        return
    st, msg = _create_status(GOOD.replace(**bad))
    assert st == status, msg
    assert needle in msg


def test_null_handles_are_rejected():
    L = _lib.lib()
    assert L.tsp_rk4_steps(None, 1) == _lib.INVALID_ARG
    assert L.tsp_sync(None) == _lib.INVALID_ARG
    assert L.tsp_destroy(None) == _lib.OK
    assert L.tsp_create(None, None, None) == _lib.INVALID_ARG
    assert L.tsp_init_expr(None, None, None, None) == _lib.INVALID_ARG
    assert L.tsp_level_init_expr(None, None, None, None) == _lib.INVALID_ARG
    assert L.tsp_get_stream(None, None) == _lib.INVALID_ARG
    assert L.tsp_batch_get_stream(None, None) == _lib.INVALID_ARG
    assert L.tsp_timer_create(None, None) == _lib.INVALID_ARG
    assert L.tsp_timer_start(None) == _lib.INVALID_ARG
    assert L.tsp_timer_stop(None, None) == _lib.INVALID_ARG
    assert L.tsp_timer_destroy(None) == _lib.OK
    assert L.tsp_rk4_steps_noalloc(None, None, None, 1) == _lib.INVALID_ARG


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setenv("TSP_LIBRARY", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.lib()


def test_batch_validation():
    from treesap_amd import Batch

    good = Case(system="phipi", n=(16, 16, 16), stencil_n=2, ghost=2, ko_order=2, dt=0.01)
    with pytest.raises(_lib.TspError, match="batched mode runs the"):
        Batch([good, good.replace(system="tensor", stencil_n=2, ghost=2, ko_order=2)])
    with pytest.raises(_lib.TspError, match="no batched kernel"):
        Batch([good.replace(stencil_n=3, ghost=4, ko_order=4)])
    with pytest.raises(_lib.TspError, match="instance 1: .*stencil_n"):
        Batch([good, good.replace(stencil_n=0)])
    # 700^3 with two components: 5.9 GB per register, valid for one solver but
    # beyond the batched kernels' 4 GiB buffer window
    with pytest.raises(_lib.TspError, match="instance 1: batched instances need registers under 4 GiB"):
        Batch([good, good.replace(n=(700, 700, 700))])
    L = _lib.lib()
    assert L.tsp_batch_create(None, 0, None, None) == _lib.INVALID_ARG
    assert L.tsp_batch_rk4_steps(None, 1) == _lib.INVALID_ARG
    assert L.tsp_batch_destroy(None) == _lib.OK


def test_custom_model_validation():
    """A TSP_SYS_CUSTOM descriptor (generic kernel) states its component count and carries no
    KO or ABL decorator of the library's (they belong to the caller's functor); the library's
    stage kernels refuse it."""
    L = _lib.lib()

    def status(**kw):
        d = _lib.TspDesc()
        assert L.tsp_desc_default(C.byref(d), 4, 16) == _lib.OK
        assert d.system == 4 and d.ncomp == 1
        for k, v in kw.items():
            setattr(d, k, v)
        h = C.c_void_p()
        st = L.tsp_create(C.byref(d), None, C.byref(h))
        return st, L.tsp_last_error().decode()

    st, msg = status(ncomp=0)
    assert st == _lib.INVALID_ARG and "ncomp in 1..12" in msg
    st, msg = status(ncomp=13)
    assert st == _lib.INVALID_ARG and "ncomp in 1..12" in msg
    st, msg = status(ncomp=3, ko_order=2)
    assert st == _lib.INVALID_ARG and "decorators are part of its functor" in msg
    st, msg = status(ncomp=3, abl_thickness=2)
    assert st == _lib.INVALID_ARG and "decorators are part of its functor" in msg
    assert L.tsp_rk4_steps_custom(None, 1, None, None) == _lib.INVALID_ARG


def test_level_validation():
    """tsp_level_create rejects bad block sizes and the tensor system before any HIP call."""
    L = _lib.lib()
    h = C.c_void_p()
    d = make_desc(Case(system="phipi", n=(16, 16, 16), stencil_n=2, ghost=2, ko_order=2, dt=0.01))
    assert L.tsp_level_create(C.byref(d), (C.c_int32 * 3)(8, 0, 8), None, C.byref(h)) == _lib.INVALID_ARG
    assert L.tsp_level_create(C.byref(d), None, None, C.byref(h)) == _lib.INVALID_ARG
    d2 = make_desc(Case(system="tensor", n=(16, 16, 16), stencil_n=2, ghost=2, ko_order=2, dt=0.01))
    assert L.tsp_level_create(C.byref(d2), (C.c_int32 * 3)(8, 8, 8), None, C.byref(h)) == _lib.UNSUPPORTED
    assert "(phi,Pi), (phi,u) and NonLinear" in L.tsp_last_error().decode()


def test_custom_boundary_validation():
    """A boundary functor the library does not carry (TSP_BC_CUSTOM) is a valid descriptor, but
    levels and batches, whose exchange kernels carry NoOp and Dirichlet boundaries only, refuse
    it before any HIP call; an unknown bc id stays invalid."""
    L = _lib.lib()
    h = C.c_void_p()
    d = make_desc(Case(system="phipi", n=(16, 16, 16), stencil_n=2, ghost=2, ko_order=2, dt=0.01))
    d.bc = 2
    assert L.tsp_level_create(C.byref(d), (C.c_int32 * 3)(8, 8, 8), None, C.byref(h)) == _lib.UNSUPPORTED
    assert "NoOp and Dirichlet boundaries only" in L.tsp_last_error().decode()
    assert L.tsp_batch_create(C.byref(d), 1, None, C.byref(h)) == _lib.UNSUPPORTED
    assert "NoOp and Dirichlet boundaries only" in L.tsp_last_error().decode()
    d.bc = 3
    assert L.tsp_level_create(C.byref(d), (C.c_int32 * 3)(8, 8, 8), None, C.byref(h)) == _lib.INVALID_ARG
    assert "unknown bc id" in L.tsp_last_error().decode()
    assert L.tsp_set_bc_callback(None, None, None) == _lib.INVALID_ARG
