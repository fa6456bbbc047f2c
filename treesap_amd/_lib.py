"""ctypes binding of the C ABI declared in include/treesap_hip.h.

The library ``libtreesap_hip.so`` is loaded from this package directory. There
is no fallback: if the library is missing or cannot be loaded, every call
raises, so a GPU run can never silently fall back to another implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libtreesap_hip.so"
HEADER = PKG.parent / "include" / "treesap_hip.h"

TSP_MAX_COMP = 12

# tsp_status
OK, INVALID_ARG, OOM, HIP_ERROR, UNSUPPORTED = range(5)
STATUS_NAMES = {0: "TSP_OK", 1: "TSP_INVALID_ARG", 2: "TSP_OOM", 3: "TSP_HIP_ERROR", 4: "TSP_UNSUPPORTED"}

# tsp_system
SYSTEMS = {"phipi": 0, "phiu": 1, "nonlinear": 2, "tensor": 3}
NCOMP = {"phipi": 2, "phiu": 4, "nonlinear": 2, "tensor": 12}
# tsp_bc, tsp_profile
BCS = {"noop": 0, "dirichlet": 1}
PROFILES = {"linear": 0, "quadratic": 1, "cubic": 2}
# tsp_role
ROLE_OLD, ROLE_NEW, ROLE_ARG, ROLE_ARG_NEXT = range(4)


class TspDesc(C.Structure):
    _fields_ = [
        ("system", C.c_int32),
        ("n", C.c_int32 * 3),
        ("stencil_n", C.c_int32),
        ("ghost", C.c_int32),
        ("ko_order", C.c_int32),
        ("bc", C.c_int32),
        ("ko_strength", C.c_double),
        ("prob_lo", C.c_double * 3),
        ("prob_hi", C.c_double * 3),
        ("periodic", C.c_int32 * 3),
        ("abl_thickness", C.c_int32),
        ("abl_profile", C.c_int32),
        ("ncomp", C.c_int32),
        ("abl_max", C.c_double),
        ("target", C.c_double * TSP_MAX_COMP),
        ("gamma", C.c_double),
        ("dt", C.c_double),
    ]


class TspKernelCfg(C.Structure):
    _fields_ = [
        ("ghost_mode", C.c_int32),
        ("tile_y", C.c_int32),
        ("zchunk", C.c_int32),
        ("xcd_remap", C.c_int32),
        ("zalt", C.c_int32),
        ("halo_ring", C.c_int32),
        ("stage_pairs", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class TspBatchCfg(C.Structure):
    _fields_ = [
        ("tile_y", C.c_int32),
        ("zchunk", C.c_int32),
        ("grid_blocks", C.c_int32),
        ("zalt", C.c_int32),
        ("halo_ring", C.c_int32),
        ("narrow_tiles", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


# tsp_expr_code: postfix programs of initial-data expressions (tsp_init_expr)
EXPR_CODES = {name: i for i, name in enumerate(
    ["num", "x", "y", "z", "add", "sub", "mul", "div", "pow", "atan2", "min", "max", "neg", "sqrt", "exp",
     "log", "log10", "sin", "cos", "tan", "asin", "acos", "atan", "sinh", "cosh", "tanh", "abs", "floor",
     "ceil", "heaviside"])}
TSP_EXPR_MAX_STACK = 16


class TspExprOp(C.Structure):
    _fields_ = [("code", C.c_int32), ("reserved", C.c_int32), ("value", C.c_double)]


class TspError(RuntimeError):
    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_VP = C.c_void_p
_DP = C.POINTER(C.c_double)
_I64P = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
PROTOTYPES = {
    "tsp_last_error": (C.c_char_p, []),
    "tsp_version": (C.c_char_p, []),
    "tsp_desc_default": (C.c_int, [C.POINTER(TspDesc), C.c_int32, C.c_int32]),
    "tsp_create": (C.c_int, [C.POINTER(TspDesc), _VP, C.POINTER(_VP)]),
    "tsp_destroy": (C.c_int, [_VP]),
    "tsp_set_state": (C.c_int, [_VP, _VP, C.c_int32]),
    "tsp_get_state": (C.c_int, [_VP, _VP, C.c_int32]),
    "tsp_init_expr": (C.c_int, [_VP, _VP, _VP, _VP]),
    "tsp_fill_ghosts": (C.c_int, [_VP, C.c_int32]),
    "tsp_get_register": (C.c_int, [_VP, C.c_int32, _VP]),
    "tsp_rk4_steps": (C.c_int, [_VP, C.c_int64]),
    "tsp_rk4_steps_noalloc": (C.c_int, [_VP, _VP, _VP, C.c_int64]),
    "tsp_euler_steps": (C.c_int, [_VP, C.c_int64]),
    "tsp_get_time": (C.c_int, [_VP, _DP, _I64P]),
    "tsp_get_kernel": (C.c_int, [_VP, C.POINTER(TspKernelCfg)]),
    "tsp_set_kernel": (C.c_int, [_VP, C.POINTER(TspKernelCfg)]),
    "tsp_autotune": (C.c_int, [_VP, C.c_int32, C.POINTER(TspKernelCfg)]),
    "tsp_graph_rk4_steps": (C.c_int, [_VP, C.c_int64]),
    "tsp_graph_prepare": (C.c_int, [_VP]),
    "tsp_profile_rk4": (C.c_int, [_VP, C.c_int64, _DP, _DP]),
    "tsp_time_rk4": (C.c_int, [_VP, C.c_int64, C.c_int32, _DP]),
    "tsp_sync": (C.c_int, [_VP]),
    "tsp_get_stream": (C.c_int, [_VP, C.POINTER(_VP)]),
    "tsp_batch_get_stream": (C.c_int, [_VP, C.POINTER(_VP)]),
    "tsp_timer_create": (C.c_int, [_VP, C.POINTER(_VP)]),
    "tsp_timer_start": (C.c_int, [_VP]),
    "tsp_timer_stop": (C.c_int, [_VP, C.POINTER(C.c_double)]),
    "tsp_timer_destroy": (C.c_int, [_VP]),
    "tsp_register_view": (C.c_int, [_VP, C.c_int32, C.POINTER(_VP), _I64P, _I64P, _I64P]),
    "tsp_rk4_bytes_per_cell": (C.c_int64, [C.POINTER(TspDesc)]),
    "tsp_batch_create": (C.c_int, [C.POINTER(TspDesc), C.c_int32, _VP, C.POINTER(_VP)]),
    "tsp_batch_destroy": (C.c_int, [_VP]),
    "tsp_batch_size": (C.c_int, [_VP, C.POINTER(C.c_int32), _I64P]),
    "tsp_batch_set_state": (C.c_int, [_VP, C.c_int32, _VP, C.c_int32]),
    "tsp_batch_get_state": (C.c_int, [_VP, C.c_int32, _VP, C.c_int32]),
    "tsp_batch_get_time": (C.c_int, [_VP, C.c_int32, _DP, _I64P]),
    "tsp_batch_get_kernel": (C.c_int, [_VP, C.POINTER(TspBatchCfg)]),
    "tsp_batch_set_kernel": (C.c_int, [_VP, C.POINTER(TspBatchCfg)]),
    "tsp_batch_get_group_kernel": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, C.POINTER(TspBatchCfg)]),
    "tsp_batch_set_group_kernel": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, C.POINTER(TspBatchCfg)]),
    "tsp_batch_rk4_steps": (C.c_int, [_VP, C.c_int64]),
    "tsp_batch_graph_rk4_steps": (C.c_int, [_VP, C.c_int64]),
    "tsp_batch_graph_prepare": (C.c_int, [_VP]),
    "tsp_batch_euler_steps": (C.c_int, [_VP, C.c_int64]),
    "tsp_batch_autotune": (C.c_int, [_VP, C.c_int32, C.POINTER(TspBatchCfg)]),
    "tsp_level_create": (C.c_int, [C.POINTER(TspDesc), C.POINTER(C.c_int32), _VP, C.POINTER(_VP)]),
    "tsp_level_boxes": (C.c_int, [_VP, C.POINTER(C.c_int32)]),
    "tsp_level_set_state": (C.c_int, [_VP, _VP, C.c_int32]),
    "tsp_level_get_state": (C.c_int, [_VP, _VP, C.c_int32]),
    "tsp_level_init_expr": (C.c_int, [_VP, _VP, _VP, _VP]),
    "tsp_batch_time_rk4": (C.c_int, [_VP, C.c_int64, C.c_int32, _DP]),
    "tsp_batch_sync": (C.c_int, [_VP]),
    "tsp_adjoint_create": (C.c_int, [_VP, C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int32, C.POINTER(_VP)]),
    "tsp_adjoint_destroy": (C.c_int, [_VP]),
    "tsp_adjoint_forward": (C.c_int, [_VP]),
    "tsp_adjoint_trace": (C.c_int, [_VP, _VP, C.c_int32]),
    "tsp_adjoint_checkpoint": (C.c_int, [_VP, C.c_int32, _VP, C.c_int32]),
    "tsp_adjoint_backward": (C.c_int, [_VP, _VP, C.c_int32, _VP, C.c_int32, C.c_int32]),
    "tsp_adjoint_apply": (C.c_int, [_VP, _VP, _VP, C.c_int32]),
    # stage callback: int (*)(void *user, const tsp_stage_view *view, void *hip_stream)
    "tsp_rk4_steps_custom": (C.c_int, [_VP, C.c_int64, _VP, _VP]),
    "tsp_euler_steps_custom": (C.c_int, [_VP, C.c_int64, _VP, _VP]),
    "tsp_set_bc_callback": (C.c_int, [_VP, _VP, _VP]),
    "tsp_batch_adjoint_create": (C.c_int, [_VP, C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int32,
                                           C.POINTER(_VP)]),
    "tsp_batch_adjoint_destroy": (C.c_int, [_VP]),
    "tsp_batch_adjoint_forward": (C.c_int, [_VP]),
    "tsp_batch_adjoint_trace": (C.c_int, [_VP, _VP, C.c_int32]),
    "tsp_batch_adjoint_backward": (C.c_int, [_VP, _VP, C.c_int32, _VP, C.c_int32, C.c_int32]),
}

_lib = None


def _share_torch_runtime() -> None:
    """Load PyTorch's HIP runtime before the library, when PyTorch is installed.

    PyTorch's ROCm libraries depend on ``libamdhip64.so`` found next to them
    (RPATH $ORIGIN), whose soname is ``libamdhip64.so.7``; this library depends
    on ``libamdhip64.so.7``. Loading PyTorch first makes the dynamic loader
    resolve our dependency to the runtime that is already mapped, so the
    process holds one HIP runtime: device pointers, streams and
    ``torch.cuda.synchronize()`` are then shared with PyTorch. Loaded the other
    way round the process would map two HIP/HSA runtimes.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib() -> C.CDLL:
    """Load the HIP library (once). Raises if it is missing: there is no fallback."""
    global _lib
    if _lib is None:
        _share_torch_runtime()
        path = Path(os.environ.get("TSP_LIBRARY", LIB_PATH))
        if not path.exists():
            raise RuntimeError(
                f"HIP library {path} is missing; build it with `python -m treesap_amd.build` "
                "(treesap_amd has no CPU fallback)")
        handle = C.CDLL(str(path))
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int, where: str) -> None:
    if status != OK:
        msg = lib().tsp_last_error()
        raise TspError(status, where, msg.decode() if msg else "")


def declared_symbols() -> list[str]:
    """Function names declared in include/treesap_hip.h (parsed from the header text)."""
    import re

    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(tsp_[a-z_0-9]+)\s*\(", text, re.M)))
