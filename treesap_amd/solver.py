"""Host-side solver object over the C ABI.

``Solver`` plays the role of ``treesap::SingleLevelRK4Integrator`` /
``SingleLevelEulerIntegrator`` (integrators/SingleLevelRK4Integrator.hpp:39-167,
integrators/SingleLevelEulerIntegrator.hpp:31-89) for the registered systems:
construct from a model description, upload the initial condition (the
``initialise()`` step, SingleLevelIntegrator.hpp:75-93), then ``step()``.
Unlike the reference it exposes the state (``get_state``) so results can be
checked. All compute runs in the HIP library; this module only moves arguments.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .workloads import Case, make_desc


def _is_torch_cuda(x) -> bool:
    return type(x).__module__.startswith("torch") and getattr(x, "is_cuda", False)


def _torch_ready(t) -> None:
    """The library works on its own stream: finish PyTorch's pending work on `t` first."""
    import torch

    torch.cuda.current_stream(t.device).synchronize()


class Solver:
    def __init__(self, case: Case, stream=None):
        self.case = case
        self._desc = make_desc(case)
        handle = C.c_void_p()
        st = _lib.lib().tsp_create(C.byref(self._desc), C.c_void_p(stream), C.byref(handle))
        _lib.check(st, "tsp_create")
        self._h = handle

    # -- lifecycle ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().tsp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # -- state -------------------------------------------------------------
    def set_state(self, state) -> None:
        """initialise(): upload the valid-region state (ncomp, nz, ny, nx) float64."""
        if _is_torch_cuda(state):
            if tuple(state.shape) != self.case.shape or str(state.dtype) != "torch.float64":
                raise ValueError(f"state must be float64 of shape {self.case.shape}")
            state = state.contiguous()
            _torch_ready(state)
            _lib.check(_lib.lib().tsp_set_state(self._h, C.c_void_p(state.data_ptr()), 1), "tsp_set_state")
            return
        arr = np.ascontiguousarray(state, dtype=np.float64)
        if arr.shape != self.case.shape:
            raise ValueError(f"state must have shape {self.case.shape}, got {arr.shape}")
        _lib.check(_lib.lib().tsp_set_state(self._h, arr.ctypes.data_as(C.c_void_p), 0), "tsp_set_state")

    def get_state(self, out=None):
        """state_old (valid region) as a numpy array, or into a float64 torch CUDA tensor."""
        if out is not None and _is_torch_cuda(out):
            if not out.is_contiguous() or tuple(out.shape) != self.case.shape or str(out.dtype) != "torch.float64":
                raise ValueError("out must be a contiguous float64 tensor of the state shape")
            _torch_ready(out)
            _lib.check(_lib.lib().tsp_get_state(self._h, C.c_void_p(out.data_ptr()), 1), "tsp_get_state")
            return out
        arr = np.empty(self.case.shape, dtype=np.float64)
        _lib.check(_lib.lib().tsp_get_state(self._h, arr.ctypes.data_as(C.c_void_p), 0), "tsp_get_state")
        return arr

    def get_register(self, role: int) -> np.ndarray:
        """Whole ghosted register (ncomp, nz+2G, ny+2G, nx+2G)."""
        G = self.case.ghost
        n = self.case.n
        arr = np.empty((self.case.ncomp, n[2] + 2 * G, n[1] + 2 * G, n[0] + 2 * G), dtype=np.float64)
        _lib.check(_lib.lib().tsp_get_register(self._h, role, arr.ctypes.data_as(C.c_void_p)), "tsp_get_register")
        return arr

    def fill_ghosts(self, role: int = _lib.ROLE_OLD) -> None:
        _lib.check(_lib.lib().tsp_fill_ghosts(self._h, role), "tsp_fill_ghosts")

    # -- stepping ----------------------------------------------------------
    def rk4_steps(self, nsteps: int) -> None:
        _lib.check(_lib.lib().tsp_rk4_steps(self._h, int(nsteps)), "tsp_rk4_steps")

    def euler_steps(self, nsteps: int) -> None:
        _lib.check(_lib.lib().tsp_euler_steps(self._h, int(nsteps)), "tsp_euler_steps")

    def graph_rk4_steps(self, nsteps: int) -> None:
        _lib.check(_lib.lib().tsp_graph_rk4_steps(self._h, int(nsteps)), "tsp_graph_rk4_steps")

    def graph_prepare(self) -> None:
        """Capture the two-step graphs of both step parities (nothing is launched)."""
        _lib.check(_lib.lib().tsp_graph_prepare(self._h), "tsp_graph_prepare")

    def step(self) -> None:
        """Integrator::step() with the RK4 scheme."""
        self.rk4_steps(1)

    def sync(self) -> None:
        _lib.check(_lib.lib().tsp_sync(self._h), "tsp_sync")

    @property
    def time(self) -> float:
        t = C.c_double()
        s = C.c_int64()
        _lib.check(_lib.lib().tsp_get_time(self._h, C.byref(t), C.byref(s)), "tsp_get_time")
        return t.value

    @property
    def steps(self) -> int:
        t = C.c_double()
        s = C.c_int64()
        _lib.check(_lib.lib().tsp_get_time(self._h, C.byref(t), C.byref(s)), "tsp_get_time")
        return s.value

    # -- kernel configuration ---------------------------------------------
    @property
    def kernel(self) -> dict:
        cfg = _lib.TspKernelCfg()
        _lib.check(_lib.lib().tsp_get_kernel(self._h, C.byref(cfg)), "tsp_get_kernel")
        return cfg.as_dict()

    @kernel.setter
    def kernel(self, cfg: dict) -> None:
        cur = self.kernel
        cur.update(cfg)
        c = _lib.TspKernelCfg(**cur)
        _lib.check(_lib.lib().tsp_set_kernel(self._h, C.byref(c)), "tsp_set_kernel")

    def autotune(self, steps: int = 2) -> dict:
        cfg = _lib.TspKernelCfg()
        _lib.check(_lib.lib().tsp_autotune(self._h, int(steps), C.byref(cfg)), "tsp_autotune")
        return cfg.as_dict()

    def profile_rk4(self, nsteps: int):
        """Eager steps with HIP events around each launch: (per-stage ms[4], ghost-fill ms)."""
        stage = (C.c_double * 4)()
        fill = C.c_double()
        _lib.check(_lib.lib().tsp_profile_rk4(self._h, int(nsteps), stage, C.byref(fill)), "tsp_profile_rk4")
        return list(stage), fill.value

    def time_rk4(self, nsteps: int, use_graph: bool = True) -> float:
        """Device time (ms) of `nsteps` RK4 steps between two HIP events on the solver's stream."""
        ms = C.c_double()
        _lib.check(_lib.lib().tsp_time_rk4(self._h, int(nsteps), int(bool(use_graph)), C.byref(ms)), "tsp_time_rk4")
        return ms.value

    def rk4_bytes_per_cell(self) -> int:
        return int(_lib.lib().tsp_rk4_bytes_per_cell(C.byref(self._desc)))


class Batch:
    """Many independent (phi, Pi), (phi, u) and NonLinear models advanced in lockstep by one launch per
    RK4 stage and (system, stencil width, KO order) group."""

    def __init__(self, cases, stream=None):
        self.cases = list(cases)
        arr = (_lib.TspDesc * len(self.cases))(*[make_desc(c) for c in self.cases])
        self._descs = arr
        handle = C.c_void_p()
        st = _lib.lib().tsp_batch_create(arr, len(self.cases), C.c_void_p(stream), C.byref(handle))
        _lib.check(st, "tsp_batch_create")
        self._h = handle

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().tsp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.cases)

    @property
    def handle(self):
        return self._h

    @property
    def cells(self) -> int:
        n = C.c_int32()
        c = C.c_int64()
        _lib.check(_lib.lib().tsp_batch_size(self._h, C.byref(n), C.byref(c)), "tsp_batch_size")
        return c.value

    def set_state(self, index: int, state) -> None:
        case = self.cases[index]
        if _is_torch_cuda(state):
            if tuple(state.shape) != case.shape or str(state.dtype) != "torch.float64":
                raise ValueError(f"state must be float64 of shape {case.shape}")
            state = state.contiguous()
            _torch_ready(state)
            _lib.check(_lib.lib().tsp_batch_set_state(self._h, index, C.c_void_p(state.data_ptr()), 1),
                       "tsp_batch_set_state")
            return
        arr = np.ascontiguousarray(state, dtype=np.float64)
        if arr.shape != case.shape:
            raise ValueError(f"state must have shape {case.shape}, got {arr.shape}")
        _lib.check(_lib.lib().tsp_batch_set_state(self._h, index, arr.ctypes.data_as(C.c_void_p), 0),
                   "tsp_batch_set_state")

    def get_state(self, index: int, out=None):
        """State of instance `index` as a numpy array, or into a float64 torch CUDA tensor."""
        if out is not None and _is_torch_cuda(out):
            if not out.is_contiguous() or tuple(out.shape) != self.cases[index].shape or str(out.dtype) != "torch.float64":
                raise ValueError("out must be a contiguous float64 tensor of the instance's state shape")
            _torch_ready(out)
            _lib.check(_lib.lib().tsp_batch_get_state(self._h, index, C.c_void_p(out.data_ptr()), 1),
                       "tsp_batch_get_state")
            return out
        arr = np.empty(self.cases[index].shape, dtype=np.float64)
        _lib.check(_lib.lib().tsp_batch_get_state(self._h, index, arr.ctypes.data_as(C.c_void_p), 0),
                   "tsp_batch_get_state")
        return arr

    def time(self, index: int) -> float:
        t = C.c_double()
        s = C.c_int64()
        _lib.check(_lib.lib().tsp_batch_get_time(self._h, index, C.byref(t), C.byref(s)), "tsp_batch_get_time")
        return t.value

    @property
    def steps(self) -> int:
        t = C.c_double()
        s = C.c_int64()
        _lib.check(_lib.lib().tsp_batch_get_time(self._h, 0, C.byref(t), C.byref(s)), "tsp_batch_get_time")
        return s.value

    @property
    def kernel(self) -> dict:
        cfg = _lib.TspBatchCfg()
        _lib.check(_lib.lib().tsp_batch_get_kernel(self._h, C.byref(cfg)), "tsp_batch_get_kernel")
        return cfg.as_dict()

    @kernel.setter
    def kernel(self, cfg: dict) -> None:
        cur = self.kernel
        cur.update(cfg)
        c = _lib.TspBatchCfg(**cur)
        _lib.check(_lib.lib().tsp_batch_set_kernel(self._h, C.byref(c)), "tsp_batch_set_kernel")

    def group_kernel(self, stencil_n: int, ko_order: int, system: str = "phipi") -> dict:
        """Kernel configuration of one (system, stencil width, KO order) group."""
        cfg = _lib.TspBatchCfg()
        _lib.check(_lib.lib().tsp_batch_get_group_kernel(self._h, _lib.SYSTEMS[system], int(stencil_n),
                                                           int(ko_order), C.byref(cfg)),
                   "tsp_batch_get_group_kernel")
        return cfg.as_dict()

    def set_group_kernel(self, stencil_n: int, ko_order: int, cfg: dict, system: str = "phipi") -> None:
        """Set one group's kernel configuration (fields not given keep their current values)."""
        cur = self.group_kernel(stencil_n, ko_order, system)
        cur.update(cfg)
        c = _lib.TspBatchCfg(**cur)
        _lib.check(_lib.lib().tsp_batch_set_group_kernel(self._h, _lib.SYSTEMS[system], int(stencil_n),
                                                           int(ko_order), C.byref(c)),
                   "tsp_batch_set_group_kernel")

    def rk4_steps(self, nsteps: int) -> None:
        _lib.check(_lib.lib().tsp_batch_rk4_steps(self._h, int(nsteps)), "tsp_batch_rk4_steps")

    def graph_rk4_steps(self, nsteps: int) -> None:
        _lib.check(_lib.lib().tsp_batch_graph_rk4_steps(self._h, int(nsteps)), "tsp_batch_graph_rk4_steps")

    def graph_prepare(self) -> None:
        """Capture the two-step graphs of both step parities (nothing is launched)."""
        _lib.check(_lib.lib().tsp_batch_graph_prepare(self._h), "tsp_batch_graph_prepare")

    def time_rk4(self, nsteps: int, use_graph: bool = True) -> float:
        ms = C.c_double()
        _lib.check(_lib.lib().tsp_batch_time_rk4(self._h, int(nsteps), int(bool(use_graph)), C.byref(ms)),
                   "tsp_batch_time_rk4")
        return ms.value

    def sync(self) -> None:
        _lib.check(_lib.lib().tsp_batch_sync(self._h), "tsp_batch_sync")

    def euler_steps(self, nsteps: int) -> None:
        _lib.check(_lib.lib().tsp_batch_euler_steps(self._h, int(nsteps)), "tsp_batch_euler_steps")

    def autotune(self, steps: int = 6) -> dict:
        """The library's batch autotuner (tsp_batch_autotune): every (system, stencil width, KO order)
        group tuned on its own; returns the largest group's configuration (group_kernel gives the others)."""
        cfg = _lib.TspBatchCfg()
        _lib.check(_lib.lib().tsp_batch_autotune(self._h, int(steps), C.byref(cfg)), "tsp_batch_autotune")
        return cfg.as_dict()


class Level(Batch):
    """One level split into boxes of at most `block` cells per axis (the BoxArray of
    SingleLevelIntegrator(model, block_size), SingleLevelIntegrator.hpp:47-49,172-176), each box its own
    registers, ghosts exchanged between boxes before every stage; (phi, Pi), (phi, u) or NonLinear. Results are
    bit-identical to one box of the whole domain."""

    def __init__(self, case: Case, block, stream=None):
        self.case = case
        self._desc = make_desc(case)
        blk = (C.c_int32 * 3)(*[int(b) for b in block])
        handle = C.c_void_p()
        st = _lib.lib().tsp_level_create(C.byref(self._desc), blk, C.c_void_p(stream), C.byref(handle))
        _lib.check(st, "tsp_level_create")
        self._h = handle
        n = C.c_int32()
        _lib.check(_lib.lib().tsp_batch_size(self._h, C.byref(n), None), "tsp_batch_size")
        lohi = (C.c_int32 * (6 * n.value))()
        _lib.check(_lib.lib().tsp_level_boxes(self._h, lohi), "tsp_level_boxes")
        self.boxes = [tuple(lohi[6 * b: 6 * b + 6]) for b in range(n.value)]
        self.cases = [case.replace(n=tuple(hi - lo + 1 for lo, hi in zip(bx[:3], bx[3:])), periodic=(0, 0, 0))
                      for bx in self.boxes]

    def set_state(self, state) -> None:  # noqa: D401 -- the whole domain
        """The domain's valid state (ncomp, nz, ny, nx) float64, numpy or torch CUDA."""
        if _is_torch_cuda(state):
            if tuple(state.shape) != self.case.shape or str(state.dtype) != "torch.float64":
                raise ValueError(f"state must be float64 of shape {self.case.shape}")
            state = state.contiguous()
            _torch_ready(state)
            _lib.check(_lib.lib().tsp_level_set_state(self._h, C.c_void_p(state.data_ptr()), 1), "tsp_level_set_state")
            return
        arr = np.ascontiguousarray(state, dtype=np.float64)
        if arr.shape != self.case.shape:
            raise ValueError(f"state must have shape {self.case.shape}, got {arr.shape}")
        _lib.check(_lib.lib().tsp_level_set_state(self._h, arr.ctypes.data_as(C.c_void_p), 0), "tsp_level_set_state")

    def get_state(self, out=None):
        if out is not None and _is_torch_cuda(out):
            if not out.is_contiguous() or tuple(out.shape) != self.case.shape or str(out.dtype) != "torch.float64":
                raise ValueError("out must be a contiguous float64 tensor of the state shape")
            _torch_ready(out)
            _lib.check(_lib.lib().tsp_level_get_state(self._h, C.c_void_p(out.data_ptr()), 1), "tsp_level_get_state")
            return out
        arr = np.empty(self.case.shape, dtype=np.float64)
        _lib.check(_lib.lib().tsp_level_get_state(self._h, arr.ctypes.data_as(C.c_void_p), 0), "tsp_level_get_state")
        return arr

    @property
    def time(self) -> float:
        return Batch.time(self, 0)
