"""Discrete adjoint of the RK4 stepping and its torch.autograd binding.

``Adjoint`` wraps the tsp_adjoint_* entry points (include/treesap_hip.h): a
forward pass that records phi at receiver cells after every RK4 step and
stores checkpoints, and a backward pass that returns dJ/du^0 for
J = sum_n j_n(phi^n at the receivers) through the hand-written transposed
RK4 stage kernels. ``ReceiverLoss`` exposes
    L(phi_0) = sum_b sum_{n=1..N} sum_r (phi_b^n[r] - d_b^n[r])^2
(Pi_0 = 0) as a torch.autograd.Function on float64 CUDA tensors; its backward
runs the adjoint. The reference has no adjoint; its forward stepping is
integrators/SingleLevelRK4Integrator.hpp:95-157.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .solver import Solver, _is_torch_cuda, _torch_ready


def _check_cuda(t, shape, name):
    """A device tensor handed to the library: contiguous float64 of exactly `shape` (the library
    reads or writes that many doubles through its pointer)."""
    if tuple(t.shape) != tuple(shape) or str(t.dtype) != "torch.float64" or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float64 tensor of shape {tuple(shape)}, got "
                         f"{tuple(t.shape)} {t.dtype}{'' if t.is_contiguous() else ' (non-contiguous)'}")


class Adjoint:
    def __init__(self, solver: Solver, receivers, nsteps: int, checkpoint_every: int = 50):
        rec = np.ascontiguousarray(np.asarray(receivers, dtype=np.int32).reshape(-1, 3))
        self.solver = solver
        self.receivers = rec
        self.nrec = rec.shape[0]
        self.nsteps = int(nsteps)
        self.checkpoint_every = int(checkpoint_every)
        handle = C.c_void_p()
        st = _lib.lib().tsp_adjoint_create(solver.handle, rec.ctypes.data_as(C.POINTER(C.c_int32)), self.nrec,
                                           self.nsteps, self.checkpoint_every, C.byref(handle))
        _lib.check(st, "tsp_adjoint_create")
        self._h = handle

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().tsp_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nslots(self) -> int:
        K = self.checkpoint_every
        return self.nsteps // K + 1 + (1 if self.nsteps % K else 0)

    def forward(self) -> None:
        _lib.check(_lib.lib().tsp_adjoint_forward(self._h), "tsp_adjoint_forward")

    def trace(self, out=None):
        if out is not None and _is_torch_cuda(out):
            _check_cuda(out, (self.nsteps, self.nrec), "out")
            _torch_ready(out)
            _lib.check(_lib.lib().tsp_adjoint_trace(self._h, C.c_void_p(out.data_ptr()), 1), "tsp_adjoint_trace")
            return out
        arr = np.empty((self.nsteps, self.nrec))
        _lib.check(_lib.lib().tsp_adjoint_trace(self._h, arr.ctypes.data_as(C.c_void_p), 0), "tsp_adjoint_trace")
        return arr

    def checkpoint(self, slot: int) -> np.ndarray:
        arr = np.empty(self.solver.case.shape)
        _lib.check(_lib.lib().tsp_adjoint_checkpoint(self._h, int(slot), arr.ctypes.data_as(C.c_void_p), 0),
                   "tsp_adjoint_checkpoint")
        return arr

    def backward(self, dtrace, recompute: bool = False, out=None):
        """dJ/du^0 for dtrace[n-1, r] = dJ/dphi^n(receiver r)."""
        if _is_torch_cuda(dtrace):
            dtrace = dtrace.contiguous()
            _check_cuda(dtrace, (self.nsteps, self.nrec), "dtrace")
            _torch_ready(dtrace)
            dptr, dev = C.c_void_p(dtrace.data_ptr()), 1
        else:
            dtrace = np.ascontiguousarray(dtrace, dtype=np.float64)
            if dtrace.shape != (self.nsteps, self.nrec):
                raise ValueError(f"dtrace must have shape {(self.nsteps, self.nrec)}")
            dptr, dev = dtrace.ctypes.data_as(C.c_void_p), 0
        if out is not None and _is_torch_cuda(out):
            _check_cuda(out, self.solver.case.shape, "out")
            _torch_ready(out)
            _lib.check(_lib.lib().tsp_adjoint_backward(self._h, dptr, dev, C.c_void_p(out.data_ptr()), 1,
                                                       int(bool(recompute))), "tsp_adjoint_backward")
            return out
        grad = np.empty(self.solver.case.shape)
        _lib.check(_lib.lib().tsp_adjoint_backward(self._h, dptr, dev, grad.ctypes.data_as(C.c_void_p), 0,
                                                   int(bool(recompute))), "tsp_adjoint_backward")
        return grad

    def apply(self, lam: np.ndarray) -> np.ndarray:
        """G^T lam for one RK4 step (valid layout)."""
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        out = np.empty_like(lam)
        _lib.check(_lib.lib().tsp_adjoint_apply(self._h, lam.ctypes.data_as(C.c_void_p),
                                                out.ctypes.data_as(C.c_void_p), 0), "tsp_adjoint_apply")
        return out


class ReceiverProblem:
    """B independent (phi, Pi) solves sharing one model, receiver set and step count; one solver and one
    adjoint object per source."""

    def __init__(self, case, receivers, nsteps: int, batch: int, checkpoint_every: int = 50):
        if case.system != "phipi":
            raise ValueError("the adjoint covers the (phi, Pi) system")
        self.case = case
        self.nsteps = int(nsteps)
        self.batch = int(batch)
        self.solvers = [Solver(case) for _ in range(self.batch)]
        self.adjoints = [Adjoint(s, receivers, nsteps, checkpoint_every) for s in self.solvers]
        self.nrec = self.adjoints[0].nrec

    def close(self):
        for a in self.adjoints:
            a.close()
        for s in self.solvers:
            s.close()


def _receiver_loss_function():
    import torch

    class ReceiverLoss(torch.autograd.Function):
        """L(phi0) = sum_b sum_n sum_r (phi_b^n[r] - target[b, n-1, r])^2 with Pi0 = 0.

        phi0: (B, nz, ny, nx) float64 CUDA; target: (B, nsteps, nrec) float64 CUDA."""

        @staticmethod
        def forward(ctx, phi0, target, problem, recompute=False):
            B = phi0.shape[0]
            traces = []
            for b in range(B):
                st = torch.zeros(problem.case.shape, dtype=torch.float64, device=phi0.device)
                st[0] = phi0[b]
                problem.solvers[b].set_state(st)
                problem.adjoints[b].forward()
                tr = torch.empty((problem.nsteps, problem.nrec), dtype=torch.float64, device=phi0.device)
                problem.adjoints[b].trace(out=tr)
                traces.append(tr)
            resid = torch.stack(traces) - target
            ctx.save_for_backward(resid)
            ctx.problem = problem
            ctx.recompute = recompute
            return (resid * resid).sum()

        @staticmethod
        def backward(ctx, grad_out):
            (resid,) = ctx.saved_tensors
            problem = ctx.problem
            dtrace = (2.0 * grad_out) * resid
            grads = []
            for b in range(resid.shape[0]):
                g = torch.empty(problem.case.shape, dtype=torch.float64, device=resid.device)
                problem.adjoints[b].backward(dtrace[b].contiguous(), recompute=ctx.recompute, out=g)
                grads.append(g[0])
            return torch.stack(grads), None, None, None

    return ReceiverLoss


_RL = None


def receiver_loss(phi0, target, problem: ReceiverProblem, recompute: bool = False):
    """Differentiable receiver loss (see ReceiverLoss)."""
    global _RL
    if _RL is None:
        _RL = _receiver_loss_function()
    return _RL.apply(phi0, target, problem, recompute)
