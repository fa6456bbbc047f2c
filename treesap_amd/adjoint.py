"""Discrete adjoint of the RK4 stepping and its torch.autograd binding.

``Adjoint`` wraps the tsp_adjoint_* entry points (include/treesap_hip.h): a
forward pass that records phi at receiver cells after every RK4 step and
stores checkpoints, and a backward pass that returns dJ/du^0 for
J = sum_n j_n(phi^n at the receivers) through the hand-written transposed
RK4 stage kernels. ``BatchAdjoint`` does the same for every instance of a
``Batch`` at once (tsp_batch_adjoint_*), each launch covering all instances.
``ReceiverLoss`` exposes
    L(phi_0) = sum_b sum_{n=1..N} sum_r (phi_b^n[r] - d_b^n[r])^2
(Pi_0 = 0) as a torch.autograd.Function on float64 CUDA tensors, the B
sources solved as one batch; its backward runs the batched adjoint. The reference has no adjoint; its forward stepping is
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


class BatchAdjoint:
    """Discrete adjoint of a whole Batch (tsp_batch_adjoint_*): one forward pass and one reverse
    sweep for every instance at once; receivers shared by all instances. Traces and dtrace are
    (count, nsteps, nrec); gradients are the instances' states concatenated (instance-major)."""

    def __init__(self, batch, receivers, nsteps: int, checkpoint_every: int = 50):
        rec = np.ascontiguousarray(np.asarray(receivers, dtype=np.int32).reshape(-1, 3))
        self.batch = batch
        self.receivers = rec
        self.nrec = rec.shape[0]
        self.nsteps = int(nsteps)
        self.checkpoint_every = int(checkpoint_every)
        self.count = len(batch)
        handle = C.c_void_p()
        st = _lib.lib().tsp_batch_adjoint_create(batch.handle, rec.ctypes.data_as(C.POINTER(C.c_int32)), self.nrec,
                                                 self.nsteps, self.checkpoint_every, C.byref(handle))
        _lib.check(st, "tsp_batch_adjoint_create")
        self._h = handle

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().tsp_batch_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def grad_elems(self) -> int:
        return sum(int(np.prod(c.shape)) for c in self.batch.cases)

    def forward(self) -> None:
        _lib.check(_lib.lib().tsp_batch_adjoint_forward(self._h), "tsp_batch_adjoint_forward")

    def trace(self, out=None):
        shape = (self.count, self.nsteps, self.nrec)
        if out is not None and _is_torch_cuda(out):
            _check_cuda(out, shape, "out")
            _torch_ready(out)
            _lib.check(_lib.lib().tsp_batch_adjoint_trace(self._h, C.c_void_p(out.data_ptr()), 1),
                       "tsp_batch_adjoint_trace")
            return out
        arr = np.empty(shape)
        _lib.check(_lib.lib().tsp_batch_adjoint_trace(self._h, arr.ctypes.data_as(C.c_void_p), 0),
                   "tsp_batch_adjoint_trace")
        return arr

    def backward(self, dtrace, recompute: bool = False, out=None):
        """dJ/du^0 of every instance (flat, instance-major) for dtrace[i, n-1, r] = dJ/dphi_i^n(receiver r)."""
        shape = (self.count, self.nsteps, self.nrec)
        if _is_torch_cuda(dtrace):
            dtrace = dtrace.contiguous()
            _check_cuda(dtrace, shape, "dtrace")
            _torch_ready(dtrace)
            dptr, dev = C.c_void_p(dtrace.data_ptr()), 1
        else:
            dtrace = np.ascontiguousarray(dtrace, dtype=np.float64)
            if dtrace.shape != shape:
                raise ValueError(f"dtrace must have shape {shape}")
            dptr, dev = dtrace.ctypes.data_as(C.c_void_p), 0
        if out is not None and _is_torch_cuda(out):
            _check_cuda(out, (self.grad_elems,), "out")
            _torch_ready(out)
            _lib.check(_lib.lib().tsp_batch_adjoint_backward(self._h, dptr, dev, C.c_void_p(out.data_ptr()), 1,
                                                             int(bool(recompute))), "tsp_batch_adjoint_backward")
            return out
        grad = np.empty(self.grad_elems)
        _lib.check(_lib.lib().tsp_batch_adjoint_backward(self._h, dptr, dev, grad.ctypes.data_as(C.c_void_p), 0,
                                                         int(bool(recompute))), "tsp_batch_adjoint_backward")
        return grad


class ReceiverProblem:
    """B independent (phi, Pi) solves sharing one model, receiver set and step count, run as one Batch
    with one BatchAdjoint (every launch covers all B sources)."""

    def __init__(self, case, receivers, nsteps: int, batch: int, checkpoint_every: int = 50):
        from .solver import Batch

        if case.system != "phipi":
            raise ValueError("the adjoint covers the (phi, Pi) system")
        self.case = case
        self.nsteps = int(nsteps)
        self.batch = int(batch)
        self.solver = Batch([case] * self.batch)
        self.adjoint = BatchAdjoint(self.solver, receivers, nsteps, checkpoint_every)
        self.nrec = self.adjoint.nrec

    def close(self):
        self.adjoint.close()
        self.solver.close()


def _receiver_loss_function():
    import torch

    class ReceiverLoss(torch.autograd.Function):
        """L(phi0) = sum_b sum_n sum_r (phi_b^n[r] - target[b, n-1, r])^2 with Pi0 = 0.

        phi0: (B, nz, ny, nx) float64 CUDA; target: (B, nsteps, nrec) float64 CUDA."""

        @staticmethod
        def forward(ctx, phi0, target, problem, recompute=False):
            B = phi0.shape[0]
            st = torch.zeros(problem.case.shape, dtype=torch.float64, device=phi0.device)
            for b in range(B):
                st[0] = phi0[b]
                problem.solver.set_state(b, st)
            problem.adjoint.forward()
            tr = torch.empty((B, problem.nsteps, problem.nrec), dtype=torch.float64, device=phi0.device)
            problem.adjoint.trace(out=tr)
            resid = tr - target
            ctx.save_for_backward(resid)
            ctx.problem = problem
            ctx.recompute = recompute
            return (resid * resid).sum()

        @staticmethod
        def backward(ctx, grad_out):
            (resid,) = ctx.saved_tensors
            problem = ctx.problem
            dtrace = ((2.0 * grad_out) * resid).contiguous()
            B = resid.shape[0]
            g = torch.empty(problem.adjoint.grad_elems, dtype=torch.float64, device=resid.device)
            problem.adjoint.backward(dtrace, recompute=ctx.recompute, out=g)
            return g.view(B, *problem.case.shape)[:, 0], None, None, None

    return ReceiverLoss


_RL = None


def receiver_loss(phi0, target, problem: ReceiverProblem, recompute: bool = False):
    """Differentiable receiver loss (see ReceiverLoss)."""
    global _RL
    if _RL is None:
        _RL = _receiver_loss_function()
    return _RL.apply(phi0, target, problem, recompute)
