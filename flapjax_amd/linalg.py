"""Circulation solve (replaces jnp.linalg.solve at aero/uvlm.py:985 / :1634).

Two forms:
  * ``solve``           FP64 LU with partial pivoting + triangular solves;
  * ``RefinedSolver``   a kept factor: the LU (FP64, or FP32 on the matrix
                        cores) of the matrix at the last (re)factorisation and
                        its explicit inverse M; every solve is x0 = M b then
                        FP64 refinement against the CURRENT matrix
                        (r = b - A x, x += M r) until LAPACK dsgesv's stop test
                        ||r|| <= sqrt(n) eps ||A|| ||x|| passes.  Each solve is
                        a few mat-vecs over all CUs.  The stop-test outcome is
                        copied to pinned host memory behind the solve, so the
                        host learns it without draining the stream; when a
                        solve fails (the matrix drifted too far from the
                        factored one) the caller refactors and solves again.
"""

from __future__ import annotations

import torch

from . import _kernels as K
from . import _lib as L
from .utils import to_device


def lu_factor(A):
    """FP64 factor of a copy of A (K.Factor: lu, piv, aux, info)."""
    return K.lu_factor(to_device(A))


def lu_solve(fac, b, trans=False):
    return K.lu_solve(fac, to_device(b), trans=trans)


def solve(A, b):
    """x = A^{-1} b in FP64 on the device."""
    return K.lu_solve(K.lu_factor(to_device(A)), to_device(b))


class RefinedSolver:
    """Kept-factor solver with FP64 refinement (see the module docstring).

    ``submit(A, b)`` enqueues a solve and the asynchronous copy of its status;
    ``confirm()`` waits for that copy only and returns the corrections made
    (negative: the stop test never passed).  ``solve(A, b)`` does both and, on
    failure with ``reuse``, refactors on A and solves once more."""

    def __init__(self, max_iter: int | None = None, tol: float | None = None, reuse: bool = True,
                 factor_dtype=torch.float64):
        self.factor_dtype = factor_dtype
        self.max_iter = max_iter if max_iter is not None else (3 if factor_dtype == torch.float64 else 8)
        self.tol = tol
        self.reuse = reuse
        self._inv = None
        self._ws = None
        self._iters = None
        self._iters_host = None
        self._event = None
        self.factorisations = 0
        self.last_iters = None

    @property
    def inverse(self):
        return self._inv

    def factor(self, A):
        A = to_device(A)
        if self.factor_dtype == torch.float32:
            fac = K.lu_factor_f32(A.to(torch.float32))
        else:
            fac = K.lu_factor(A)
        self._inv = K.lu_inverse(fac)
        self._info = fac.info
        self.factorisations += 1

    def _buffers(self, n, dev):
        if self._ws is None or self._ws.device != dev or self._n != n:
            self._n = n
            self._ws = torch.empty(L.lib().fj_refine_workspace_bytes(n) // 8 + 1, dtype=torch.float64, device=dev)
            self._iters = torch.zeros(1, dtype=torch.int32, device=dev)
            self._iters_host = torch.zeros(1, dtype=torch.int32, pin_memory=True)
            self._event = torch.cuda.Event()

    def submit(self, A, b, out=None):
        A = to_device(A)
        b = to_device(b)
        n = A.shape[0]
        if self._inv is None or not self.reuse or self._inv.shape != A.shape:
            self.factor(A)
        self._buffers(n, A.device)
        x, _ = K.solve_refined_inv(A, self._inv, b, self.max_iter, self.tol, out=out, workspace=self._ws,
                                   iters=self._iters)
        self._iters_host.copy_(self._iters, non_blocking=True)
        self._event.record()
        return x

    def confirm(self) -> int:
        self._event.synchronize()
        self.last_iters = int(self._iters_host[0])
        return self.last_iters

    def solve(self, A, b, out=None):
        x = self.submit(A, b, out=out)
        if self.confirm() < 0 and self.reuse:
            self.factor(A)
            x = self.submit(A, b, out=x)
            self.confirm()
        if self.last_iters < 0:
            if int(self._info.item()) != 0:
                raise L.FlapjaxLibError("singular matrix in the circulation solve")
            raise L.FlapjaxLibError(f"refinement did not pass its stop test in {self.max_iter} corrections")
        return x
