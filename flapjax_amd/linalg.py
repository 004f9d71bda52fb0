"""Circulation solve (replaces jnp.linalg.solve at aero/uvlm.py:985 / :1634).

Two forms:
  * ``solve``           FP64 LU with partial pivoting + triangular solves;
  * ``RefinedSolver``   LU (FP32 on the matrix cores by default, or FP64)
                        with FP64 iterative refinement against the FP64
                        matrix (residual b - A x in FP64); the factor can be
                        reused across steps while the matrix stays close
                        (rigid motion), refactoring when refinement stops
                        converging.
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
    """FP32-factor + FP64-refinement solver with optional factor reuse."""

    def __init__(self, max_iter: int = 10, tol: float | None = None, reuse: bool = True,
                 factor_dtype=torch.float32):
        self.max_iter = max_iter
        self.tol = tol
        self.reuse = reuse
        self.factor_dtype = factor_dtype
        self._fac = None
        self.factorisations = 0
        self.last_iters = None

    def factor(self, A):
        if self.factor_dtype == torch.float32:
            self._fac = K.lu_factor_f32(A.to(torch.float32))
        else:
            self._fac = K.lu_factor(A)
        self.factorisations += 1

    def solve(self, A, b, check: bool = True):
        A = to_device(A)
        b = to_device(b)
        if self._fac is None or not self.reuse or self._fac.lu.shape != A.shape:
            self.factor(A)
        x, iters = K.solve_refined(A, self._fac, b, self.max_iter, self.tol)
        if check:
            it = int(iters.item())
            if it < 0 and self.reuse:
                # the stored factor is too far from A: refactor and retry once
                self.factor(A)
                x, iters = K.solve_refined(A, self._fac, b, self.max_iter, self.tol)
                it = int(iters.item())
            if it < 0:
                raise L.FlapjaxLibError(f"refinement did not converge in {self.max_iter} sweeps")
            self.last_iters = it
        self.iters_tensor = iters
        return x
