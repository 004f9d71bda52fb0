"""Differentiation form: hand-written adjoint kernels behind torch.autograd.

* ``PairSum``     - v = sum_s g_s K(x; y0_s, y1_s); backward runs the
                    adjoint Biot-Savart kernels (fj_pair_sum_backward) for
                    dL/dx, dL/dy0, dL/dy1, dL/dg.
* ``LinearSolve`` - x = A^{-1} b by the device LU; backward solves the
                    transposed system with the same factors (fj_lu_solve,
                    trans=1): dL/db = A^{-T} xbar, dL/dA = -(A^{-T} xbar) x^T.
* ``aic_solve``   - Gamma = A(geometry)^{-1} b without ever differentiating
                    the assembled matrix: in reverse mode the geometry
                    gradient of -lambda^T A Gamma is the adjoint of the pair
                    sum v_ind(c; zeta, Gamma) . n weighted by lambda, so it
                    reuses PairSum's adjoint kernels (``AdjointSolve``
                    supplies lambda = A^-T xbar).

These replace what the reference obtains from jax.vjp of its residuals
(aero/uvlm.py:1529-2047, coupled/gradients/coupled.py:540-629).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _kernels as K
from . import _lib as L
from .utils import SegmentKernel, as_segment_kernel

F64 = torch.float64
_CONSTS: dict = {}


def device_const(x, device):
    """A small FP64 constant (mirror plane, flow direction, ...) on the
    device, copied there once: a copy from pageable host memory on every use
    waits for the stream to drain, a stall per call in the reverse passes."""
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=F64)
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    key = (a.tobytes(), a.shape, str(torch.device(device)))
    t = _CONSTS.get(key)
    if t is None:
        if len(_CONSTS) > 4096:
            _CONSTS.clear()
        t = torch.as_tensor(a, device=device)
        _CONSTS[key] = t
    return t


class PairSum(torch.autograd.Function):
    @staticmethod
    def forward(ctx, targets, y0, y1, g, kernel: SegmentKernel):
        kern = as_segment_kernel(kernel)
        tflat = targets.detach().reshape(-1, 3).contiguous()
        ends = torch.stack((y0.detach().reshape(-1, 3), y1.detach().reshape(-1, 3)), dim=1).contiguous()
        seg = K.segments_from_endpoints(ends, g.detach().reshape(-1).contiguous(), kern.kind, kern.eps)
        v = K.pair_sum(seg, tflat, kern.kind, kern.eps, kern.rcut)
        ctx.save_for_backward(tflat, seg)
        ctx.kern = kern
        ctx.shapes = (targets.shape, y0.shape, y1.shape, g.shape)
        return v.reshape(targets.shape)

    @staticmethod
    def backward(ctx, gv):
        tflat, seg = ctx.saved_tensors
        kern = ctx.kern
        ts, y0s, y1s, gs = ctx.shapes
        gv = gv.reshape(-1, 3).contiguous().to(F64)
        need_x = ctx.needs_input_grad[0]
        need_s = any(ctx.needs_input_grad[1:4])
        gx = torch.empty_like(tflat) if need_x else None
        gseg = torch.empty((seg.shape[0], 7), dtype=F64, device=seg.device) if need_s else None
        # the workspace selects the one-pass adjoint kernel (both sides at once)
        nb = L.lib().fj_pair_sum_backward_workspace_bytes(tflat.shape[0], seg.shape[0])
        ws = torch.empty(nb // 8 + 1, dtype=F64, device=seg.device)
        rc = L.lib().fj_pair_sum_backward(L.ptr(seg), seg.shape[0], L.ptr(tflat), tflat.shape[0], L.ptr(gv),
                                          kern.kind, kern.eps, kern.rcut, L.ptr(gx) if need_x else None, 0,
                                          L.ptr(gseg) if need_s else None, L.ptr(ws), L.stream())
        L.check(rc, "fj_pair_sum_backward")
        g_t = gx.reshape(ts) if need_x else None
        g0 = gseg[:, 0:3].reshape(y0s) if need_s else None
        g1 = gseg[:, 3:6].reshape(y1s) if need_s else None
        gg = gseg[:, 6].reshape(gs) if need_s else None
        return g_t, g0, g1, gg, None


def pair_sum(targets, y0, y1, g, kernel=None):
    return PairSum.apply(targets, y0, y1, g, kernel)


def grid_filaments(zeta, gamma):
    """Differentiable de-duplicated filaments of a vertex grid (aero/aic.py:277-287)."""
    gm = torch.diff(torch.nn.functional.pad(gamma, (1, 1, 0, 0)), dim=1)
    gn = -torch.diff(torch.nn.functional.pad(gamma, (0, 0, 1, 1)), dim=0)
    y0 = torch.cat((zeta[:-1].reshape(-1, 3), zeta[:, :-1].reshape(-1, 3)))
    y1 = torch.cat((zeta[1:].reshape(-1, 3), zeta[:, 1:].reshape(-1, 3)))
    return y0, y1, torch.cat((gm.reshape(-1), gn.reshape(-1)))


def mirror_grid(zeta, mirror):
    p = device_const(mirror[0], zeta.device)
    nrm = device_const(mirror[1], zeta.device)
    return zeta - 2.0 * ((zeta - p) * nrm).sum(-1, keepdim=True) * nrm


def v_ind(targets, zetas, gammas, kernel=None, mirror=None):
    """Differentiable compute_v_ind (aero/aic.py:248-310) on PairSum."""
    ys0, ys1, gs = [], [], []
    for z, g in zip(zetas, gammas):
        if z.numel() == 0 or g.numel() == 0:
            continue
        y0, y1, ge = grid_filaments(z, g)
        ys0.append(y0), ys1.append(y1), gs.append(ge)
        if mirror is not None:
            zm = mirror_grid(z, mirror)
            y0, y1, ge = grid_filaments(zm, g)
            ys0.append(y0), ys1.append(y1), gs.append(-ge)
    if not gs:
        return torch.zeros_like(targets)
    return pair_sum(targets, torch.cat(ys0), torch.cat(ys1), torch.cat(gs), kernel)


class LinearSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, b):
        fac = K.lu_factor(A.detach().contiguous())
        x = K.lu_solve(fac, b.detach().contiguous())
        ctx.save_for_backward(fac.lu, fac.piv, fac.aux, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        LU, piv, aux, x = ctx.saved_tensors
        lam = K.lu_solve((LU, piv, aux), gx.contiguous().to(F64), trans=True)
        gA = -torch.outer(lam, x) if ctx.needs_input_grad[0] else None
        return gA, lam


def linear_solve(A, b):
    return LinearSolve.apply(A, b)


def _aic_action(c, nc, zetas_b, zetas_w, gamma_vec, shapes, kernel, mirror):
    """(A Gamma)_t = n_t . v_ind(c_t; bound rings with Gamma [+ static wake
    rings carrying the TE row of Gamma]) -- the action of the assembled AIC
    (aero/aic.py:137-205) written with differentiable pair sums."""
    gammas, zs = [], []
    off = 0
    for i, (m, n) in enumerate(shapes):
        g = gamma_vec[off:off + m * n].reshape(m, n)
        off += m * n
        zs.append(zetas_b[i])
        gammas.append(g)
        if zetas_w is not None and zetas_w[i].shape[0] > 1:
            zs.append(zetas_w[i])
            gammas.append(g[-1:].expand(zetas_w[i].shape[0] - 1, n))
    v = v_ind(c, zs, gammas, kernel, mirror)
    return (v * nc).sum(-1)


class _KeptInverseSolve:
    """A^-T g by FP64 refinement around M^T, M the kept inverse the forward
    solve used; a refinement that never passes its stop test is counted in
    ``solver.adjoint_failures`` (a device counter the caller reads once
    after the reverse pass)."""

    def __init__(self, A, M, solver):
        self.A, self.M, self.solver = A, M, solver

    def transposed(self, g):
        lam, it = K.solve_refined_inv(self.A.t().contiguous(), self.M.t().contiguous(), g, max_iter=8)
        sv = self.solver
        fails = (it < 0).to(torch.int32)
        prev = getattr(sv, "adjoint_failures", None)
        sv.adjoint_failures = fails if prev is None else prev + fails
        return lam


class _LUSolve:
    """A^-T g with the LU factor of A (fj_lu_solve, trans = 1)."""

    def __init__(self, fac):
        self.fac = fac

    def transposed(self, g):
        return K.lu_solve(self.fac, g, trans=True)


class AdjointSolve(torch.autograd.Function):
    """Zero in the forward pass; the backward returns A^-T of its incoming
    gradient (``handle.transposed``).
    ``aic_solve`` adds it, with its value cancelled, to carry the implicit
    derivative of the circulation solve through the caller's own graph."""

    @staticmethod
    def forward(ctx, r, handle):
        ctx.handle = handle
        return torch.zeros_like(r)

    @staticmethod
    def backward(ctx, g):
        return ctx.handle.transposed(g.contiguous().to(F64)), None


def aic_solve(b, c, nc, zetas_b, zetas_w=None, kernel=None, mirror=None, solver=None):
    """Differentiable circulation solve Gamma = A^-1 b, A = AIC(c, nc,
    zeta_b[, zeta_w]) assembled on the device (c, nc packed (N_b, 3));
    differentiable in b and in the geometry without ever differentiating the
    assembled matrix.

    The value is the device solve x (``solver``: a ``linalg.RefinedSolver``,
    the production form: its kept inverse of an earlier matrix and FP64
    refinement against the current A, refactoring when refinement fails; None:
    an LU of A).  The derivative is the implicit one, dx = -A^-1 (dA x - db):
    the returned tensor is x + (s - s.detach()) with s = -AdjointSolve(r),
    r = A(geometry) x - b written with differentiable pair sums (so the
    reverse pass gets lambda = A^-T xbar from AdjointSolve, then dL/db =
    lambda and dL/dgeometry = -lambda^T d(A x)/dgeometry from the pair sums'
    adjoint kernels), all inside the caller's graph."""
    from .aic import compute_aic_solve

    shapes = tuple((z.shape[0] - 1, z.shape[1] - 1) for z in zetas_b)
    nb = len(shapes)
    kern = as_segment_kernel(kernel)
    with torch.no_grad():
        cs, ns, off = [], [], 0
        for (m, n) in shapes:
            cs.append(c.detach()[off:off + m * n].reshape(m, n, 3))
            ns.append(nc.detach()[off:off + m * n].reshape(m, n, 3))
            off += m * n
        mp = mn = None
        if mirror is not None:
            mp, mn = mirror
        zb = [z.detach() for z in zetas_b]
        zw = [z.detach() for z in zetas_w] if zetas_w is not None else None
        A = compute_aic_solve(cs, ns, zb, zw, [kern] * nb, [kern] * nb if zw is not None else None,
                              mirror_point=mp, mirror_normal=mn)
        if solver is not None:
            x = solver.solve(A, b.detach().contiguous())
            handle = _KeptInverseSolve(A, solver.inverse, solver)
        else:
            fac = K.lu_factor(A, overwrite=True)
            x = K.lu_solve(fac, b.detach().contiguous())
            handle = _LUSolve(fac)
    inputs = [b, c, nc] + list(zetas_b) + (list(zetas_w) if zetas_w is not None else [])
    if not torch.is_grad_enabled() or not any(t.requires_grad for t in inputs):
        return x
    r = _aic_action(c, nc, list(zetas_b), list(zetas_w) if zetas_w is not None else None, x, shapes, kern,
                    mirror) - b
    s = -AdjointSolve.apply(r, handle)
    return x + (s - s.detach())