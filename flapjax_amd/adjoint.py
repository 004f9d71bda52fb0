"""Differentiable UVLM time marching for adjoint gradients.

A rigid/prescribed-motion UVLM run (aero/uvlm.py:824-1117 semantics: free or
prescribed wake, bound-only dynamic AIC with the wake upwash on the right-hand
side, Kutta-Joukowski forces, filtered circulation rate and unsteady force)
written with differentiable pieces only: the pair sums are
``autograd.PairSum`` (hand-written adjoint Biot-Savart kernels), the
circulation solve is ``autograd.aic_solve`` (transposed LU solve + adjoint pair
sums for the geometry), and the O(panels) stencils are torch expressions.
The time history is checkpointed every ``checkpoint_every`` steps and the
steps in between are recomputed in the reverse pass
(torch.utils.checkpoint), so memory grows with the number of checkpoints, not
with the number of steps.

Objective used by the reference's adjoint examples (time-integrated lift):
    J = sum_n dt * sum_vertices (f_steady + f_unsteady)_z
Design variables here: a twist angle per spanwise station (rotation of the
local grid about the y axis through the elastic axis) and the freestream.
"""

from __future__ import annotations

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from . import autograd as AG
from .utils import as_segment_kernel

F64 = torch.float64


def _cross(a, b):
    return torch.linalg.cross(a, b, dim=-1)


def colloc(z):
    a = 0.5 * (z[:-1] + z[1:])
    return 0.5 * (a[:, :-1] + a[:, 1:])


def normals(z):
    return _cross(z[1:, 1:] - z[:-1, :-1], z[1:, :-1] - z[:-1, 1:])


def split_to_vertex(a, axis):
    pads = [0, 0] * (a.dim() - axis - 1)
    lo = torch.nn.functional.pad(0.5 * a, pads + [0, 1])
    hi = torch.nn.functional.pad(0.5 * a, pads + [1, 0])
    return lo + hi


def twisted_grid(zeta0, twist):
    """Rotate each spanwise station of the local grid about y by twist_j."""
    c, s = torch.cos(twist), torch.sin(twist)
    x, y, z = zeta0[..., 0], zeta0[..., 1], zeta0[..., 2]
    return torch.stack((c * x + s * z, y, -s * x + c * z), dim=-1)


def apply_hg(hg, z0):
    """R_j z0_ij + x_j (algebra/se3.py:209-216) per spanwise station j."""
    return torch.einsum("jab,ijb->ija", hg[:, :3, :3], z0) + hg[None, :, :3, 3]


class DifferentiableUVLM:
    """Single-surface rigid-wing UVLM with prescribed beam poses."""

    def __init__(self, m, n, m_star, dt, rho=1.225, free_wake=True, gamma_dot_relaxation=0.7, kernel=None,
                 checkpoint_every=10):
        self.m, self.n, self.m_star = m, n, m_star
        self.dt = float(dt)
        self.rho = float(rho)
        self.free_wake = free_wake
        self.relax = float(gamma_dot_relaxation)
        self.kernel = as_segment_kernel(kernel)
        self.k = int(checkpoint_every)

    def _flow(self, u_inf, x):
        return u_inf.expand_as(x)

    def _forces(self, zb, zbd, gb, gw, zw, u_inf):
        m, n = self.m, self.n
        mid_c = 0.5 * (zb[:-1] + zb[1:])
        mid_s = 0.5 * (zb[:, :-1] + zb[:, 1:])
        mids = torch.cat((mid_c.reshape(-1, 3), mid_s.reshape(-1, 3)))
        v = AG.v_ind(mids, [zb, zw], [gb, gw], self.kernel) + u_inf
        vc = v[: m * (n + 1)].reshape(m, n + 1, 3) - 0.5 * (zbd[:-1] + zbd[1:])
        vs = v[m * (n + 1):].reshape(m + 1, n, 3) - 0.5 * (zbd[:, :-1] + zbd[:, 1:])
        gc = torch.nn.functional.pad(gb, (0, 1)) - torch.nn.functional.pad(gb, (1, 0))
        gs = torch.nn.functional.pad(-gb, (0, 0, 0, 1)) + torch.nn.functional.pad(gb, (0, 0, 1, 0))
        if gw.shape[0] > 0:
            gs = gs - torch.nn.functional.pad(gw[:1], (0, 0, m, 0))
        fc = self.rho * gc[..., None] * _cross(vc, zb[1:] - zb[:-1])
        fs = self.rho * gs[..., None] * _cross(vs, zb[:, 1:] - zb[:, :-1])
        return split_to_vertex(fc, 0) + split_to_vertex(fs, 1)

    def static(self, zb, u_inf):
        """Static solve with the flat wake lumped onto the TE (aero/uvlm.py:895-1012)."""
        m, n, ms = self.m, self.n, self.m_star
        udir = u_inf / torch.linalg.norm(u_inf)
        s = torch.arange(ms + 1, dtype=F64, device=zb.device) * (self.dt * torch.linalg.norm(u_inf))
        zw = zb[-1][None] + s[:, None, None] * udir
        c, nc = colloc(zb), normals(zb)
        b = -(self._flow(u_inf, c) * nc).sum(-1).reshape(-1)
        gb = AG.aic_solve(b, c.reshape(-1, 3), nc.reshape(-1, 3), [zb], [zw], self.kernel).reshape(m, n)
        gw = gb[-1:].expand(ms, n)
        return zw, gw, gb

    def step(self, state, zb_n, zb_nm1, zbd_n, u_inf):
        zw, gw, gb, gbd = state
        m, n = self.m, self.n
        base = zw[:-1]
        if self.free_wake:
            v = AG.v_ind(base.reshape(-1, 3), [zb_nm1, zw], [gb, gw], self.kernel).reshape(base.shape) + u_inf
        else:
            v = self._flow(u_inf, base)
        zw_n = torch.cat((zb_n[-1:], base + self.dt * v), dim=0)
        gw_n = torch.cat((gb[-1:], gw[:-1]), dim=0)
        c, nc = colloc(zb_n), normals(zb_n)
        cd = colloc(zbd_n)
        vbc = self._flow(u_inf, c) - cd + AG.v_ind(c.reshape(-1, 3), [zw_n], [gw_n], self.kernel).reshape(c.shape)
        b = -(vbc * nc).sum(-1).reshape(-1)
        gb_n = AG.aic_solve(b, c.reshape(-1, 3), nc.reshape(-1, 3), [zb_n], None, self.kernel).reshape(m, n)
        gbd_n = (gb_n - gb) / self.dt * self.relax + gbd * (1.0 - self.relax)
        f_st = self._forces(zb_n, zbd_n, gb_n, gw_n, zw_n, u_inf)
        f_un = split_to_vertex(split_to_vertex(self.rho * gbd_n[..., None] * nc, 0), 1)
        lift = f_st[..., 2].sum() + f_un[..., 2].sum()
        return (zw_n, gw_n, gb_n, gbd_n), lift

    def objective(self, twist, u_inf, zeta0, hg_t, hgd_t):
        """J = dt * sum over dynamic steps of the total vertical force."""
        z0 = twisted_grid(zeta0, twist)
        n_tstep = hg_t.shape[0]
        zb0 = apply_hg(hg_t[0], z0)
        zw, gw, gb = self.static(zb0, u_inf)
        state = (zw, gw, gb, torch.zeros_like(gb))

        def chunk(i0, i1, *st):
            st = tuple(st)
            total = torch.zeros((), dtype=F64, device=zeta0.device)
            for i in range(i0, i1):
                zb_n = apply_hg(hg_t[i], z0)
                zb_nm1 = apply_hg(hg_t[i - 1], z0)
                zbd_n = apply_hg(hgd_t[i], z0)
                st, lift = self.step(st, zb_n, zb_nm1, zbd_n, u_inf)
                total = total + self.dt * lift
            return (*st, total)

        J = torch.zeros((), dtype=F64, device=zeta0.device)
        for i0 in range(1, n_tstep, self.k):
            i1 = min(n_tstep, i0 + self.k)
            out = checkpoint(chunk, i0, i1, *state, use_reentrant=False)
            state, J = out[:4], J + out[4]
        return J


def gradient(model: DifferentiableUVLM, twist, u_inf, zeta0, hg_t, hgd_t):
    """(J, dJ/dtwist, dJ/du_inf) by the adjoint (reverse) pass."""
    tw = twist.detach().clone().requires_grad_(True)
    u = u_inf.detach().clone().requires_grad_(True)
    J = model.objective(tw, u, zeta0, hg_t, hgd_t)
    g_tw, g_u = torch.autograd.grad(J, (tw, u))
    return J.detach(), g_tw, g_u
