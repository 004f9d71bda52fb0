"""Gradients of the coupled aeroelastic solve: the counterpart of the
reference's coupled adjoint (coupled/gradients/coupled.py: static_adjoint
:195-353, dynamic_adjoint :993-1384) for a single lifting surface on a beam.

``DifferentiableAeroelastic`` reruns ``CoupledAeroelastic``'s static and
dynamic solves (coupled/coupled.py:154-433, structure/beam.py:2612-2725)
with differentiable pieces and lets torch's reverse pass produce the
discrete adjoint:

* aerodynamics: collocation points, normals, the static wake, wake
  convection, the right-hand side and the Kutta-Joukowski and unsteady forces
  as torch expressions of the vertex grids; every pair sum is
  ``autograd.PairSum`` (hand-written adjoint Biot-Savart kernels) and every
  circulation solve ``autograd.aic_solve`` (device AIC and LU, transposed LU
  solve in the reverse pass);
* structure: every Newton solve is ``beam_adjoint.StaticBeamSolve`` /
  ``DynamicBeamSolve`` (device Newton forward, implicit derivative backward);
* coupling: the force projection (aero/utils.py:666-697) and the
  generalised-alpha bookkeeping (structure/time_integration.py:53-150) as
  torch expressions.  The fluid-structure iterations are differentiated as
  they are run (the Aitken factor of the static iteration is held constant in
  the reverse pass, as a fixed-point derivative allows), so the gradient is
  that of the converged coupled solution when the iterations converge.

The time history is checkpointed every ``checkpoint_every`` steps and
recomputed in the reverse pass (torch.utils.checkpoint).

Design variables: one stiffness scaling per entry of the beam's k_cs table
and one twist angle per spanwise station after the root (the local section
rotated about the local y axis through the reference axis, structure-fixed),
as in the reference's adjoint example (models/ifasd_aviation_2026/
cantilever_wing_adjoint.py: k_cs and zeta_b0 as design variables).

Objective (``objective``): J = dt sum_n [lift_n + w_tip (z_tip,n - z_tip,0)^2]
over the dynamic steps, lift_n the vertical component of the total (steady +
unsteady) aerodynamic force on the surface at step n and z_tip the vertical
position of the last beam node.
"""

from __future__ import annotations

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from . import autograd as AG
from . import beam_adjoint as BA
from .adjoint import colloc, normals, split_to_vertex, twisted_grid
from .beam import ConvergenceStatus, TimeIntegrator
from .flowfields import OneMinusCosine
from .linalg import RefinedSolver

F64 = torch.float64


def _cross(a, b):
    return torch.linalg.cross(a, b, dim=-1)


def flow_at(ff, x, t):
    """The flowfield velocity at points x (..., 3) and time t as a torch
    expression (aero/flowfields.py:88-181)."""
    dev = x.device
    u = AG.device_const(ff.u_inf, dev)
    base = u.expand_as(x) if ff.relative_motion else torch.zeros_like(x)
    if isinstance(ff, OneMinusCosine):
        rel = x - AG.device_const(ff.gust_x0, dev)
        if ff.relative_motion:
            rel = rel - u * float(t)
        s = (rel * AG.device_const(ff.gust_travel_direction, dev)).sum(-1)
        amp = ff.gust_amplitude * 0.5 * (1.0 - torch.cos(np.pi * s / ff.gust_length))
        inside = (s > 0.0) & (s < 2.0 * ff.gust_length)
        amp = torch.where(inside, amp, torch.zeros_like(amp))
        base = base + amp[..., None] * AG.device_const(ff.gust_amplitude_direction, dev)
    elif type(ff).__name__ != "Constant":
        raise NotImplementedError(f"{type(ff).__name__} has no differentiable form")
    return base


class DifferentiableAeroelastic:
    """A ``CoupledAeroelastic`` case with one surface, rerun differentiably."""

    def __init__(self, coupled, prescribed_dofs, checkpoint_every: int = 10, aic_reuse: bool | None = None):
        """checkpoint_every: steps per recomputed chunk of the time history
        (0: keep the whole history, no recomputation; the history of a few
        hundred steps of a 16 x 80 case is a few tens of GB).  aic_reuse: solve
        the circulation systems with kept inverses and FP64 refinement, as
        the production solve does (None: as the case's UVLM is set)."""
        aero, beam = coupled.aero, coupled.structure
        if aero.n_surf != 1:
            raise NotImplementedError("the coupled adjoint takes a single lifting surface")
        if any(d is not None for d in aero.delta_w):
            raise NotImplementedError("the coupled adjoint takes a uniform wake discretisation")
        if aero.grid_func is not None:
            raise NotImplementedError("the coupled adjoint takes no control surfaces")
        gd = aero.grid_disc[0]
        if gd.m_star < 1:
            raise NotImplementedError("the coupled adjoint needs a wake")
        self.cw, self.aero, self.beam = coupled, aero, beam
        self.m, self.n, self.m_star = gd.m, gd.n, gd.m_star
        self.dof = torch.as_tensor(aero._dof_host[0], dtype=torch.long, device=beam.hg0.device)
        if len(set(aero._dof_host[0].tolist())) != self.n + 1:
            raise NotImplementedError("the coupled adjoint takes one beam node per spanwise station")
        self.zeta0 = aero.zeta_b0[0]
        self.mirror = aero.mirror
        self.kernel = aero.kernel
        self.ff = aero.flowfield
        self.dt = float(aero.dt)
        self.rho = float(self.ff.rho)
        self.relax = float(aero.gamma_dot_relaxation)
        self.unsteady = bool(aero.include_unsteady_force)
        self.free_wake = bool(aero.free_wake)
        self.ti = TimeIntegrator(beam.spectral_radius, self.dt)
        self.solver = BA.BeamSolver(beam, prescribed_dofs, self.ti)
        self.k = int(checkpoint_every)
        reuse = aero.aic_reuse if aic_reuse is None else bool(aic_reuse)
        # up to 8 corrections per solve, the limit the production solve moves
        # to once a deforming wing outgrows 3 (uvlm.UVLM._run_checked): one
        # kept inverse then serves the whole run instead of ~1 refactorisation
        # per step
        self.aic_solvers = ({True: RefinedSolver(max_iter=8), False: RefinedSolver(max_iter=8)} if reuse
                            else {True: None, False: None})
        free = torch.ones(beam.n_dof, dtype=torch.bool, device=beam.hg0.device)
        if self.solver.prescribed:
            free[torch.as_tensor(self.solver.prescribed, device=free.device)] = False
        self.free = free.view(-1, 6)
        u = self.ff.u_inf_dir * self.dt * self.ff.u_inf_mag
        self.wake_s = torch.as_tensor(np.arange(self.m_star + 1)[:, None] * u, dtype=F64, device=beam.hg0.device)
        self.fsi_iters = []

    # ------------------------------------------------------------ geometry
    def section_grid(self, twist):
        """Local grid with a twist angle per station (station 0 untwisted)."""
        tw = torch.cat((torch.zeros(1, dtype=F64, device=twist.device), twist))
        return twisted_grid(self.zeta0, tw)

    def grid(self, hg, z0):
        """aero/uvlm.py:537-560: R z0 + x per station."""
        g = hg[self.dof]
        return torch.einsum("jab,ijb->ija", g[:, :3, :3], z0) + g[None, :, :3, 3]

    # ---------------------------------------------------------------- aero
    def forces(self, zb, zbd, gb, gw, zw, t):
        """Kutta-Joukowski vertex forces (aero/utils.py:126-223)."""
        m, n = self.m, self.n
        mid_c = 0.5 * (zb[:-1] + zb[1:])
        mid_s = 0.5 * (zb[:, :-1] + zb[:, 1:])
        mids = torch.cat((mid_c.reshape(-1, 3), mid_s.reshape(-1, 3)))
        v = AG.v_ind(mids, [zb, zw], [gb, gw], self.kernel, self.mirror) + flow_at(self.ff, mids, t)
        vc = v[: m * (n + 1)].reshape(m, n + 1, 3)
        vs = v[m * (n + 1):].reshape(m + 1, n, 3)
        if zbd is not None:
            vc = vc - 0.5 * (zbd[:-1] + zbd[1:])
            vs = vs - 0.5 * (zbd[:, :-1] + zbd[:, 1:])
        pad = torch.nn.functional.pad
        gc = pad(gb, (0, 1)) - pad(gb, (1, 0))
        gs = pad(-gb, (0, 0, 0, 1)) + pad(gb, (0, 0, 1, 0)) - pad(gw[:1], (0, 0, m, 0))
        fc = self.rho * gc[..., None] * _cross(vc, zb[1:] - zb[:-1])
        fs = self.rho * gs[..., None] * _cross(vs, zb[:, 1:] - zb[:, :-1])
        return split_to_vertex(fc, 0) + split_to_vertex(fs, 1)

    def aero_static(self, zb, t):
        """Static solve with the flat wake lumped onto the trailing edge
        (aero/uvlm.py:895-1012): state (zeta_w, gamma_w, gamma_b, gamma_dot)
        and the steady vertex forces."""
        m, n, ms = self.m, self.n, self.m_star
        zw = zb[-1][None] + self.wake_s[:, None, :]
        c, nc = colloc(zb), normals(zb)
        b = -(flow_at(self.ff, c, t) * nc).sum(-1).reshape(-1)
        gb = AG.aic_solve(b, c.reshape(-1, 3), nc.reshape(-1, 3), [zb], [zw], self.kernel, self.mirror,
                          self.aic_solvers[True]).reshape(m, n)
        gw = gb[-1:].expand(ms, n)
        return (zw, gw, gb, torch.zeros_like(gb)), self.forces(zb, None, gb, gw, zw, t)

    def aero_step(self, st, zb_n, zb_nm1, zbd_n, t_n):
        """One dynamic step (aero/uvlm.py:888-1103): wake shift and
        convection, bound-only AIC with the wake upwash on the right-hand
        side, filtered circulation rate, steady and unsteady forces."""
        zw, gw, gb, gbd = st
        m, n = self.m, self.n
        base = zw[:-1]
        v = flow_at(self.ff, base, t_n)
        if self.free_wake:
            v = v + AG.v_ind(base.reshape(-1, 3), [zb_nm1, zw], [gb, gw], self.kernel, self.mirror).reshape(base.shape)
        zw_n = torch.cat((zb_n[-1:], base + self.dt * v), dim=0)
        gw_n = torch.cat((gb[-1:], gw[:-1]), dim=0)
        c, nc, cd = colloc(zb_n), normals(zb_n), colloc(zbd_n)
        vbc = flow_at(self.ff, c, t_n) - cd + AG.v_ind(c.reshape(-1, 3), [zw_n], [gw_n], self.kernel,
                                                        self.mirror).reshape(c.shape)
        b = -(vbc * nc).sum(-1).reshape(-1)
        gb_n = AG.aic_solve(b, c.reshape(-1, 3), nc.reshape(-1, 3), [zb_n], None, self.kernel,
                            self.mirror, self.aic_solvers[False]).reshape(m, n)
        gbd_n = (gb_n - gb) / self.dt * self.relax + gbd * (1.0 - self.relax)
        f_st = self.forces(zb_n, zbd_n, gb_n, gw_n, zw_n, t_n)
        f_un = split_to_vertex(split_to_vertex(self.rho * gbd_n[..., None] * nc, 0), 1)
        return (zw_n, gw_n, gb_n, gbd_n), f_st, f_un

    # ------------------------------------------------------------ coupling
    def project(self, f, hg, z0):
        """aero/utils.py:666-697: per beam node the strip's force and its
        moment about the node, global frame."""
        r = hg[self.dof][:, :3, :3]
        arm = torch.einsum("jab,ijb->ija", r, z0)
        out = torch.zeros((hg.shape[0], 6), dtype=F64, device=hg.device)
        return out.index_put((self.dof,), torch.cat((f.sum(0), _cross(arm, f).sum(0)), -1))

    @staticmethod
    def _to_global(f, hg):
        r = hg[..., :3, :3]
        return torch.cat(((r @ f[..., :3, None])[..., 0], (r @ f[..., 3:, None])[..., 0]), -1)

    def static(self, k_scale, z0, t=0.0, fsi_relaxation=0.6, fsi_omega_min=1e-2):
        """CoupledAeroelastic.static_solve (coupled/coupled.py:154-305): the
        Aitken-relaxed fixed point of aero forces and static beam solves."""
        hg0 = self.beam.hg0
        hg = hg0
        f_app_local = torch.zeros((hg0.shape[0], 6), dtype=F64, device=hg0.device)
        ast, f_st = self.aero_static(self.grid(hg, z0), t)
        status = ConvergenceStatus(self.cw.fsi_convergence_settings)
        omega_prev = float(fsi_relaxation)
        r_prev = torch.zeros_like(f_app_local)
        zero = torch.zeros_like(f_app_local)
        while not status.get_status():
            f_new = self.project(f_st, hg, z0)
            f_app = self._to_global(f_app_local, hg)
            r_curr = torch.where(self.free, f_new - f_app, zero)
            dr = (r_curr - r_prev).detach()
            den, num = (float(v) for v in torch.stack((torch.vdot(dr.reshape(-1), dr.reshape(-1)),
                                                       torch.vdot(r_prev.detach().reshape(-1), dr.reshape(-1)))))
            om_a = min(max(-omega_prev * num / (den if den > 0.0 else 1.0), fsi_omega_min), 1.0)
            omega = om_a if status.i_iter >= 2 else float(fsi_relaxation)
            f_aero = f_new if status.i_iter == 0 else f_app + omega * r_curr
            hg_new = BA.StaticBeamSolve.apply(f_aero, k_scale, self.solver)
            with torch.no_grad():
                rel = BA.log_se3(BA.hg_inv(hg) @ hg_new)
                tot = BA.log_se3(BA.hg_inv(hg_new) @ hg0)
                delta = rel if status.i_iter == 0 else rel / omega
                nv = torch.stack((delta.norm(), tot.norm(), r_curr.norm(), torch.where(self.free, f_new, zero).abs().max(),
                                  torch.isnan(delta).any().to(F64))).cpu().numpy()
            status.update_norms(*nv[:4], nv[4] != 0.0)
            hg = hg_new
            f_app_local = BA.dead_to_local(f_aero, hg_new)
            ast, f_st = self.aero_static(self.grid(hg, z0), t)
            omega_prev, r_prev = omega, r_curr
        self.fsi_iters.append(status.i_iter)
        return hg, ast, f_st

    def dynamic_step(self, i, k_scale, z0, hg_p, v_p, vd_p, acc_p, f_g_p, ast):
        """One generalised-alpha step with its FSI iterations
        (structure/beam.py:2612-2725 with the hooks of coupled/coupled.py:
        355-420); returns the step-i state."""
        ti, sv = self.ti, self.solver
        af = ti.alpha_f
        t_n = i * self.dt + self.t0
        phi_a, v_a, vd_a = ti.predict_alpha(v_p, vd_p, acc_p)
        f_nm1 = f_g_p
        f_prev = f_nm1
        zb_nm1 = self.grid(hg_p, z0)
        status = ConvergenceStatus(self.cw.fsi_convergence_settings)
        while not status.get_status():
            phi_n = phi_a / (1.0 - af)
            v_n = (v_a - af * v_p) / (1.0 - af)
            hg_n = hg_p @ BA.exp_se3(phi_n)
            hg_dot = hg_n @ BA.tilde4(v_n)
            st_n, f_st, f_un = self.aero_step(ast, self.grid(hg_n, z0), zb_nm1, self.grid(hg_dot, z0), t_n)
            f_n = self.project(f_st + f_un if self.unsteady else f_st, hg_n, z0)
            f_al = ti.f_alpha(f_nm1, f_n)
            phi_s = BA.DynamicBeamSolve.apply(hg_p, phi_a, v_a, vd_a, f_al, k_scale, sv)
            phi_new, v_a, vd_a = BA.DynamicBeamSolve.full_state(sv, phi_s, phi_a, v_a, vd_a)
            with torch.no_grad():
                dd = sv.reduce(phi_a - phi_new)
                df = sv.reduce(f_al - f_prev)
                nv = torch.stack((dd.norm(), sv.reduce(phi_new).norm(), df.norm(), sv.reduce(f_al).abs().max(),
                                  torch.isnan(dd).any().to(F64))).cpu().numpy()
            status.update_norms(*nv[:4], nv[4] != 0.0)
            phi_a, f_prev = phi_new, f_al
        self.fsi_iters.append(status.i_iter)
        # calculate_q_n_from_q_alpha + calculate_a_n (structure/time_integration.py:53-84)
        v_i = (v_a - af * v_p) / (1.0 - af)
        vd_i = (vd_a - af * vd_p) / (1.0 - af)
        acc_i = ((1.0 - af) * vd_i + af * vd_p - ti.alpha_m * acc_p) / (1.0 - ti.alpha_m)
        hg_i = hg_p @ BA.exp_se3(phi_a / (1.0 - af))
        f_tot = f_st + f_un if self.unsteady else f_st
        f_g_i = self.project(f_tot, hg_i, z0)
        lift = f_st[..., 2].sum() + f_un[..., 2].sum()
        return hg_i, v_i, vd_i, acc_i, f_g_i, st_n, lift

    # ------------------------------------------------------------ objective
    def objective(self, k_scale, twist, n_tstep, tip_weight=0.0, t_init=0.0):
        """Static equilibrium, then n_tstep - 1 coupled steps;
        J = dt sum_n [lift_n + tip_weight (z_tip,n - z_tip,0)^2]."""
        self.t0 = float(t_init)
        self.fsi_iters = []
        z0 = self.section_grid(twist)
        hg, ast, f_st = self.static(k_scale, z0, self.t0)
        nn = hg.shape[0]
        zeros = torch.zeros((nn, 6), dtype=F64, device=hg.device)
        f_g = self.project(f_st, hg, z0)
        z_tip0 = hg[-1, 2, 3]
        state = (hg, zeros, zeros, zeros, f_g, *ast)

        def chunk(i0, i1, ks, z0_, *st):
            hg_, v_, vd_, acc_, fg_, *a_ = st
            a_ = tuple(a_)
            total = torch.zeros((), dtype=F64, device=hg_.device)
            for i in range(i0, i1):
                hg_, v_, vd_, acc_, fg_, a_, lift = self.dynamic_step(i, ks, z0_, hg_, v_, vd_, acc_, fg_, a_)
                total = total + self.dt * (lift + tip_weight * (hg_[-1, 2, 3] - z_tip0) ** 2)
            return (hg_, v_, vd_, acc_, fg_, *a_, total)

        if self.k <= 0:
            return chunk(1, n_tstep, k_scale, z0, *state)[-1]
        J = torch.zeros((), dtype=F64, device=hg.device)
        for i0 in range(1, n_tstep, self.k):
            i1 = min(n_tstep, i0 + self.k)
            out = checkpoint(chunk, i0, i1, k_scale, z0, *state, use_reentrant=False)
            state, J = out[:-1], J + out[-1]
        return J

    def gradient(self, k_scale, twist, n_tstep, tip_weight=0.0):
        """(J, dJ/dk_scale, dJ/dtwist) by the reverse pass."""
        ks = torch.as_tensor(k_scale, dtype=F64, device=self.beam.hg0.device).detach().clone().requires_grad_(True)
        tw = torch.as_tensor(twist, dtype=F64, device=self.beam.hg0.device).detach().clone().requires_grad_(True)
        before = getattr(self.beam, "stiffness_scaling", None)
        for sv in self.aic_solvers.values():
            if sv is not None:
                sv.adjoint_failures = None
        try:
            J = self.objective(ks, tw, n_tstep, tip_weight)
            g_k, g_t = torch.autograd.grad(J, (ks, tw))
            fails = [sv.adjoint_failures for sv in self.aic_solvers.values()
                     if sv is not None and sv.adjoint_failures is not None]
            if fails and int(sum(f.sum() for f in fails)) > 0:
                raise RuntimeError("a transposed circulation solve did not pass its refinement stop test")
        finally:
            # the solves rescale the beam's device tables; leave them as found
            self.beam.set_stiffness_scaling(np.ones(self.beam.k_cs.shape[0]) if before is None else before)
        return J.detach(), g_k, g_t
