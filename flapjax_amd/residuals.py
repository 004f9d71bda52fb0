"""Discrete-adjoint residual layer of the UVLM (aero/uvlm.py:1529-2047) and a
matrix-free adjoint driver with GMRES and the frozen-wake preconditioner
(coupled/gradients/coupled.py:848-990, 1234-1249), on the device.

The residuals are the reference's, written with the differentiable device
pieces (``autograd.PairSum`` for every Biot-Savart sum, ``autograd.aic_solve``
for the circulation solve, torch stencils), so reverse passes through them
run the adjoint Biot-Savart kernels and the transposed LU solve:

  gamma_b_res    A_b(zeta_b)^-1 [-(U - c_dot + v_w(c; zeta_w, gamma_w)) . n] - gamma_b
  wake_prop_res  W_zeta(zeta_b,n, zeta_w,n-1, ...) - zeta_w,n,  W_gamma(gamma_b,n-1, gamma_w,n-1) - gamma_w,n
  gamma_b_dot_res  g/dt (gamma_b,n - gamma_b,n-1) + (1 - g) gamma_dot,n-1 - gamma_dot,n
  f_aero_res     R^T project(f_steady + f_unsteady) - f_aero_beam,n   (local frame, solve DOFs)

The beam poses enter as hg / hg_dot (the reference maps its varphi / v to
them with the beam's own functions first).  Vectors are the reference's
ravelled per-surface lists (surfaces concatenated).

``PrescribedAdjoint`` solves the discrete adjoint of a prescribed-motion run
with these residuals: per time step, backwards, the transposed step Jacobian
system (dR_n/dq_n)^T lambda_n = -(dj_n/dq_n)^T - (dR_n+1/dq_n)^T lambda_n+1 by
GMRES on reverse passes, preconditioned by the LU of the Jacobian without its
wake blocks (the reference's frozen-wake preconditioner), and assembles
dJ/dx = dj/dx + sum lambda_n^T dR_n/dx plus the static initial state's part.
"""

from __future__ import annotations

import numpy as np
import torch

from . import autograd as AG
from .adjoint import DifferentiableUVLM, apply_hg, colloc, normals, split_to_vertex, twisted_grid
from .coupled_adjoint import flow_at

F64 = torch.float64


def _cross(a, b):
    return torch.linalg.cross(a, b, dim=-1)


def _split(vec, shapes):
    out, off = [], 0
    for s in shapes:
        k = int(np.prod(s))
        out.append(vec[off:off + k].reshape(s))
        off += k
    return out


def _cat(lst):
    return torch.cat([a.reshape(-1) for a in lst]) if lst else torch.zeros(0, dtype=F64)


def steady_forces(zb, zbd, gb, gw, rho, v_func):
    """calculate_steady_forcing (aero/utils.py:126-223), differentiable."""
    m, n = gb.shape
    mid_c = 0.5 * (zb[:-1] + zb[1:])
    mid_s = 0.5 * (zb[:, :-1] + zb[:, 1:])
    v = v_func(torch.cat((mid_c.reshape(-1, 3), mid_s.reshape(-1, 3))))
    vc = v[: m * (n + 1)].reshape(m, n + 1, 3) - 0.5 * (zbd[:-1] + zbd[1:])
    vs = v[m * (n + 1):].reshape(m + 1, n, 3) - 0.5 * (zbd[:, :-1] + zbd[:, 1:])
    pad = torch.nn.functional.pad
    gc = pad(gb, (0, 1)) - pad(gb, (1, 0))
    gs = pad(-gb, (0, 0, 0, 1)) + pad(gb, (0, 0, 1, 0))
    if gw.shape[0] > 0:
        gs = gs - pad(gw[:1], (0, 0, m, 0))
    fc = rho * gc[..., None] * _cross(vc, zb[1:] - zb[:-1])
    fs = rho * gs[..., None] * _cross(vs, zb[:, 1:] - zb[:, :-1])
    return split_to_vertex(fc, 0) + split_to_vertex(fs, 1)


class UVLMResiduals:
    """The residual functions of a ``UVLM`` (aero/uvlm.py:1529-2047).

    ``zeta_b0`` (list of local grids) and ``flow`` (callable (x, t) -> U)
    default to the UVLM's design variables; passing differentiable ones
    makes the residuals differentiable in them (design gradients)."""

    def __init__(self, uvlm):
        if any(dw is not None for dw in getattr(uvlm, "delta_w", [None])):
            raise NotImplementedError("variable wake discretisation is not on the residual device path")
        if getattr(uvlm, "grid_func", None) is not None:
            raise NotImplementedError("control surfaces are not on the residual device path")
        self.uvlm = uvlm
        self.kernel = uvlm.kernel
        self.mirror = uvlm.mirror
        self.dt = uvlm.dt
        self.relax = uvlm.gamma_dot_relaxation
        self.free_wake = uvlm.free_wake
        self.unsteady = uvlm.include_unsteady_force
        self.gd = uvlm.grid_disc
        self.dof = [torch.as_tensor(d, dtype=torch.long, device=uvlm.hg_ref.device) for d in uvlm._dof_host]
        self.shapes_b = [(g.m, g.n) for g in self.gd]
        self.shapes_w = [(g.m_star, g.n) for g in self.gd]
        self.shapes_zw = [(g.m_star + 1, g.n + 1, 3) for g in self.gd]

    # ------------------------------------------------------------- helpers
    def _flow(self, flow):
        if flow is not None:
            return flow
        ff = self.uvlm.flowfield
        return lambda x, t: flow_at(ff, x, t)

    def _grids(self, hg, zeta_b0):
        z0 = self.uvlm.zeta_b0 if zeta_b0 is None else zeta_b0
        return [apply_hg(hg[d], z) for d, z in zip(self.dof, z0)]

    # --------------------------------------------------------- residuals
    def gamma_b_res_func(self, t_n, hg_n, hg_dot_n, gamma_b_n, gamma_w_n, zeta_w_n, zeta_b0=None, flow=None):
        """aero/uvlm.py:1529-1642."""
        flow = self._flow(flow)
        zb = self._grids(hg_n, zeta_b0)
        zbd = self._grids(hg_dot_n, zeta_b0) if hg_dot_n is not None else [torch.zeros_like(z) for z in zb]
        # hg_dot acts on [zeta0; 1] (aero/uvlm.py:562-621): translation rate included by apply_hg
        gw = _split(gamma_w_n, self.shapes_w)
        zw = _split(zeta_w_n, self.shapes_zw)
        cs = [colloc(z).reshape(-1, 3) for z in zb]
        ns = [normals(z).reshape(-1, 3) for z in zb]
        c, nc = torch.cat(cs), torch.cat(ns)
        vbc = torch.cat([(flow(colloc(z), t_n) - colloc(zd)).reshape(-1, 3) for z, zd in zip(zb, zbd)])
        vbc = vbc + AG.v_ind(c, zw, gw, self.kernel, self.mirror)
        g = AG.aic_solve(-(vbc * nc).sum(-1), c, nc, zb, None, self.kernel, self.mirror)
        return g - gamma_b_n

    def wake_prop_res_func(self, t_n, hg_nm1, hg_n, gamma_b_nm1, gamma_w_nm1, gamma_w_n, zeta_w_nm1, zeta_w_n,
                           zeta_b0=None, flow=None):
        """aero/uvlm.py:1644-1751: (zeta_w residual, gamma_w residual)."""
        flow = self._flow(flow)
        zb_nm1 = self._grids(hg_nm1, zeta_b0)
        zb_n = self._grids(hg_n, zeta_b0)
        gb1 = _split(gamma_b_nm1, self.shapes_b)
        gw1 = _split(gamma_w_nm1, self.shapes_w)
        zw1 = _split(zeta_w_nm1, self.shapes_zw)
        zs, gs = [], []
        for s in range(len(self.gd)):
            base = zw1[s][:-1]
            v = flow(base, t_n)
            if self.free_wake and base.numel():
                v = v + AG.v_ind(base.reshape(-1, 3), zb_nm1 + zw1, gb1 + gw1, self.kernel,
                                 self.mirror).reshape(base.shape)
            zs.append(torch.cat((zb_n[s][-1:], base + self.dt * v), dim=0))
            gs.append(torch.cat((gb1[s][-1:], gw1[s][:-1]), dim=0))
        return _cat(zs) - zeta_w_n, _cat(gs) - gamma_w_n

    def gamma_b_dot_res_func(self, gamma_b_nm1, gamma_b_n, gamma_b_dot_nm1, gamma_b_dot_n):
        """aero/uvlm.py:1753-1801."""
        g = self.relax
        return g / self.dt * (gamma_b_n - gamma_b_nm1) + (1.0 - g) * gamma_b_dot_nm1 - gamma_b_dot_n

    def forces(self, t_n, hg_n, hg_dot_n, gamma_b_n, gamma_w_n, gamma_b_dot_n, zeta_w_n, zeta_b0=None, flow=None,
               block_grid_gradients=False):
        """Steady + unsteady vertex forces of a state (aero/uvlm.py:1870-1905)."""
        flow = self._flow(flow)
        zb = self._grids(hg_n, zeta_b0)
        zbd = self._grids(hg_dot_n, zeta_b0)
        zw = _split(zeta_w_n, self.shapes_zw)
        if block_grid_gradients:
            zb = [z.detach() for z in zb]
            zw = [z.detach() for z in zw]
        gb = _split(gamma_b_n, self.shapes_b)
        gw = _split(gamma_w_n, self.shapes_w)
        gbd = _split(gamma_b_dot_n, self.shapes_b)
        rho = self.uvlm.flowfield.rho

        def v_tot(x):
            return flow(x, t_n) + AG.v_ind(x, zb + zw, gb + gw, self.kernel, self.mirror)

        out = []
        for s in range(len(self.gd)):
            f = steady_forces(zb[s], zbd[s], gb[s], gw[s], rho, v_tot)
            if self.unsteady:
                f = f + split_to_vertex(split_to_vertex(rho * gbd[s][..., None] * normals(zb[s]), 0), 1)
            out.append(f)
        return out, zb

    def f_aero_res_func(self, t_n, hg_n, hg_dot_n, gamma_b_n, gamma_w_n, gamma_b_dot_n, zeta_w_n, f_aero_beam_n,
                        block_grid_gradients=False, solve_dofs=None, zeta_b0=None, flow=None):
        """aero/uvlm.py:1803-1940: the beam-node force of the state in the
        local frame, minus the structural state's f_aero (solve DOFs)."""
        f, _ = self.forces(t_n, hg_n, hg_dot_n, gamma_b_n, gamma_w_n, gamma_b_dot_n, zeta_w_n, zeta_b0, flow,
                           block_grid_gradients)
        z0 = self.uvlm.zeta_b0 if zeta_b0 is None else zeta_b0
        n_nodes = hg_n.shape[0]
        glob = torch.zeros((n_nodes, 6), dtype=F64, device=hg_n.device)
        for fs, d, x0 in zip(f, self.dof, z0):
            r = hg_n[d][:, :3, :3]
            arm = torch.einsum("jab,ijb->ija", r, x0)
            glob = glob.index_put((d,), torch.cat((fs.sum(0), _cross(arm, fs).sum(0)), -1))
        rt = hg_n[:, :3, :3].transpose(-1, -2)
        loc = torch.cat(((rt @ glob[:, :3, None])[..., 0], (rt @ glob[:, 3:, None])[..., 0]), -1)
        res = loc.reshape(-1) - f_aero_beam_n.reshape(-1)
        return res if solve_dofs is None else res[torch.as_tensor(solve_dofs, device=res.device)]

    def timestep_residual(self, t_n, hg_nm1, hg_n, hg_dot_n, q_n, q_nm1, f_aero_beam_n=None, approx_grads=False,
                          zeta_b0=None, flow=None):
        """aero/uvlm.py:1942-2047: [gamma_b, gamma_w, gamma_b_dot, zeta_w(, f_aero)]
        residuals; q_n / q_nm1 are AeroStates with flat tensors (or per-surface lists)."""
        qn = {k: _cat(getattr(q_n, k)) if isinstance(getattr(q_n, k), (list, tuple)) else getattr(q_n, k)
              for k in ("gamma_b", "gamma_w", "gamma_b_dot", "zeta_w")}
        qp = {k: _cat(getattr(q_nm1, k)) if isinstance(getattr(q_nm1, k), (list, tuple)) else getattr(q_nm1, k)
              for k in ("gamma_b", "gamma_w", "gamma_b_dot", "zeta_w")}
        zr, gr = self.wake_prop_res_func(t_n, hg_nm1, hg_n, qp["gamma_b"], qp["gamma_w"], qn["gamma_w"],
                                         qp["zeta_w"], qn["zeta_w"], zeta_b0, flow)
        parts = [self.gamma_b_res_func(t_n, hg_n, hg_dot_n, qn["gamma_b"], qn["gamma_w"], qn["zeta_w"], zeta_b0, flow),
                 gr, self.gamma_b_dot_res_func(qp["gamma_b"], qn["gamma_b"], qp["gamma_b_dot"], qn["gamma_b_dot"]), zr]
        if f_aero_beam_n is not None:
            parts.append(self.f_aero_res_func(t_n, hg_n, hg_dot_n, qn["gamma_b"], qn["gamma_w"], qn["gamma_b_dot"],
                                              qn["zeta_w"], f_aero_beam_n, approx_grads, None, zeta_b0, flow))
        return torch.cat(parts)


# ---------------------------------------------------------------------------
# GMRES (restarted, right-preconditioned), jax.scipy.sparse.linalg.gmres's role
# ---------------------------------------------------------------------------
def gmres(matvec, b, x0=None, tol=1e-10, atol=0.0, restart=50, maxiter=50, precond=None):
    """Solve A x = b with restarted GMRES (Saad & Schultz 1986: modified
    Gram-Schmidt Arnoldi, Givens rotations), right preconditioned by
    ``precond`` (A M y = b, x = M y).  Returns (x, Arnoldi steps, relative
    residual |b - A x| / |b|)."""
    M = precond if precond is not None else (lambda v: v)
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    bnorm = float(torch.linalg.norm(b))
    if bnorm == 0.0:
        return torch.zeros_like(b), 0, 0.0
    thr = max(tol * bnorm, atol)
    total = 0
    r = b - matvec(x)
    beta = float(torch.linalg.norm(r))
    for _ in range(maxiter):
        if beta <= thr:
            break
        V, Z = [r / beta], []
        H = np.zeros((restart + 1, restart))
        cs, sn = np.zeros(restart), np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        k_used = 0
        for k in range(restart):
            z = M(V[k])
            Z.append(z)
            w = matvec(z)
            for i in range(k + 1):
                H[i, k] = float(torch.dot(w, V[i]))
                w = w - H[i, k] * V[i]
            hn = float(torch.linalg.norm(w))
            H[k + 1, k] = hn
            for i in range(k):
                t = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
                H[i + 1, k] = -sn[i] * H[i, k] + cs[i] * H[i + 1, k]
                H[i, k] = t
            den = float(np.hypot(H[k, k], H[k + 1, k]))
            cs[k], sn[k] = (1.0, 0.0) if den == 0.0 else (H[k, k] / den, H[k + 1, k] / den)
            H[k, k], H[k + 1, k] = den, 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            k_used = k + 1
            total += 1
            if abs(g[k + 1]) <= thr or hn == 0.0:
                break
            V.append(w / hn)
        y = np.linalg.solve(np.triu(H[:k_used, :k_used]), g[:k_used])
        for i in range(k_used):
            x = x + float(y[i]) * Z[i]
        r = b - matvec(x)
        beta = float(torch.linalg.norm(r))
    return x, total, beta / bnorm


# ---------------------------------------------------------------------------
# discrete adjoint of a prescribed-motion run
# ---------------------------------------------------------------------------
class PrescribedAdjoint:
    """dJ/d(twist, U_inf) of J = dt sum_n (lift_n) over a prescribed-motion
    single-surface run, by the discrete adjoint of the timestep residuals
    (coupled/gradients/coupled.py:993-1384 restricted to the aerodynamic
    states): per step a GMRES solve of the transposed step Jacobian on
    reverse passes, preconditioned by the frozen-wake LU."""

    def __init__(self, uvlm, gmres_tol=1e-12, restart=30, precond=True, i_ts_preconditioner=1):
        if uvlm.n_surf != 1 or uvlm.mirror is not None:
            raise NotImplementedError("one unmirrored surface (the twist design variable is per station of it)")
        self.uvlm = uvlm
        self.res = UVLMResiduals(uvlm)
        gd = uvlm.grid_disc[0]
        self.m, self.n, self.ms = gd.m, gd.n, gd.m_star
        self.model = DifferentiableUVLM(gd.m, gd.n, gd.m_star, uvlm.dt, uvlm.flowfield.rho, uvlm.free_wake,
                                        uvlm.gamma_dot_relaxation, uvlm.kernel)
        self.tol, self.restart, self.use_precond = gmres_tol, restart, precond
        self.i_pre = i_ts_preconditioner
        self.gmres_iters = []

    # state vector q = [gamma_b, gamma_w, gamma_b_dot, zeta_w] (the reference's order)
    def _sizes(self):
        m, n, ms = self.m, self.n, self.ms
        return [m * n, ms * n, m * n, 3 * (ms + 1) * (n + 1)]

    def _pack(self, st):
        zw, gw, gb, gbd = st
        return torch.cat((gb.reshape(-1), gw.reshape(-1), gbd.reshape(-1), zw.reshape(-1)))

    def _unpack(self, q):
        a, b, c, _ = self._sizes()
        return q[:a], q[a:a + b], q[a + b:a + b + c], q[a + b + c:]

    def _flow(self, u_inf):
        return lambda x, t: u_inf.expand_as(x)

    def _residual(self, i, q_n, q_nm1, z0, u_inf, hg_t, hgd_t, t):
        from .data_structures import AeroStates

        a = self._unpack(q_n)
        b = self._unpack(q_nm1)
        return self.res.timestep_residual(t[i], hg_t[i - 1], hg_t[i], hgd_t[i], AeroStates(*a), AeroStates(*b),
                                          zeta_b0=[z0], flow=self._flow(u_inf))

    def _objective(self, q, z0, u_inf, hg, hgd, t):
        gb, gw, gbd, zw = self._unpack(q)
        f, _ = self.res.forces(t, hg, hgd, gb, gw, gbd, zw, [z0], self._flow(u_inf))
        return self.uvlm.dt * f[0][..., 2].sum()

    def _preconditioner(self, q_n, q_nm1, i, z0, u_inf, hg_t, hgd_t, t):
        """Frozen-wake preconditioner (coupled/gradients/coupled.py:848-990):
        the Jacobian of the residual over the non-wake states (gamma_b,
        gamma_b_dot) with the wake blocks replaced by -I; LU of its transpose."""
        a, b, c, d = self._sizes()
        keep = torch.cat((torch.arange(a), torch.arange(a + b, a + b + c))).to(q_n.device)
        q = q_n.detach().clone().requires_grad_(True)
        with torch.enable_grad():
            r = self._residual(i, q, q_nm1, z0, u_inf, hg_t, hgd_t, t)
            rows = []
            for k in keep.tolist():
                (gq,) = torch.autograd.grad(r[k], q, retain_graph=True)
                rows.append(gq[keep])
        jac = torch.stack(rows)  # d r_keep / d q_keep
        lu = torch.linalg.lu_factor(jac.T.contiguous())

        def apply(v):
            out = -v.clone()
            out[keep] = torch.linalg.lu_solve(lu[0], lu[1], v[keep][:, None])[:, 0]
            return out

        return apply

    def gradient(self, twist, u_inf, zeta0, hg_t, hgd_t):
        dev = zeta0.device
        hg_t = hg_t.to(dev)
        hgd_t = hgd_t.to(dev)
        n_t = hg_t.shape[0]
        t = np.arange(n_t) * self.uvlm.dt
        tw = twist.detach().clone().requires_grad_(True)
        uu = u_inf.detach().clone().requires_grad_(True)
        # forward: the static state and the steps (no graph kept)
        with torch.no_grad():
            z0 = twisted_grid(zeta0, tw)
            st = self.model.static(apply_hg(hg_t[0], z0), uu)
            st = (st[0], st[1], st[2], torch.zeros_like(st[2]))
            qs = [self._pack(st)]
            for i in range(1, n_t):
                st, _ = self.model.step(st, apply_hg(hg_t[i], z0), apply_hg(hg_t[i - 1], z0),
                                        apply_hg(hgd_t[i], z0), uu)
                qs.append(self._pack(st))
        J = 0.0
        g_tw = torch.zeros_like(tw)
        g_u = torch.zeros_like(uu)
        lam_next = None
        precond = None
        self.gmres_iters = []
        for i in range(n_t - 1, 0, -1):
            # right-hand side: -(dj_i/dq_i) - (dR_i+1/dq_i)^T lam_i+1
            q_i = qs[i].detach().clone().requires_grad_(True)
            with torch.enable_grad():
                z0 = twisted_grid(zeta0, tw)
                j_i = self._objective(q_i, z0, uu, hg_t[i], hgd_t[i], t[i])
                gj = torch.autograd.grad(j_i, (q_i, tw, uu))
            J += float(j_i.detach())
            rhs = -gj[0]
            g_tw += gj[1]
            g_u += gj[2]
            if lam_next is not None:
                q_i2 = qs[i].detach().clone().requires_grad_(True)
                with torch.enable_grad():
                    r_next = self._residual(i + 1, qs[i + 1], q_i2, twisted_grid(zeta0, tw), uu, hg_t, hgd_t, t)
                    (gq,) = torch.autograd.grad(r_next, q_i2, grad_outputs=lam_next)
                rhs = rhs - gq
            # the transposed Jacobian of R_i in q_i, matrix-free
            q_v = qs[i].detach().clone().requires_grad_(True)
            with torch.enable_grad():
                r_i = self._residual(i, q_v, qs[i - 1], twisted_grid(zeta0, tw).detach(), uu.detach(), hg_t, hgd_t, t)

            def matvec_t(v, r_i=r_i, q_v=q_v):
                (g,) = torch.autograd.grad(r_i, q_v, grad_outputs=v, retain_graph=True)
                return g

            if self.use_precond and precond is None:
                precond = self._preconditioner(qs[i], qs[i - 1], i, twisted_grid(zeta0, tw).detach(), uu.detach(),
                                               hg_t, hgd_t, t)
            lam, iters, relres = gmres(matvec_t, rhs, x0=lam_next, tol=self.tol, restart=self.restart,
                                       precond=precond if self.use_precond else None)
            self.gmres_iters.append((iters, relres))
            # design part: lam^T dR_i/dx
            with torch.enable_grad():
                r_x = self._residual(i, qs[i], qs[i - 1], twisted_grid(zeta0, tw), uu, hg_t, hgd_t, t)
                gx = torch.autograd.grad(r_x, (tw, uu), grad_outputs=lam, allow_unused=True)
            g_tw += gx[0] if gx[0] is not None else 0.0
            g_u += gx[1] if gx[1] is not None else 0.0
            lam_next = lam
        # initial state: q_0(x) from the static solve, weighted by (dR_1/dq_0)^T lam_1
        if lam_next is not None:
            q0 = qs[0].detach().clone().requires_grad_(True)
            with torch.enable_grad():
                r1 = self._residual(1, qs[1], q0, twisted_grid(zeta0, tw), uu, hg_t, hgd_t, t)
                (w0,) = torch.autograd.grad(r1, q0, grad_outputs=lam_next)
                z0 = twisted_grid(zeta0, tw)
                st0 = self.model.static(apply_hg(hg_t[0], z0), uu)
                q0x = self._pack((st0[0], st0[1], st0[2], torch.zeros_like(st0[2])))
                g0 = torch.autograd.grad(q0x, (tw, uu), grad_outputs=w0, allow_unused=True)
            g_tw += g0[0] if g0[0] is not None else 0.0
            g_u += g0[1] if g0[1] is not None else 0.0
        return J, g_tw, g_u
