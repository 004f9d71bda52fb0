"""Linearised UVLM about a steady reference (aero/linear/linear_uvlm.py:79-1039,
aero/linear/data_structures.py, utils/linear.py:680-822).

The discrete-time state-space model x_{n+1} = A x_n + B u_{n+1},
y_n = C x_n + D u_{n+1} is the Jacobian of one UVLM step about the
reference state, exactly as the reference builds it with jacrev of its
``step``: the step here is written with the differentiable device pieces
(``autograd.PairSum`` -- the adjoint Biot-Savart kernels -- and
``autograd.aic_solve`` -- the device LU and its transposed solve), and every
row of [A B; C D] is one reverse pass through it.

Inputs u (total values at step n+1): bound grid zeta_b, its velocity
zeta_b_dot, and optionally a bound-vertex upwash nu_b and a wake-vertex
upwash nu_w.  States x: gamma_b, gamma_w, gamma_b at n-1 (unsteady force),
and for a prescribed/free wake the wake grid zeta_w and the bound grid.
Outputs y: f_steady and f_unsteady vertex forces.  ``run`` marches the
matrices on the device with perturbation inputs and returns totals
(reference + perturbation), as the reference's AeroLinearResult.

Not built: variable wake discretisation (delta_w), control surfaces and the
optional input/output projections of ``linearise``.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Literal
from warnings import warn

import numpy as np
import torch

from . import autograd as AG
from .adjoint import colloc, normals, split_to_vertex
from .coupled_adjoint import flow_at
from .utils import as_segment_kernel, biot_savart_cutoff, to_device

F64 = torch.float64
LinearWakeType = Literal["frozen", "prescribed", "free"]


def _cross(a, b):
    return torch.linalg.cross(a, b, dim=-1)


# ---------------------------------------------------------------------------
# aero/linear/data_structures.py
# ---------------------------------------------------------------------------
@dataclass
class AeroInputUnflattened:
    zeta_b: list
    zeta_b_dot: list
    nu_b: list | None = None
    nu_w: list | None = None


@dataclass
class AeroStateUnflattened:
    gamma_b: list
    gamma_w: list
    gamma_b_nm1: list | None = None
    zeta_w: list | None = None
    zeta_b: list | None = None


@dataclass
class AeroOutputUnflattened:
    f_steady: list
    f_unsteady: list | None = None


@dataclass
class AeroLinearResult:
    """aero/linear/data_structures.py:38-80: perturbation and total histories."""

    reference: object
    u_t: AeroInputUnflattened
    x_t: AeroStateUnflattened
    y_t: AeroOutputUnflattened
    u_t_tot: AeroInputUnflattened
    x_t_tot: AeroStateUnflattened
    y_t_tot: AeroOutputUnflattened
    n_tstep: int
    n_surf: int
    t: np.ndarray
    surf_b_names: list
    surf_w_names: list


# ---------------------------------------------------------------------------
# utils/linear.py:680-822
# ---------------------------------------------------------------------------
class LinearSystem:
    """Discrete (or continuous) state-space matrices on the device."""

    def __init__(self, a, b, c, d, dt, continuous_time: bool = False, removed_u_np1: bool = False):
        self.a, self.b, self.c, self.d = a, b, c, d
        self.dt = float(dt)
        self.n_inputs = b.shape[1]
        self.n_states = a.shape[0]
        self.n_outputs = c.shape[0]
        self.continuous_time = continuous_time
        self.removed_u_np1 = removed_u_np1

    def remove_u_np1(self) -> None:
        """utils/linear.py:714-724: D <- C B + D, B <- A B."""
        if self.removed_u_np1:
            warn("u_np1 has already been removed from the system. Skipping.")
            return
        self.d = self.c @ self.b + self.d
        self.b = self.a @ self.b
        self.removed_u_np1 = True

    def continuous_to_discrete(self, method: str = "tustin") -> "LinearSystem":
        """utils/linear.py:726-773 (zero-order hold or Tustin)."""
        if not self.continuous_time:
            warn("System is already discrete-time.")
            return self
        n, k = self.a.shape[0], self.b.shape[1]
        eye = torch.eye(n, dtype=F64, device=self.a.device)
        if method == "zoh":
            aug = torch.zeros((n + k, n + k), dtype=F64, device=self.a.device)
            aug[:n, :n], aug[:n, n:] = self.a, self.b
            e = torch.linalg.matrix_exp(self.dt * aug)
            a_d, b_d = e[:n, :n], e[:n, n:]
        elif method == "tustin":
            inv = torch.linalg.inv(eye - 0.5 * self.dt * self.a)
            a_d, b_d = inv @ (eye + 0.5 * self.dt * self.a), inv @ (self.b * self.dt)
        else:
            raise ValueError(f"unknown discretisation method {method!r}")
        return LinearSystem(a_d, b_d, self.c, self.d, self.dt)

    def run(self, u, x0=None):
        """utils/linear.py:775-822: x_i = A x_{i-1} + B u_i (u_{i-1} once u_np1
        is removed), y_i = C x_i + D u_i; u (n_tstep, n_inputs)."""
        if self.continuous_time:
            raise ValueError("discretise the system first (continuous_to_discrete)")
        if u.dim() != 2 or u.shape[1] != self.n_inputs:
            raise ValueError(f"u must have shape (n_tstep, {self.n_inputs}), got {tuple(u.shape)}")
        n_t = u.shape[0]
        x = torch.zeros((n_t, self.n_states), dtype=F64, device=u.device)
        if x0 is not None:
            if tuple(x0.shape) != (self.n_states,):
                raise ValueError(f"x0 must have shape ({self.n_states},)")
            x[0] = x0
        bu = (u[:-1] if self.removed_u_np1 else u[1:]) @ self.b.T  # (n_t - 1, n_states)
        for i in range(1, n_t):
            x[i] = self.a @ x[i - 1] + bu[i - 1]
        y = x @ self.c.T + u @ self.d.T
        return x, y


# ---------------------------------------------------------------------------
# the differentiable step
# ---------------------------------------------------------------------------
def _forces(zb, zbd, gb, gw, rho, v_func, nu_b):
    """calculate_steady_forcing (aero/utils.py:126-223) with the additive
    vertex upwash of the linear model (v_inputs)."""
    m, n = gb.shape
    mid_c = 0.5 * (zb[:-1] + zb[1:])
    mid_s = 0.5 * (zb[:, :-1] + zb[:, 1:])
    v = v_func(torch.cat((mid_c.reshape(-1, 3), mid_s.reshape(-1, 3))))
    vc = v[: m * (n + 1)].reshape(m, n + 1, 3) - 0.5 * (zbd[:-1] + zbd[1:])
    vs = v[m * (n + 1):].reshape(m + 1, n, 3) - 0.5 * (zbd[:, :-1] + zbd[:, 1:])
    if nu_b is not None:
        vc = vc + 0.5 * (nu_b[:-1] + nu_b[1:])
        vs = vs + 0.5 * (nu_b[:, :-1] + nu_b[:, 1:])
    pad = torch.nn.functional.pad
    gc = pad(gb, (0, 1)) - pad(gb, (1, 0))
    gs = pad(-gb, (0, 0, 0, 1)) + pad(gb, (0, 0, 1, 0))
    if gw.shape[0] > 0:
        gs = gs - pad(gw[:1], (0, 0, m, 0))
    fc = rho * gc[..., None] * _cross(vc, zb[1:] - zb[:-1])
    fs = rho * gs[..., None] * _cross(vs, zb[:, 1:] - zb[:, :-1])
    return split_to_vertex(fc, 0) + split_to_vertex(fs, 1)


class LinearUVLM:
    """aero/linear/linear_uvlm.py:79-1039."""

    def __init__(self, case, reference, wake_type: LinearWakeType = "frozen", bound_upwash: bool = True,
                 wake_upwash: bool = True, unsteady_force: bool = True, *, skip_checks: bool = False,
                 skip_linearisation: bool = False):
        if wake_type not in ("frozen", "prescribed", "free"):
            raise ValueError(f"unknown wake type {wake_type!r}")
        self.prescribed_wake, self.free_wake = {"frozen": (False, False), "prescribed": (True, False),
                                                "free": (True, True)}[wake_type]
        self.unsteady_force = bool(unsteady_force)
        self.bound_upwash = bool(bound_upwash)
        self.wake_upwash = bool(wake_upwash)
        if any(dw is not None for dw in getattr(case, "delta_w", [None])):
            raise NotImplementedError("variable wake discretisation is not on the linear device path")
        if getattr(case, "grid_func", None) is not None:
            raise NotImplementedError("control-surface inputs are not on the linear device path")
        self.case = case
        self.dt = case.dt
        self._reference = reference
        i0 = 0 if reference.n_tstep == 1 else None
        if i0 is None:
            raise ValueError("the reference must be a single-step (static) AeroCase")
        self.ref = {k: [a[0] for a in getattr(reference, k)] for k in
                    ("zeta_b", "zeta_b_dot", "zeta_w", "gamma_b", "gamma_b_dot", "gamma_w", "f_steady",
                     "f_unsteady")}
        self.t_ref = float(reference.t[0])
        if not skip_checks:
            if max(float(z.abs().max()) for z in self.ref["zeta_b_dot"]) > 1e-6:
                warn("Reference bound surface velocities are non-zero. Ensure that the reference state is steady "
                     "for linearisation.")
            if max(float(g.abs().max()) for g in self.ref["gamma_b_dot"]) > 1e-6:
                warn("Reference bound circulation time derivative is non-zero. Ensure that the reference state is "
                     "steady for linearisation.")
        self.kernel = case.kernel
        # the force velocity uses the cut-off kernel (aero/linear/linear_uvlm.py:147-152)
        self.force_kernel = as_segment_kernel(biot_savart_cutoff)
        self.mirror = case.mirror
        self.n_surf = reference.n_surf
        self._layout()
        self.sys = None if skip_linearisation else self.linearise()

    @property
    def reference(self):
        return self._reference

    # ---------------------------------------------------------------- layout
    def _layout(self):
        shp = lambda k: [tuple(a.shape) for a in self.ref[k]]  # noqa: E731
        zb, zw, gb, gw = shp("zeta_b"), shp("zeta_w"), shp("gamma_b"), shp("gamma_w")
        self.inputs = [("zeta_b", zb), ("zeta_b_dot", zb)]
        if self.bound_upwash:
            self.inputs.append(("nu_b", zb))
        if self.wake_upwash:
            self.inputs.append(("nu_w", zw))
        self.states = [("gamma_b", gb), ("gamma_w", gw)]
        if self.unsteady_force:
            self.states.append(("gamma_b_nm1", gb))
        if self.prescribed_wake:
            self.states += [("zeta_w", zw), ("zeta_b", zb)]
        self.outputs = [("f_steady", zb)]
        if self.unsteady_force:
            self.outputs.append(("f_unsteady", zb))

    @staticmethod
    def _size(spec):
        return sum(int(np.prod(s)) for _, shapes in spec for s in shapes)

    @staticmethod
    def _unpack(vec, spec, lead=()):
        """vec (*lead, total) -> {name: [per-surface (*lead, *shape)]}."""
        out, off = {}, 0
        for name, shapes in spec:
            lst = []
            for s in shapes:
                k = int(np.prod(s))
                lst.append(vec[..., off:off + k].reshape(*lead, *s))
                off += k
            out[name] = lst
        return out

    @staticmethod
    def _pack(d, spec, lead=()):
        parts = []
        for name, shapes in spec:
            for a, s in zip(d[name], shapes):
                parts.append(a.reshape(*lead, int(np.prod(s))))
        return torch.cat(parts, dim=-1)

    def _ref_values(self, spec):
        ref = dict(self.ref)
        ref["gamma_b_nm1"] = self.ref["gamma_b"]
        ref["nu_b"] = [torch.zeros_like(z) for z in self.ref["zeta_b"]]
        ref["nu_w"] = [torch.zeros_like(z) for z in self.ref["zeta_w"]]
        return {name: ref[name] for name, _ in spec}

    # ------------------------------------------------------------------ step
    def step(self, u_np1: dict, x_n: dict):
        """aero/linear/linear_uvlm.py:322-430 on total values: (x_{n+1}, y_n)."""
        ref, dt, ff = self.ref, self.dt, self.case.flowfield
        t = self.t_ref
        zb_np1, zbd_np1 = u_np1["zeta_b"], u_np1["zeta_b_dot"]
        nu_b, nu_w = u_np1.get("nu_b"), u_np1.get("nu_w")
        gb_n, gw_n = x_n["gamma_b"], x_n["gamma_w"]
        gbd_n = ([(a - b) / dt for a, b in zip(gb_n, x_n["gamma_b_nm1"])] if self.unsteady_force
                 else ref["gamma_b_dot"])
        zb_n, zw_n = (x_n["zeta_b"], x_n["zeta_w"]) if self.prescribed_wake else (ref["zeta_b"], ref["zeta_w"])
        # wake propagation (aero/uvlm.py:912-941, aero/utils.py:400-450) with
        # the wake upwash input added to the new wake grid (aero/uvlm.py:943-947)
        zw_np1, gw_np1 = [], []
        for s in range(self.n_surf):
            base = zw_n[s][:-1]
            v = flow_at(ff, base, t)
            if self.free_wake and base.numel():
                v = v + AG.v_ind(base.reshape(-1, 3), list(zb_n) + list(zw_n), list(gb_n) + list(gw_n),
                                 self.kernel, self.mirror).reshape(base.shape)
            zw = torch.cat((zb_np1[s][-1:], base + dt * v), dim=0)
            if nu_w is not None:
                zw = zw + dt * nu_w[s]
            zw_np1.append(zw)
            gw_np1.append(torch.cat((gb_n[s][-1:], gw_n[s][:-1]), dim=0))
        # circulation solve (aero/uvlm.py:949-993): bound-only AIC, wake upwash
        # and the bound upwash input on the right-hand side
        cs, ncs, bs = [], [], []
        for s in range(self.n_surf):
            c, nc = colloc(zb_np1[s]), normals(zb_np1[s])
            vbc = flow_at(ff, c, t) - colloc(zbd_np1[s])
            if nu_b is not None:
                vbc = vbc + colloc(nu_b[s])
            cs.append(c.reshape(-1, 3))
            ncs.append(nc.reshape(-1, 3))
            bs.append(vbc)
        c_all, nc_all = torch.cat(cs), torch.cat(ncs)
        v_w = AG.v_ind(c_all, zw_np1, gw_np1, self.kernel, self.mirror)
        v_bc = torch.cat([b.reshape(-1, 3) for b in bs]) + v_w
        rhs = -(v_bc * nc_all).sum(-1)
        g_vec = AG.aic_solve(rhs, c_all, nc_all, list(zb_np1), None, self.kernel, self.mirror)
        gb_np1, off = [], 0
        for g in gb_n:
            k = g.numel()
            gb_np1.append(g_vec[off:off + k].reshape(g.shape))
            off += k
        # outputs at n: forces of the state at n on the grid at n+1
        # (aero/linear/linear_uvlm.py:387-420)
        rho = ff.rho
        zs, gs = list(zb_np1) + list(zw_n), list(gb_n) + list(gw_n)

        def v_out(x):
            return flow_at(ff, x, t) + AG.v_ind(x, zs, gs, self.force_kernel, self.mirror)

        f_st = [_forces(zb_np1[s], zbd_np1[s], gb_n[s], gw_n[s], rho, v_out, nu_b[s] if nu_b is not None else None)
                for s in range(self.n_surf)]
        x_np1 = {"gamma_b": gb_np1, "gamma_w": gw_np1}
        y_n = {"f_steady": f_st}
        if self.unsteady_force:
            x_np1["gamma_b_nm1"] = list(gb_n)
            y_n["f_unsteady"] = [split_to_vertex(split_to_vertex(rho * gbd_n[s][..., None] * normals(zb_np1[s]), 0), 1)
                                 for s in range(self.n_surf)]
        if self.prescribed_wake:
            x_np1["zeta_w"] = zw_np1
            x_np1["zeta_b"] = list(zb_np1)
        return x_np1, y_n

    # ------------------------------------------------------------- linearise
    def linearise(self) -> LinearSystem:
        """aero/linear/linear_uvlm.py:783-866: [A B; C D] = d(x_{n+1}, y_n) /
        d(x_n, u_{n+1}) at the reference, one reverse pass per row."""
        nx, nu = self._size(self.states), self._size(self.inputs)
        ny = self._size(self.outputs)
        x0 = self._pack(self._ref_values(self.states), self.states).detach().clone().requires_grad_(True)
        u0 = self._pack(self._ref_values(self.inputs), self.inputs).detach().clone().requires_grad_(True)
        with torch.enable_grad():
            x_np1, y_n = self.step(self._unpack(u0, self.inputs), self._unpack(x0, self.states))
            out = torch.cat((self._pack(x_np1, self.states), self._pack(y_n, self.outputs)))
        jac = torch.zeros((nx + ny, nx + nu), dtype=F64, device=x0.device)
        for i in range(nx + ny):
            if not out[i].requires_grad:
                continue
            gx, gu = torch.autograd.grad(out[i], (x0, u0), retain_graph=i + 1 < nx + ny, allow_unused=True)
            if gx is not None:
                jac[i, :nx] = gx
            if gu is not None:
                jac[i, nx:] = gu
        return LinearSystem(jac[:nx, :nx].contiguous(), jac[:nx, nx:].contiguous(), jac[nx:, :nx].contiguous(),
                            jac[nx:, nx:].contiguous(), self.dt)

    # ------------------------------------------------------------------- run
    def run(self, u: AeroInputUnflattened, x0: AeroStateUnflattened | None = None, flowfield=None):
        """aero/linear/linear_uvlm.py:868-990: march perturbation inputs u
        ((n_tstep, ...) per surface); a flow field adds its change from the
        reference time at the reference bound / wake vertices as nu_b / nu_w."""
        if self.sys is None:
            self.sys = self.linearise()
        if self.prescribed_wake and self.sys.removed_u_np1:
            warn("Wake perturbations coordinates at the trailing edge are zero when removing u_np1 from the system.")
        zb_t = [to_device(z) for z in u.zeta_b]
        n_t = zb_t[0].shape[0]
        t = self.t_ref + np.arange(n_t) * self.dt
        dev = zb_t[0].device
        ut = {"zeta_b": zb_t, "zeta_b_dot": [to_device(z) for z in u.zeta_b_dot]}
        for name, ref_key, on in (("nu_b", "zeta_b", self.bound_upwash), ("nu_w", "zeta_w", self.wake_upwash)):
            if not on:
                continue
            given = getattr(u, name)
            if given is None and flowfield is None:
                warn(f"No flowfield or {'bound' if name == 'nu_b' else 'wake'} upwash perturbations provided. "
                     f"Assuming zero {'bound' if name == 'nu_b' else 'wake'} upwash perturbations.")
            vals = [torch.zeros((n_t, *z.shape), dtype=F64, device=dev) for z in self.ref[ref_key]] \
                if given is None else [to_device(a).clone() for a in given]
            if flowfield is not None:
                for s, z in enumerate(self.ref[ref_key]):
                    u0 = flowfield.vmap_call(z, t[0])
                    for i in range(n_t):
                        vals[s][i] += flowfield.vmap_call(z, t[i]) - u0
            ut[name] = vals
        u_vec = self._pack(ut, self.inputs, (n_t,))
        x0_vec = None if x0 is None else self._pack(vars(x0), self.states)
        x_vec, y_vec = self.sys.run(u_vec, x0_vec)
        xt = self._unpack(x_vec, self.states, (n_t,))
        yt = self._unpack(y_vec, self.outputs, (n_t,))

        def total(d, spec):
            ref = self._ref_values(spec)
            return {k: [a + r[None] for a, r in zip(d[k], ref[k])] for k in d}

        ut_tot = total(ut, self.inputs)
        mk_u = lambda d: AeroInputUnflattened(d["zeta_b"], d["zeta_b_dot"], d.get("nu_b"), d.get("nu_w"))  # noqa
        mk_x = lambda d: AeroStateUnflattened(d["gamma_b"], d["gamma_w"], d.get("gamma_b_nm1"), d.get("zeta_w"),  # noqa
                                              d.get("zeta_b"))
        mk_y = lambda d: AeroOutputUnflattened(d["f_steady"], d.get("f_unsteady"))  # noqa
        return AeroLinearResult(self.reference, u, mk_x(xt), mk_y(yt), mk_u(ut_tot), mk_x(total(xt, self.states)),
                                mk_y(total(yt, self.outputs)), n_t, self.n_surf, t,
                                list(self.reference.surf_b_names), list(self.reference.surf_w_names))
