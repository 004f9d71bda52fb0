"""Grid discretisation and solution container (aero/data_structures.py:29-615).

``AeroCase`` keeps the reference's field names; every per-surface history is
a device tensor with a leading time axis, so a whole run stays resident in
HBM (cfg 16x64 with 200 wake rows and 200 steps: ~70 MB).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch


@dataclass
class GridDiscretisation:
    """aero/data_structures.py:29-40: chordwise m, spanwise n, wake rows m_star."""

    m: int
    n: int
    m_star: int


FIELDS = ("zeta_b", "zeta_b_dot", "zeta_w", "gamma_b", "gamma_b_dot", "gamma_w", "f_steady", "f_unsteady",
          "alpha", "c", "nc")


def field_shapes(gd: GridDiscretisation) -> dict:
    m, n, ms = gd.m, gd.n, gd.m_star
    return {
        "zeta_b": (m + 1, n + 1, 3),
        "zeta_b_dot": (m + 1, n + 1, 3),
        "zeta_w": (ms + 1, n + 1, 3),
        "gamma_b": (m, n),
        "gamma_b_dot": (m, n),
        "gamma_w": (ms, n),
        "f_steady": (m + 1, n + 1, 3),
        "f_unsteady": (m + 1, n + 1, 3),
        "alpha": (n,),
        "c": (m, n, 3),
        "nc": (m, n, 3),
    }


@dataclass
class AeroStates:
    """aero/gradients/data_structures.py:113-191: the aerodynamic state of one
    time step that the adjoint residuals carry (per-surface lists)."""

    gamma_b: list
    gamma_w: list
    gamma_b_dot: list
    zeta_w: list


@dataclass
class AeroSurfacePlot:
    """One surface at one time step (aero/data_structures.py:617-689): the
    fields the VTK output writes."""

    zeta_b: torch.Tensor
    zeta_b_dot: torch.Tensor
    zeta_w: torch.Tensor
    gamma_b: torch.Tensor
    gamma_b_dot: torch.Tensor
    gamma_w: torch.Tensor
    f_steady: torch.Tensor
    f_unsteady: torch.Tensor
    alpha: torch.Tensor
    surf_b_name: str
    surf_w_name: str
    i_ts: int


@dataclass
class AeroCase:
    """Solution history (aero/data_structures.py:43-615): per field a list over
    surfaces of (n_tstep, ...) device tensors; ``t`` is a host array.

    A static solve returns a one-step history (the reference returns a
    snapshot without the time axis): its arrays are ``field[i_surf][0]``, and
    the accessors below take ``i_ts = None`` for it, as for a snapshot.  The
    solver settings the accessors need (kernels, mirror plane, flow field,
    dof mapping) are attached by the UVLM that made the case."""

    grid_disc: tuple
    n_tstep: int
    t: np.ndarray
    zeta_b: list = field(default_factory=list)
    zeta_b_dot: list = field(default_factory=list)
    zeta_w: list = field(default_factory=list)
    gamma_b: list = field(default_factory=list)
    gamma_b_dot: list = field(default_factory=list)
    gamma_w: list = field(default_factory=list)
    f_steady: list = field(default_factory=list)
    f_unsteady: list = field(default_factory=list)
    alpha: list = field(default_factory=list)
    c: list = field(default_factory=list)
    nc: list = field(default_factory=list)
    surf_b_names: list = field(default_factory=list)
    surf_w_names: list = field(default_factory=list)
    kernels: list = field(default_factory=list)
    mirror_point: object = None
    mirror_normal: object = None
    flowfield: object = None
    dof_mapping: list = field(default_factory=list)
    static_horseshoe: bool = False
    free_wake: bool = False
    gamma_dot_relaxation: float = 0.7
    batch_size: object = None
    # control inputs (aero/data_structures.py: cs_ang, cs_vel): {name: (n_tstep,)}
    cs_ang: dict = field(default_factory=dict)
    cs_vel: dict = field(default_factory=dict)

    def __post_init__(self):
        # aero/uvlm.py:330-331 naming
        if not self.surf_b_names:
            self.surf_b_names = [f"surf_{i}_bound" for i in range(len(self.grid_disc))]
        if not self.surf_w_names:
            self.surf_w_names = [f"surf_{i}_wake" for i in range(len(self.grid_disc))]

    @classmethod
    def zeros(cls, grid_disc, n_tstep: int, device) -> "AeroCase":
        case = cls(grid_disc=tuple(grid_disc), n_tstep=n_tstep, t=np.zeros(n_tstep))
        for gd in grid_disc:
            for name, shp in field_shapes(gd).items():
                getattr(case, name).append(torch.zeros((n_tstep, *shp), dtype=torch.float64, device=device))
        return case

    # -------------------------------------------------------------- settings
    def settings_from(self, other: "AeroCase") -> "AeroCase":
        """Copy the solver settings (not the arrays) of another case."""
        for k in ("surf_b_names", "surf_w_names", "kernels", "mirror_point", "mirror_normal", "flowfield",
                  "dof_mapping", "static_horseshoe", "free_wake", "gamma_dot_relaxation", "batch_size"):
            setattr(self, k, getattr(other, k))
        return self

    @property
    def n_surf(self) -> int:
        return len(self.grid_disc)

    @property
    def is_batched(self) -> bool:
        """More than one time step (a one-step static result acts as the
        reference's snapshot)."""
        return self.n_tstep > 1

    def _ts(self, i_ts) -> int:
        if i_ts is None:
            if self.is_batched:
                raise ValueError("i_ts must be provided for batched AeroCase")
            return 0
        i = int(i_ts)
        if not -self.n_tstep <= i < self.n_tstep:
            raise IndexError(f"time step {i} out of range for {self.n_tstep} steps")
        return i % self.n_tstep

    # ------------------------------------------------ aero/data_structures.py
    def get_states(self, i_ts=None) -> AeroStates:
        """aero/data_structures.py:240-262."""
        i = self._ts(i_ts)
        return AeroStates([g[i] for g in self.gamma_b], [g[i] for g in self.gamma_w],
                          [g[i] for g in self.gamma_b_dot], [z[i] for z in self.zeta_w])

    def gamma_full(self, i_ts=None) -> list:
        """aero/data_structures.py:264-279: bound then wake ring strengths."""
        i = self._ts(i_ts)
        return [g[i] for g in self.gamma_b] + [g[i] for g in self.gamma_w]

    def zeta_full(self, i_ts=None) -> list:
        """aero/data_structures.py:281-296: bound then wake vertex grids."""
        i = self._ts(i_ts)
        return [z[i] for z in self.zeta_b] + [z[i] for z in self.zeta_w]

    def set_arraylist_at_ts(self, attr: str, values, i_ts: int) -> None:
        """aero/data_structures.py:298-308 (in place)."""
        if not self.is_batched:
            raise TypeError("set_arraylist_at_ts only supported for batched AeroCase")
        i = self._ts(i_ts)
        for dst, v in zip(getattr(self, attr), values):
            dst[i].copy_(torch.as_tensor(v, dtype=dst.dtype, device=dst.device))

    def get_surf_snapshot(self, i_ts: int, i_surf: int) -> AeroSurfacePlot:
        """aero/data_structures.py:310-333."""
        if not self.is_batched:
            raise TypeError("get_surf_snapshot only supported for batched AeroCase")
        return self._surface(self._ts(i_ts), i_surf)

    def get_surface(self, idx: int) -> AeroSurfacePlot:
        """aero/data_structures.py:335-355."""
        if self.is_batched:
            raise TypeError("get_surface only supported for snapshot AeroCase")
        return self._surface(0, idx)

    def _surface(self, i, s) -> AeroSurfacePlot:
        return AeroSurfacePlot(self.zeta_b[s][i], self.zeta_b_dot[s][i], self.zeta_w[s][i], self.gamma_b[s][i],
                               self.gamma_b_dot[s][i], self.gamma_w[s][i], self.f_steady[s][i], self.f_unsteady[s][i],
                               self.alpha[s][i], self.surf_b_names[s], self.surf_w_names[s], i)

    def __setitem__(self, i_ts: int, snapshot: "AeroCase") -> None:
        """aero/data_structures.py:357-399: write a one-step case into step i_ts."""
        if not self.is_batched:
            raise TypeError("__setitem__ only supported for batched AeroCase")
        if snapshot.is_batched:
            raise TypeError("snapshot AeroCase must not be batched")
        i = self._ts(i_ts)
        self.t[i] = snapshot.t[0]
        for name in FIELDS:
            for dst, src in zip(getattr(self, name), getattr(snapshot, name)):
                dst[i].copy_(src[0])
        # a wake the UVLM step convected ahead into a later step is replaced
        self._fj_rollup_mark = None

    def __getitem__(self, idx: int) -> "AeroCase":
        """aero/data_structures.py:533-568: a one-step case (views) of step idx."""
        if not self.is_batched:
            raise TypeError("__getitem__ only supported for batched AeroCase")
        i = self._ts(idx)
        out = AeroCase(grid_disc=self.grid_disc, n_tstep=1, t=self.t[i:i + 1].copy())
        for name in FIELDS:
            setattr(out, name, [a[i:i + 1] for a in getattr(self, name)])
        return out.settings_from(self)

    @classmethod
    def initialise(cls, initial_snapshot: "AeroCase", n_tstep: int) -> "AeroCase":
        """aero/data_structures.py:612-615."""
        return initial_snapshot.to_dynamic(0, n_tstep)

    def project_forcing_to_beam(self, i_ts, rmat, x0_aero, include_unsteady: bool):
        """aero/data_structures.py:463-495: per beam node (global frame), the
        force and moment of the strip's vertex forces (fj_beam_project_forces)."""
        from .coupled import project_forcing_to_beam
        from .utils import to_device

        i = self._ts(i_ts)
        f = [fs[i] for fs in self.f_steady]
        extra = [fu[i] for fu in self.f_unsteady] if include_unsteady else None
        dof = [torch.as_tensor(np.asarray(d), dtype=torch.int32, device=f[0].device) for d in self.dof_mapping]
        x0 = [to_device(x) for x in x0_aero] if isinstance(x0_aero, (list, tuple)) else [to_device(x0_aero)]
        return project_forcing_to_beam(f, to_device(rmat), dof, x0, f_extra=extra)

    def _t_at(self, i_ts) -> float:
        return float(self.t[self._ts(i_ts)])

    def get_v_background(self, x_target, i_ts=None):
        """aero/data_structures.py:504-514: flow-field velocity at the points
        (a tensor (..., 3) or a list of them) at the step's time."""
        t = self._t_at(i_ts)
        if isinstance(x_target, (list, tuple)):
            return [self.flowfield.vmap_call(x, t) for x in x_target]
        return self.flowfield.vmap_call(x_target, t)

    def get_v_ind(self, x_target, i_ts=None):
        """aero/data_structures.py:516-528: induced velocity of all bound and
        wake surfaces (compute_v_ind, with the case's kernels and mirror)."""
        from . import aic

        kernels = list(self.kernels) if self.kernels else None
        if not kernels:
            raise ValueError("this AeroCase carries no kernels (made outside a UVLM solve)")
        return aic.compute_v_ind(x_target, self.zeta_full(i_ts), self.gamma_full(i_ts), kernels,
                                 self.mirror_point, self.mirror_normal, self.batch_size)

    def get_v_tot(self, x_target, i_ts=None):
        """aero/data_structures.py:530-531."""
        vi = self.get_v_ind(x_target, i_ts)
        vb = self.get_v_background(x_target, i_ts)
        if isinstance(vi, list):
            return [a + b for a, b in zip(vi, vb)]
        return vi + vb

    def snapshot(self, i_ts: int) -> dict:
        return {name: [a[i_ts] for a in getattr(self, name)] for name in FIELDS}

    def to_dynamic(self, i_ts: int, n_tstep: int) -> "AeroCase":
        """aero/data_structures.py:570-610: expand snapshot i_ts into a history."""
        dev = self.gamma_b[0].device
        out = AeroCase.zeros(self.grid_disc, n_tstep, dev)
        for name in FIELDS:
            for dst, src in zip(getattr(out, name), getattr(self, name)):
                dst[i_ts].copy_(src[i_ts if self.n_tstep > 1 else 0])
        out.t[i_ts] = self.t[i_ts if self.n_tstep > 1 else 0]
        out = out.settings_from(self)
        # the control inputs held over the run (aero/data_structures.py:570-610)
        out.cs_ang = {k: v[i_ts if v.shape[0] > 1 else 0].expand(n_tstep).clone() for k, v in self.cs_ang.items()}
        out.cs_vel = {k: v[i_ts if v.shape[0] > 1 else 0].expand(n_tstep).clone() for k, v in self.cs_vel.items()}
        return out

    def plot(self, directory, plot_bound: bool = True, plot_wake: bool = True, index=None):
        """VTK output (aero/data_structures.py:401-461): per-surface ``.vts``
        files for a single-step case, per-step files and one ``.pvd`` per
        surface and kind for a history."""
        from .vtk_output import plot_case

        return plot_case(self, directory, plot_bound, plot_wake, index)

    def to_numpy(self) -> dict:
        res = {"t": self.t.copy()}
        for name in FIELDS:
            res[name] = [a.detach().cpu().numpy() for a in getattr(self, name)]
        return res
