"""Aeroelastic coupling of the device UVLM and the device beam
(coupled/coupled.py, structure/beam.py:2319-2725, aero/utils.py:666-697).

``CoupledAeroelastic`` keeps the reference's ``set_design_variables``,
``reference_configuration``, ``static_solve`` (Aitken-relaxed fixed point)
and ``dynamic_solve`` (fluid-structure iterations inside each generalised-
alpha step, the aero step called through ``UVLM.case_solve``, the solver
protocol of aero/utils.py:736-758).  Forces cross from the aero grid to the
beam nodes with ``fj_beam_project_forces``; both solvers keep their state in
HBM, so a coupled step moves no arrays through the host (the host reads the
convergence norms only).
"""

from __future__ import annotations

from dataclasses import dataclass
from warnings import warn

import numpy as np
import torch

from . import _lib as L
from .beam import BeamStructure, ConvergenceSettings, ConvergenceStatus, StructureCase
from .data_structures import AeroCase
from .utils import cs_ang_to_cs_vel, to_device

F64 = torch.float64

# coupled/coupled.py:38-44
DEFAULT_FSI_CONVERGENCE_SETTINGS = ConvergenceSettings(25, 1e-3, 1e-5, 1e-3, 1e-5)


def project_forcing_to_beam(f_total, rmat, dof_mapping, x0_aero, f_extra=None, out=None):
    """aero/utils.py:666-697: per beam node, the sum over its chordwise strip
    of the vertex forces and of (R x0) x f (global frame); later surfaces
    overwrite shared nodes.  ``f_extra`` (optional, per surface) is added to
    ``f_total`` (steady + unsteady forcing)."""
    rmat = to_device(rmat).contiguous()
    n_nodes = rmat.shape[0]
    if out is None:
        out = torch.empty((n_nodes, 6), dtype=F64, device=rmat.device)
    lib = L.lib()
    for s, (f, dm, x0) in enumerate(zip(f_total, dof_mapping, x0_aero)):
        fb = None if f_extra is None else f_extra[s].contiguous()
        f = f.contiguous()
        dm = torch.as_tensor(dm, dtype=torch.int32, device=rmat.device)
        m1, n1 = f.shape[0], f.shape[1]
        L.check(lib.fj_beam_project_forces(n_nodes, L.ptr(rmat), m1, n1, L.ptr(f), L.ptr(fb), L.ptr(x0.contiguous()),
                                           L.ptr(dm), int(s == 0), L.ptr(out), L.stream()), "fj_beam_project_forces")
    return out


@dataclass
class AeroelasticCase:
    """coupled/data_structures.py: the structure and aero results of a solve."""

    structure: StructureCase
    aero: AeroCase


class CoupledAeroelastic:
    """coupled/coupled.py:46-486 (BaseCoupledAeroelastic + CoupledAeroelastic)."""

    def __init__(self, structure: BeamStructure, aero, fsi_convergence_settings: ConvergenceSettings =
                 DEFAULT_FSI_CONVERGENCE_SETTINGS):
        self.structure = structure
        self.aero = aero
        self.fsi_convergence_settings = fsi_convergence_settings

    # coupled/coupled.py:58-92
    def set_design_variables(self, coords, k_cs, m_cs, m_lumped, dt, flowfield, x0_aero, delta_w=None,
                             thrust_reference=None, orientation_euler=None, cs_angles_reference=None, *,
                             remove_checks: bool = False):
        self.structure.set_design_variables(coords=coords, k_cs=k_cs, m_cs=m_cs, m_lumped=m_lumped,
                                            thrust_reference=thrust_reference, orientation_euler=orientation_euler,
                                            remove_checks=remove_checks)
        self.aero.set_design_variables(dt=dt, flowfield=flowfield, delta_w=delta_w, zeta_b0=x0_aero,
                                       hg0=self.structure.hg0, reference_cs_angles=cs_angles_reference)

    def _dof_i32(self):
        # the node mappings as the projection kernel reads them (int32 on the
        # device), converted once per mapping instead of once per call
        dof = self.aero._dof
        if getattr(self, "_dof_src", None) is not dof:
            self._dof_src = dof
            self._dof32 = [d.to(torch.int32) for d in dof]
        return self._dof32

    def _project(self, case: AeroCase, i_ts: int, hg, include_unsteady: bool, out=None):
        f_un = [f[i_ts] for f in case.f_unsteady] if include_unsteady else None
        return project_forcing_to_beam([f[i_ts] for f in case.f_steady], hg[:, :3, :3], self._dof_i32(),
                                       self.aero.zeta_b0, f_extra=f_un, out=out)

    # coupled/coupled.py:122-152
    def reference_configuration(self, prescribed_dofs, horseshoe: bool = False, use_f_ext_follower: bool = False,
                                use_f_ext_dead: bool = False, t_init: float = 0.0) -> AeroelasticCase:
        return AeroelasticCase(
            structure=self.structure.reference_configuration(use_f_ext_follower=use_f_ext_follower,
                                                             use_f_ext_dead=use_f_ext_dead, use_f_aero=True,
                                                             prescribed_dofs=tuple(prescribed_dofs)),
            aero=self.aero.solve_static(t=t_init, hg=self.structure.hg0, horseshoe=horseshoe))

    # coupled/coupled.py:154-305
    def static_solve(self, prescribed_dofs, f_ext_follower=None, f_ext_dead=None, t: float = 0.0,
                     load_steps: int = 1, horseshoe: bool = False, fsi_relaxation: float = 0.6,
                     fsi_omega_min: float = 1e-2) -> AeroelasticCase:
        st_ = self.structure
        pre = st_.make_prescribed_dofs_tuple(prescribed_dofs)
        if not self.aero.flowfield.relative_motion:
            warn("Flow field relative motion is disabled. Static solve may not be accurate.")
        if not 0.0 < fsi_relaxation <= 1.0:
            raise ValueError(f"fsi_relaxation must be in (0, 1], got {fsi_relaxation}.")
        if not 0.0 < fsi_omega_min <= fsi_relaxation:
            raise ValueError(f"fsi_omega_min must be in (0, fsi_relaxation], got {fsi_omega_min}.")
        dev = st_.hg0.device
        free = torch.ones(st_.n_dof, dtype=torch.bool, device=dev)
        if pre:
            free[torch.as_tensor(pre, device=dev)] = False
        free = free.view(-1, 6)
        hg = st_.hg0.clone()
        f_app_local = torch.zeros((st_.n_nodes, 6), dtype=F64, device=dev)
        aero_case = self.aero.solve_static(t=t, hg=hg, horseshoe=horseshoe)
        struct_case = st_.reference_configuration(use_f_aero=True, prescribed_dofs=pre)
        status = ConvergenceStatus(self.fsi_convergence_settings)
        omega_prev = float(fsi_relaxation)
        r_prev = torch.zeros((st_.n_nodes, 6), dtype=F64, device=dev)
        rel = torch.empty((st_.n_nodes, 6), dtype=F64, device=dev)
        tot = torch.empty_like(rel)
        while not status.get_status():
            f_new = self._project(aero_case, 0, hg, False)
            f_app = st_._to_global(f_app_local, hg)
            r_curr = torch.where(free, f_new - f_app, torch.zeros_like(f_new))
            dr = r_curr - r_prev
            den, num = (torch.stack((torch.vdot(dr.reshape(-1), dr.reshape(-1)),
                                     torch.vdot(r_prev.reshape(-1), dr.reshape(-1)))).cpu().numpy())
            om_a = -omega_prev * num / (den if den > 0.0 else 1.0)
            om_a = min(max(om_a, fsi_omega_min), 1.0)
            omega = om_a if status.i_iter >= 2 else float(fsi_relaxation)
            f_aero = f_new if status.i_iter == 0 else (f_app + omega * r_curr).contiguous()
            s_np1 = st_.static_solve(pre, f_ext_follower=f_ext_follower, f_ext_dead=f_ext_dead, f_ext_aero=f_aero,
                                     load_steps=load_steps)
            st_._nodes(L.BEAM_NODE_REL, hg, hg_ref=s_np1.hg, h_out=rel)
            st_._nodes(L.BEAM_NODE_REL, s_np1.hg, hg_ref=st_.hg0, h_out=tot)
            delta = rel if status.i_iter == 0 else rel / omega
            nv = torch.stack((delta.norm(), tot.norm(), r_curr.norm(),
                              torch.where(free, f_new, torch.zeros_like(f_new)).abs().max(),
                              torch.isnan(delta).any().to(F64))).cpu().numpy()
            status.update_norms(*nv[:4], nv[4] != 0.0)
            hg = s_np1.hg
            struct_case = s_np1
            f_app_local = s_np1.f_ext_aero
            aero_case = self.aero.solve_static(t=t, hg=hg, horseshoe=horseshoe)
            omega_prev, r_prev = omega, r_curr
        struct_case.iters = dict(fsi=status.i_iter)
        return AeroelasticCase(structure=struct_case, aero=aero_case)

    # coupled/coupled.py:307-433
    def dynamic_solve(self, init_case: AeroelasticCase | None, prescribed_dofs, n_tstep: int, f_ext_follower=None,
                      f_ext_dead=None, t_init: float = 0.0, load_steps: int = 1, thrust_t=None, cs_ang_t=None,
                      cs_vel_t=None) -> AeroelasticCase:
        if thrust_t:
            raise NotImplementedError("thrust forces are not on the device path")
        st_ = self.structure
        pre = st_.make_prescribed_dofs_tuple(prescribed_dofs)
        dt = self.aero.dt
        t = np.arange(n_tstep) * dt + t_init
        # control inputs (coupled/coupled.py:322-342): checked, else the
        # initial case's angles held over the run; velocities by backward
        # differences when not given
        if cs_ang_t is not None:
            for key in cs_ang_t:
                if np.shape(cs_ang_t[key])[0] != n_tstep:
                    raise ValueError(f"Inconsistent number of time steps for control surface input cs_ang_t['{key}']")
                if cs_vel_t is not None and key not in cs_vel_t:
                    raise ValueError(f"Missing control velocity for control surface input cs_vel_t['{key}']")
            cs_ang_t = {k: to_device(v) for k, v in cs_ang_t.items()}
        elif init_case is not None:
            cs_ang_t = {k: to_device(v).reshape(-1)[0].expand(n_tstep).clone() for k, v in init_case.aero.cs_ang.items()}
        if cs_vel_t is None and cs_ang_t is not None:
            cs_vel_t = cs_ang_to_cs_vel(cs_ang_t, dt)
        elif cs_vel_t is not None:
            cs_vel_t = {k: to_device(v) for k, v in cs_vel_t.items()}
        cs_ang_t = cs_ang_t or {}
        cs_vel_t = cs_vel_t or {}
        controls = bool(cs_ang_t or cs_vel_t) or self.aero.grid_func is not None
        if init_case is None:
            init_case = self.reference_configuration(pre, use_f_ext_follower=f_ext_follower is not None,
                                                     use_f_ext_dead=f_ext_dead is not None, t_init=t_init)
        struct0 = init_case.structure.to_dynamic()
        aero_case = init_case.aero.to_dynamic(0, n_tstep)
        if cs_ang_t or cs_vel_t:
            aero_case.cs_ang, aero_case.cs_vel = dict(cs_ang_t), dict(cs_vel_t)
        shp = (n_tstep, st_.n_nodes, 6)

        def series(f, name):
            if f is None:
                return None
            f = to_device(f)
            if f.dim() == 2:
                f = f.expand(shp)
            if tuple(f.shape) != shp:
                raise ValueError(f"{name} must have shape {shp}")
            return f.contiguous()

        outer = self
        unsteady = bool(self.aero.include_unsteady_force)

        class _Hooks:
            @staticmethod
            def initial_force(hg0):
                return outer._project(aero_case, 0, hg0, False)

            @staticmethod
            def solve(i, hg_n, hg_nm1, hg_dot):
                if controls:
                    outer.aero.case_solve(aero_case, i, hg_n, hg_nm1, hg_dot, False, False,
                                          cs_ang_n={k: v[i] for k, v in cs_ang_t.items()},
                                          cs_ang_nm1={k: v[i - 1] for k, v in cs_ang_t.items()},
                                          cs_vel_n={k: v[i] for k, v in cs_vel_t.items()})
                else:
                    outer.aero.case_solve(aero_case, i, hg_n, hg_nm1, hg_dot, False, False)
                return outer._project(aero_case, i, hg_n, unsteady)

            @staticmethod
            def project(i, hg_n, out=None):
                return outer._project(aero_case, i, hg_n, unsteady, out=out)

            @staticmethod
            def step_end_inputs(i):
                # one surface: what fj_beam_step_end projects (vertex forces
                # of step i, the reference grid, the station -> node map)
                if outer.aero.n_surf != 1:
                    return None
                fa = aero_case.f_steady[0][i]
                fb = aero_case.f_unsteady[0][i] if unsteady else None
                return fa, fb, outer.aero.zeta_b0[0], outer._dof_i32()[0]

        struct = st_.base_dynamic_solve(struct0, t, dt, pre, load_steps, series(f_ext_follower, "f_ext_follower"),
                                        series(f_ext_dead, "f_ext_dead"), aero=_Hooks,
                                        fsi_settings=self.fsi_convergence_settings)
        return AeroelasticCase(structure=struct, aero=aero_case)
