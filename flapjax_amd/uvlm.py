"""UVLM orchestration on the device (aero/uvlm.py:166-1406 of the reference).

Same class name, constructor keywords and entry points as the reference
(``set_design_variables``, ``solve_static``, ``solve_prescribed_dynamic``,
``case_solve``, ``base_solve_from_grid``).  Every step runs as a fixed
sequence of HIP kernels on torch's current stream with all state resident in
HBM; the host only enqueues work.

Per dynamic step (aero/uvlm.py:912-1103):
  1. bound grids at n and n-1 and grid velocities from the beam poses (K8);
  2. collocation points / normals / velocities into packed buffers (K5);
  3. free-wake rollup: filament table of [bound(n-1), wake(n-1)] and one fused
     pair-sum + convection launch per surface (K1 + K4), wake circulation shift;
  4. bound AIC (K2) and right-hand side with the new wake's upwash (K1);
  5. circulation solve (K3);
  6. filament-midpoint velocities and steady vertex forces (K1 + K6), strip
     alpha, circulation rate and unsteady force (K7).

Arithmetic forms (``precision``):
  * "fp64"  - the plain drop-in: FP64 everywhere.  The AIC is re-assembled
              every step as the reference does.  By default (``aic_reuse``)
              the static and the dynamic system each keep an FP64 factor and
              its explicit inverse; each step's system is solved by FP64
              iterative refinement against the freshly assembled AIC (LAPACK
              dsgesv's stop test), which gives the solution of that step's
              system to FP64 accuracy.  The stop-test outcome of every step is
              read back behind the step's remaining work; a step whose
              refinement failed is refactored and computed again.
              ``aic_reuse=False`` factors and solves every step literally;
  * "mixed" - FP32 far field / FP64 near field for the rollup and force pair
              sums; the AIC, the wake upwash in the right-hand side and the
              circulation solve stay FP64, so the solve is FP64-accurate for
              the given wake.
"""

from __future__ import annotations

from collections.abc import Sequence
from warnings import warn

import numpy as np
import torch

from . import _kernels as K
from . import _lib as L
from .aic import segment_tables
from .data_structures import FIELDS, AeroCase, GridDiscretisation, field_shapes
from .linalg import RefinedSolver
from .utils import HORSESHOE_LENGTH, InducedVelocity, as_segment_kernel, to_device

F64 = torch.float64
PRECISIONS = {"fp64": L.PREC_FP64, "mixed": L.PREC_MIXED}


class UVLM:
    """aero/uvlm.py:166-339 (constructor checks follow the reference)."""

    def __init__(self, grid_shapes, dof_mapping, variable_wake_disc: bool = False, mirror_point=None,
                 mirror_normal=None, kernel=None, grid_func=None, free_wake: bool = False,
                 gamma_dot_relaxation: float = 0.7, include_unsteady_force: bool = True, batch_size=64,
                 polars=None, polar_circulation_scale: float = 0.0, precision: str = "fp64",
                 aic_reuse: bool | None = None):
        if isinstance(dof_mapping, (np.ndarray, torch.Tensor)):
            dof_list = [dof_mapping]
        elif isinstance(dof_mapping, Sequence):
            dof_list = list(dof_mapping)
        else:
            raise TypeError("Invalid dof mapping type")
        self.n_surf = len(grid_shapes)
        gds = []
        for g in grid_shapes:
            if isinstance(g, GridDiscretisation):
                gds.append(g)
            elif isinstance(g, Sequence):
                if len(g) != 3:
                    raise ValueError("Grid shape tuple must have exactly three elements (m, n, m_star)")
                gds.append(GridDiscretisation(*[int(v) for v in g]))
            else:
                raise TypeError("Grid shape must be either a Sequence of three integers or a GridDiscretisation instance")
        self.grid_disc = tuple(gds)
        if len(dof_list) != self.n_surf:
            raise ValueError(f"Expected {self.n_surf} DOF mapping arrays, got {len(dof_list)}")
        self._dof_host = []
        for i, d in enumerate(dof_list):
            d = np.asarray(d.detach().cpu() if isinstance(d, torch.Tensor) else d)
            if not np.issubdtype(d.dtype, np.integer):
                raise TypeError("dof_mapping must be an integer array")
            if d.shape != (self.grid_disc[i].n + 1,):
                raise ValueError(f"dof_mapping of surface {i} must have shape ({self.grid_disc[i].n + 1},)")
            self._dof_host.append(d.astype(np.int32))
        if grid_func is not None and not callable(grid_func):
            raise TypeError("grid_func must be callable")
        # control surfaces (aero/uvlm.py:333-339): grid_func(zeta_b0, **angles)
        # returns the deflected local grids; None is the reference's identity
        self.grid_func = grid_func
        self.cs_ang0 = {}
        self.cs_vel0 = {}
        if polars is not None and any(p is not None for p in polars):
            raise NotImplementedError("polar corrections are not on the device path")
        if not 0.0 <= polar_circulation_scale <= 1.0:
            raise ValueError(f"polar_circulation_scale must be in [0, 1], got {polar_circulation_scale}.")
        if (mirror_point is None) != (mirror_normal is None):
            raise ValueError("Both mirror_point and mirror_normal must be provided to apply mirroring, or both must "
                             "be None to apply no mirroring.")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {tuple(PRECISIONS)}")
        self.kernel = as_segment_kernel(kernel)
        self.kernels_b = [self.kernel] * self.n_surf
        self.kernels_w = [self.kernel] * self.n_surf
        self.variable_wake_disc = variable_wake_disc
        self.free_wake = bool(free_wake)
        # one launch for the forces of step n and the rollup of step n + 1
        # (free wake, one surface, FP64); False keeps separate launches
        self.fuse_rollup = True
        self.gamma_dot_relaxation = float(gamma_dot_relaxation)
        self.include_unsteady_force = include_unsteady_force
        self.batch_size = batch_size
        self.precision = precision
        self._prec = PRECISIONS[precision]
        self.aic_reuse = True if aic_reuse is None else bool(aic_reuse)
        if mirror_point is None:
            self.mirror = None
            self.mirror_point = self.mirror_normal = None
        else:
            mp = np.asarray(mirror_point.detach().cpu() if isinstance(mirror_point, torch.Tensor) else mirror_point,
                            dtype=np.float64)
            mn = np.asarray(mirror_normal.detach().cpu() if isinstance(mirror_normal, torch.Tensor) else mirror_normal,
                            dtype=np.float64)
            mn = mn / np.linalg.norm(mn)
            self.mirror = (mp, mn)
            self.mirror_point, self.mirror_normal = mp, mn
        self.n_bound_panels = tuple(gd.m * gd.n for gd in self.grid_disc)
        self.n_wake_panels = tuple(gd.m_star * gd.n for gd in self.grid_disc)
        self.n_b = sum(self.n_bound_panels)
        self._b_off = np.concatenate(([0], np.cumsum(self.n_bound_panels))).astype(int)
        nseg = [K.n_segments(gd.m, gd.n) for gd in self.grid_disc]
        self._s_off = np.concatenate(([0], np.cumsum(nseg))).astype(int)
        self.flowfield = None
        self._dt = None
        self.zeta_b0 = None
        self.hg_ref = None
        # one kept factor per system kind: the static AIC carries the lumped
        # wake, the dynamic one does not
        self._solvers = {True: RefinedSolver(), False: RefinedSolver()} if self.aic_reuse else None
        self._refactor = False
        self._aic = None
        self._bufs = {}

    def linearise(self, reference, wake_type, bound_upwash: bool = True, wake_upwash: bool = True,
                  unsteady_force: bool = True):
        """aero/uvlm.py:356-385: the linearised model about a steady reference
        case (a static solve of this UVLM)."""
        from .linear_uvlm import LinearUVLM

        return LinearUVLM(self, reference=reference, wake_type=wake_type, bound_upwash=bound_upwash,
                          wake_upwash=wake_upwash, unsteady_force=unsteady_force)

    # ------------------------------------------------------------------ setup
    @property
    def dt(self):
        if self._dt is None:
            raise ValueError("Time step length dt has not been set.")
        return self._dt

    def set_design_variables(self, dt, flowfield, zeta_b0, hg0, delta_w=None, reference_cs_angles=None):
        """aero/uvlm.py:387-472."""
        if isinstance(zeta_b0, (np.ndarray, torch.Tensor)):
            z0 = [zeta_b0]
        elif isinstance(zeta_b0, Sequence):
            z0 = list(zeta_b0)
        else:
            raise TypeError("Invalid zeta_b0 type")
        if len(z0) != self.n_surf:
            raise ValueError(f"Expected {self.n_surf} aerodynamic grid coordinate arrays, got {len(z0)}")
        self.zeta_b0 = []
        for i, z in enumerate(z0):
            z = to_device(z)
            gd = self.grid_disc[i]
            if tuple(z.shape) != (gd.m + 1, gd.n + 1, 3):
                raise ValueError(f"zeta_b0 must have shape {(gd.m + 1, gd.n + 1, 3)}, got {tuple(z.shape)}")
            self.zeta_b0.append(z)
        if delta_w is None:
            dws = [None] * self.n_surf
        elif isinstance(delta_w, (np.ndarray, torch.Tensor)):
            dws = [delta_w]
        elif isinstance(delta_w, Sequence):
            dws = list(delta_w)
        else:
            raise TypeError("Invalid delta_w type")
        self.delta_w = []
        for i, dw in enumerate(dws):
            if dw is None:
                self.delta_w.append(None)
                continue
            dw = np.asarray(dw.detach().cpu() if isinstance(dw, torch.Tensor) else dw, dtype=np.float64)
            if dw.shape != (self.grid_disc[i].m_star,):
                raise ValueError(f"delta_w must have shape ({self.grid_disc[i].m_star},)")
            self.delta_w.append(dw)
        if reference_cs_angles is not None:
            self.cs_ang0 = {k: to_device(v).reshape(()) for k, v in reference_cs_angles.items()}
            self.cs_vel0 = {k: torch.zeros_like(v) for k, v in self.cs_ang0.items()}
        hg0 = to_device(hg0)
        if hg0.dim() != 3 or tuple(hg0.shape[1:]) != (4, 4):
            raise ValueError("hg0 must have shape (n_nodes, 4, 4)")
        self.hg_ref = hg0
        self._dof = [torch.as_tensor(d, device=hg0.device) for d in self._dof_host]
        if not isinstance(dt, (float, int, np.floating)):
            if isinstance(dt, (np.ndarray, torch.Tensor)) and np.ndim(dt) == 0:
                dt = float(dt)
            else:
                raise TypeError("dt must be either a float or an Array scalar")
        self._dt = float(dt)
        self.flowfield = flowfield
        # precomputed next-step rollups (see _dynamic_step) belong to one design
        self._design_version = getattr(self, "_design_version", 0) + 1
        self.zeta_b_ref = self.hg_to_zeta_b(hg0, self.cs_ang0)
        self.zeta_w_ref = self.initialise_wake(self.zeta_b_ref)

    def _cs_grids(self, cs_ang_n):
        """The local grids with the control surfaces deflected: zeta_b0, or
        grid_func(zeta_b0, **cs_ang_n) (aero/uvlm.py:545-547)."""
        if self.grid_func is None:
            return self.zeta_b0
        z = self.grid_func(self.zeta_b0, **(cs_ang_n or {}))
        if len(z) != self.n_surf:
            raise ValueError(f"grid_func must return {self.n_surf} grids, got {len(z)}")
        out = []
        for zi, z0 in zip(z, self.zeta_b0):
            zi = to_device(zi)
            if zi.shape != z0.shape:
                raise ValueError(f"grid_func must keep the grid shape {tuple(z0.shape)}, got {tuple(zi.shape)}")
            out.append(zi)
        return out

    def hg_to_zeta_b(self, hg_n, cs_ang_n=None, out=None):
        """aero/uvlm.py:537-560: every vertex of the (deflected) local grid
        moved by its beam node's frame.  cs_ang_n: {name: angle} (the
        reference angles set by set_design_variables when None)."""
        hg_n = to_device(hg_n)
        z0 = self._cs_grids(self.cs_ang0 if cs_ang_n is None else cs_ang_n)
        res = []
        for i in range(self.n_surf):
            z, _ = K.grid_from_hg(hg_n, None, self._dof[i], z0[i], zeta=None if out is None else out[i])
            res.append(z)
        return res

    def hg_dot_to_zeta_b_dot(self, hg_n, hg_dot_n, cs_ang_n=None, cs_vel_n=None, out=None):
        """aero/uvlm.py:562-621: hg_dot @ [zeta0_cs; 1] plus, with control
        surfaces, R times the forward derivative of the grid function along
        the control velocities (torch.func.jvp where the reference takes a
        jax.jvp)."""
        hg_dot_n = to_device(hg_dot_n)
        if cs_ang_n is None and cs_vel_n is None:
            cs_ang_n, cs_vel_n = self.cs_ang0, self.cs_vel0
        cs_ang_n = cs_ang_n or {}
        cs_vel_n = cs_vel_n or {}
        for key in set(cs_ang_n) | set(cs_vel_n):
            if key not in cs_vel_n or key not in cs_ang_n:
                raise ValueError(f"Missing pair of control angles and velocities for control surface key '{key}'")
        angs = {k: to_device(v).reshape(()) for k, v in cs_ang_n.items()}
        vels = {k: to_device(cs_vel_n[k]).reshape(()) for k in angs}
        z0 = self._cs_grids(angs)
        dz = None
        if self.grid_func is not None and angs:
            _, dz = torch.func.jvp(lambda a: self.grid_func(self.zeta_b0, **a), (angs,), (vels,))
        res = []
        for i in range(self.n_surf):
            z, _ = K.grid_from_hg(hg_dot_n, None, self._dof[i], z0[i], zeta=None if out is None else out[i])
            if dz is not None:
                # the rotation of each node frame applied to the local grid velocity
                hg_rot = to_device(hg_n).clone()
                hg_rot[:, :3, 3] = 0.0
                zr, _ = K.grid_from_hg(hg_rot, None, self._dof[i], to_device(dz[i]))
                z += zr
            res.append(z)
        return res

    def _wake_s(self, i):
        gd = self.grid_disc[i]
        dw = self.delta_w[i] if getattr(self, "delta_w", None) and self.delta_w[i] is not None else \
            np.ones(gd.m_star) * self.dt * self.flowfield.u_inf_mag
        return to_device(np.concatenate((np.zeros(1), np.cumsum(dw))))

    def initialise_wake(self, zeta_b=None, out=None):
        """aero/uvlm.py:662-689 (uniform discretisation)."""
        zeta_b = self.zeta_b_ref if zeta_b is None else zeta_b
        return [K.wake_initialise(to_device(zeta_b[i]), self._wake_s(i), self.flowfield.u_inf_dir,
                                  zeta_w=None if out is None else out[i]) for i in range(self.n_surf)]

    def _horseshoe_wake(self, zeta_b, i):
        """aero/uvlm.py:642-660."""
        gd = self.grid_disc[i]
        if gd.m_star == 0:
            warn("Horseshoe wake requested but m_star == 0, skipping.")
            return to_device(zeta_b)[-1:].contiguous()
        s = to_device(np.array([0.0, HORSESHOE_LENGTH]))
        return K.wake_initialise(to_device(zeta_b), s, self.flowfield.u_inf_dir)

    # ---------------------------------------------------------------- kernels
    def _packed(self, per_surface_shape_fn):
        return [per_surface_shape_fn(gd) for gd in self.grid_disc]

    def _stencils(self, zeta_b_n, zeta_b_dot_n, c_out=None, nc_out=None):
        """Collocation points and normals of all surfaces into packed (N_b, 3)
        buffers (views per surface), plus collocation velocities."""
        dev = zeta_b_n[0].device
        c_all = torch.empty((self.n_b, 3), dtype=F64, device=dev)
        nc_all = torch.empty((self.n_b, 3), dtype=F64, device=dev)
        cd_all = torch.empty((self.n_b, 3), dtype=F64, device=dev) if zeta_b_dot_n is not None else None
        for i, gd in enumerate(self.grid_disc):
            a, b = self._b_off[i], self._b_off[i + 1]
            zd = zeta_b_dot_n[i] if zeta_b_dot_n is not None else None
            L.check(L.lib().fj_colloc_normals(L.ptr(zeta_b_n[i]), L.ptr(zd) if zd is not None else None, gd.m, gd.n,
                                              L.ptr(c_all[a:b]), L.ptr(nc_all[a:b]),
                                              L.ptr(cd_all[a:b]) if cd_all is not None else None, L.stream()),
                    "fj_colloc_normals")
        return c_all, nc_all, cd_all

    def _assemble_aic(self, c_all, nc_all, zeta_b_n, zeta_w_n=None):
        dev = c_all.device
        if self._aic is None or self._aic.shape[0] != self.n_b:
            self._aic = torch.empty((self.n_b, self.n_b), dtype=F64, device=dev)
        A = self._aic
        kern = self.kernel
        for i in range(self.n_surf):  # target surface
            a, b = self._b_off[i], self._b_off[i + 1]
            blk = A[a:b]
            for j in range(self.n_surf):  # source surface
                K.aic_assemble(c_all[a:b], nc_all[a:b], zeta_b_n[j], kern.kind, kern.eps, kern.rcut, blk,
                               int(self._b_off[j]), zeta_w=None if zeta_w_n is None else zeta_w_n[j],
                               mirror=self.mirror)
        return A

    def _solve(self, A, rhs, static):
        if self._solvers is None:
            return K.lu_solve(K.lu_factor(A, overwrite=True), rhs)
        solver = self._solvers[bool(static)]
        if self._refactor:
            solver.factor(A)
        self._last_system = (A, rhs)
        return solver.submit(A, rhs)

    def _run_checked(self, step, static, i_ts):
        """Run ``step`` (a pure function of its inputs that ends with the
        circulation solve enqueued) and check the solve's stop test; if the
        kept factor no longer serves this step's AIC, refactor on it and run
        the step again."""
        res = step()
        if self._confirm(static):
            return res
        # the kept inverse no longer converges this step's system in the
        # allowed corrections: the geometry is changing (a deforming wing),
        # so allow more corrections from now on, which keeps one inverse
        # serving as the refinement's preconditioner instead of
        # refactorising at nearly every step (measured on the coupled gust
        # case: 254 refactorisations in 300 steps with 3 corrections, 1 with
        # 8; a rigid wing never gets here and keeps the shorter sweep)
        sv = self._solvers[bool(static)]
        sv.max_iter = max(sv.max_iter, 8)
        self._refactor = True
        try:
            res = step()
        finally:
            self._refactor = False
        if not self._confirm(static):
            A, rhs = self._last_system
            if bool(torch.isfinite(A).all()) and bool(torch.isfinite(rhs).all()):
                raise L.FlapjaxLibError("circulation solve failed its stop test after refactorisation "
                                        "(singular or ill-conditioned AIC)")
            # non-finite geometry or circulation: like the reference, the
            # NaN/inf propagates into the results instead of raising
            warn(f"Non-finite values in the circulation system at time step {i_ts}; results are not finite.")
        return res

    def _confirm(self, static) -> bool:
        """Stop-test outcome of the step's circulation solve (waits for the
        solve only, not for the rest of the step)."""
        if self._solvers is None:
            return True
        return self._solvers[bool(static)].confirm() >= 0

    def _tables(self, zetas, gammas):
        return segment_tables(zetas, gammas, [self.kernel] * len(zetas), self.mirror)

    # ----------------------------------------------------------------- solve
    def base_solve_from_grid(self, q_nm1, t_n, zeta_b_n, zeta_b_nm1, zeta_b_dot_n, static, horseshoe, out=None):
        """aero/uvlm.py:824-1117.  Returns (c, nc, gamma_b, gamma_w, gamma_b_dot,
        zeta_b, zeta_w, zeta_b_dot, f_steady, f_unsteady, alpha).  ``out``
        (optional) is a snapshot dict of preallocated views to write into."""
        if not (0.0 < self.gamma_dot_relaxation <= 1.0):
            warn("Gamma_dot relaxation factor not in (0, 1]")
        if not static and horseshoe:
            warn("Horseshoe wake not compatible with non-static solve. Overriding horseshoe to False.")
            horseshoe = False
        if not static and q_nm1 is None:
            raise ValueError("q_nm1 needs to be specified for dynamic solve")
        ff = self.flowfield
        t_n = float(t_n)
        dev = zeta_b_n[0].device
        o = out if out is not None else {}
        c_all, nc_all, cd_all = self._stencils(zeta_b_n, zeta_b_dot_n)
        c_n = [c_all[self._b_off[i]:self._b_off[i + 1]].reshape(gd.m, gd.n, 3) for i, gd in enumerate(self.grid_disc)]
        nc_n = [nc_all[self._b_off[i]:self._b_off[i + 1]].reshape(gd.m, gd.n, 3)
                for i, gd in enumerate(self.grid_disc)]

        # ---- wake
        if static:
            if horseshoe:
                zeta_w_n = [self._horseshoe_wake(zeta_b_n[i], i) for i in range(self.n_surf)]
            else:
                zeta_w_n = self.initialise_wake(zeta_b_n, out=o.get("zeta_w"))
            gamma_w_n = None
        else:
            zeta_w_n = o.get("zeta_w") or [torch.empty_like(z) for z in q_nm1["zeta_w"]]
            gamma_w_n = o.get("gamma_w") or [torch.empty_like(g) for g in q_nm1["gamma_w"]]
            tables_nm1 = None
            if self.free_wake:
                tables_nm1 = self._tables(list(zeta_b_nm1) + list(q_nm1["zeta_w"]),
                                          list(q_nm1["gamma_b"]) + list(q_nm1["gamma_w"]))
            for i, gd in enumerate(self.grid_disc):
                if self.delta_w[i] is not None:
                    # variable discretisation (aero/utils.py:431-491)
                    from .utils import FlowAt, _propagate_variable

                    vf = InducedVelocity(ff, t_n, tables_nm1, self.kernel, self._prec) if self.free_wake \
                        else FlowAt(ff, t_n)
                    zn, gn = _propagate_variable(q_nm1["gamma_b"][i], q_nm1["gamma_w"][i], zeta_b_n[i],
                                                 q_nm1["zeta_w"][i], self.delta_w[i], vf, self.dt, False)
                    zeta_w_n[i].copy_(zn)
                    gamma_w_n[i].copy_(gn)
                    continue
                K.wake_shift(q_nm1["gamma_b"][i], q_nm1["gamma_w"][i], zeta_b_n[i], gd.m_star, gamma_w_n[i],
                             zeta_w_n[i])
                if gd.m_star == 0:
                    continue
                if self.free_wake:
                    zw = q_nm1["zeta_w"][i]
                    base = zw[:-1].reshape(-1, 3)
                    dst = zeta_w_n[i][1:].reshape(-1, 3)
                    (kind, eps, rcut), seg = next(iter(tables_nm1.items()))
                    if len(tables_nm1) != 1:
                        raise NotImplementedError("one kernel kind expected")
                    K.pair_sum(seg, base, kind, eps, rcut, self._prec, L.OUT_CONVECT, out=dst, flow=ff, time=t_n,
                               dt=self.dt)
                else:
                    K.wake_convect(q_nm1["zeta_w"][i], None, ff, t_n, self.dt, zeta_w_n[i])

        # ---- system
        A = self._assemble_aic(c_all, nc_all, zeta_b_n, zeta_w_n if static else None)
        v_w = None
        if not static:
            tables_w = self._tables(zeta_w_n, gamma_w_n)
            if tables_w:
                v_w = None
                for (kind, eps, rcut), seg in tables_w.items():
                    v_w = K.pair_sum(seg, c_all, kind, eps, rcut, L.PREC_FP64,
                                     L.OUT_STORE if v_w is None else L.OUT_ADD, out=v_w)
        rhs = K.rhs(c_all, nc_all, cd_all if not static else None, v_w, ff, t_n)
        gvec = self._solve(A, rhs, static)
        gamma_b_n = []
        for i, gd in enumerate(self.grid_disc):
            g = gvec[self._b_off[i]:self._b_off[i + 1]].reshape(gd.m, gd.n)
            if out is not None:
                out["gamma_b"][i].copy_(g)
                g = out["gamma_b"][i]
            gamma_b_n.append(g)
        if static:
            gamma_w_n = []
            for i, gd in enumerate(self.grid_disc):
                rows = 1 if (horseshoe and gd.m_star != 0) else gd.m_star
                gw = K.static_wake_gamma(gamma_b_n[i], rows)
                gamma_w_n.append(gw)

        # ---- forces
        mids = torch.empty((int(self._s_off[-1]), 3), dtype=F64, device=dev)
        for i, gd in enumerate(self.grid_disc):
            L.check(L.lib().fj_filament_midpoints(L.ptr(zeta_b_n[i]), gd.m, gd.n,
                                                  L.ptr(mids[self._s_off[i]:self._s_off[i + 1]]), L.stream()),
                    "fj_filament_midpoints")
        tables_n = self._tables(list(zeta_b_n) + list(zeta_w_n), list(gamma_b_n) + list(gamma_w_n))
        v_mid = InducedVelocity(ff, t_n, tables_n, self.kernel, self._prec).induced(mids)
        f_steady, alpha = [], []
        for i, gd in enumerate(self.grid_disc):
            gw0 = gamma_w_n[i][0].contiguous() if gamma_w_n[i].shape[0] > 0 else None
            zd = zeta_b_dot_n[i] if zeta_b_dot_n is not None else None
            fs = K.steady_forces(zeta_b_n[i], zd, gamma_b_n[i], gw0, v_mid[self._s_off[i]:self._s_off[i + 1]], ff,
                                 t_n, ff.rho, out=None if out is None else out["f_steady"][i])
            f_steady.append(fs)
            alpha.append(K.strip_alpha(zeta_b_n[i], fs, ff, t_n, ff.rho,
                                       out=None if out is None else out["alpha"][i]))
        if static:
            gamma_b_dot_n, f_unsteady = None, None
        else:
            gamma_b_dot_n, f_unsteady = [], []
            for i in range(self.n_surf):
                gd_, fu = K.gamma_dot_unsteady(gamma_b_n[i], q_nm1["gamma_b"][i], q_nm1["gamma_b_dot"][i], nc_n[i],
                                               self.dt, self.gamma_dot_relaxation, ff.rho,
                                               gamma_dot=None if out is None else out["gamma_b_dot"][i],
                                               f_uns=None if out is None else out["f_unsteady"][i])
                gamma_b_dot_n.append(gd_)
                f_unsteady.append(fu)
        return (c_n, nc_n, gamma_b_n, gamma_w_n, gamma_b_dot_n, zeta_b_n, zeta_w_n, zeta_b_dot_n, f_steady,
                f_unsteady, alpha)

    # ------------------------------------------------------ fused dynamic step
    def _buf(self, name, shape, dtype=F64):
        """Work buffer kept across steps (re-allocated when the shape changes)."""
        b = self._bufs.get(name)
        if b is None or tuple(b.shape) != tuple(shape) or b.dtype != dtype:
            b = torch.empty(shape, dtype=dtype, device=self.hg_ref.device)
            self._bufs[name] = b
        return b

    def _grid_list(self, zetas, gammas):
        """(zeta, gamma, mirrored) in filament-table order: every non-empty
        source grid, followed by its mirror image when mirrored
        (aero/aic.py:276-308)."""
        out = []
        for z, g in zip(zetas, gammas):
            if z.numel() == 0 or g.numel() == 0:
                continue
            out.append((z, g, False))
            if self.mirror is not None:
                out.append((z, g, True))
        return out

    @staticmethod
    def _table_len(grids):
        return sum(K.n_segments(z.shape[0] - 1, z.shape[1] - 1) for z, _, _ in grids)

    def _pair_ws(self, name, n_tgt, n_seg, lattices=None):
        nbytes = K.pair_workspace_bytes(n_tgt, n_seg, self._prec, lattices)
        return self._buf(name, (nbytes // 8 + 1,))

    def _fast_step_ok(self, static, hg_nm1, hg_dot_n):
        return (not static and hg_nm1 is not None and hg_dot_n is not None
                and all(d is None for d in self.delta_w))

    def _dynamic_step(self, case, i_ts, hg_n, hg_nm1, hg_dot_n):
        """The dynamic branch of base_solve_from_grid (aero/uvlm.py:824-1117)
        for a uniform wake, written straight into ``case`` at ``i_ts`` with
        the fused entry points: one geometry launch per surface (grids at n
        and n-1, grid velocities, stencils, force midpoints), one filament
        table per stage (no concatenation; the wake table at n serves the
        upwash and, with the bound table after the solve, the forces), and
        one force launch per surface.  Same arithmetic and the same table
        order as the general path, so the same results."""
        ff = self.flowfield
        t_n = float(case.t[i_ts])
        dt = self.dt
        kern = self.kernel
        snap = case.snapshot(i_ts)
        prev = {k: [a[i_ts - 1] for a in getattr(case, k)] for k in ("gamma_b", "gamma_w", "gamma_b_dot", "zeta_w")}
        hg_n, hg_nm1, hg_dot_n = to_device(hg_n), to_device(hg_nm1), to_device(hg_dot_n)
        single = self.n_surf == 1
        n_b = self.n_b
        n_mid = int(self._s_off[-1])
        # Free wake, one surface: the forces' pair sum of this step and the
        # rollup of the next one sum over the same lattices with the same
        # circulations (bound and wake at n), so one launch serves both
        # (fj_pair_sum_lattices_split in FP64, fj_pair_sum_split in the
        # mixed form): targets = force midpoints, then the wake vertices,
        # whose convected positions go straight to the next step's wake rows.
        # The next step skips its rollup when it finds the mark of this step
        # and the same pose as its hg_nm1.
        fuse_ok = self.fuse_rollup and self.free_wake and single and self.grid_disc[0].m_star > 0
        fuse_next = fuse_ok and i_ts + 1 < len(case.t)
        rolled = fuse_ok and getattr(case, "_fj_rollup_mark", None) == (
            id(self), self._design_version, i_ts, hg_nm1.data_ptr())
        c_all = snap["c"][0].view(-1, 3) if single else self._buf("c_all", (n_b, 3))
        nc_all = snap["nc"][0].view(-1, 3) if single else self._buf("nc_all", (n_b, 3))
        cd_all = self._buf("cd_all", (n_b, 3))
        if fuse_next:
            gd0 = self.grid_disc[0]
            n_wv = gd0.m_star * (gd0.n + 1)
            tgt_all = self._buf("tgt_fused", (n_mid + n_wv, 3))
            mids = tgt_all[:n_mid]
        else:
            mids = self._buf("mids", (n_mid, 3))
        zb_nm1 = []
        for s, gd in enumerate(self.grid_disc):
            a, b = self._b_off[s], self._b_off[s + 1]
            sa, sb = self._s_off[s], self._s_off[s + 1]
            zb_nm1.append(self._buf(f"zb_nm1_{s}", (gd.m + 1, gd.n + 1, 3)))
            K.step_geometry(hg_n, hg_nm1, hg_dot_n, self._dof[s], self.zeta_b0[s], zeta_n=snap["zeta_b"][s],
                            zeta_nm1=zb_nm1[s], zeta_dot_n=snap["zeta_b_dot"][s], c=c_all[a:b], nc=nc_all[a:b],
                            c_dot=cd_all[a:b], mid=mids[sa:sb])
        # ---- the bound-on-bound AIC needs only this step's geometry: issued
        # first, so the device has a long kernel while the host prepares the
        # wake launches (aero/uvlm.py:944-993)
        A = self._assemble_aic(c_all, nc_all, snap["zeta_b"], None)
        # ---- wake: circulation shift, TE row, convection (aero/utils.py:400-577)
        zw_n, gw_n = snap["zeta_w"], snap["gamma_w"]
        seg_nm1 = None
        if self.free_wake and not rolled:
            grids = self._grid_list(zb_nm1 + prev["zeta_w"], prev["gamma_b"] + prev["gamma_w"])
            seg_nm1 = K.segments_from_grids(grids, kern.kind, kern.eps, self.mirror,
                                            out=self._buf("seg_nm1", (self._table_len(grids), L.SEG_STRIDE)))
        for s, gd in enumerate(self.grid_disc):
            K.wake_shift(prev["gamma_b"][s], prev["gamma_w"][s], snap["zeta_b"][s], gd.m_star, gw_n[s], zw_n[s])
            if gd.m_star == 0:
                continue
            if self.free_wake:
                if rolled:
                    continue  # rows 1.. were convected by the previous step's force launch
                base = prev["zeta_w"][s][:-1].reshape(-1, 3)
                lats = K.lattices_of(grids)
                K.pair_sum(seg_nm1, base, kern.kind, kern.eps, kern.rcut, self._prec, L.OUT_CONVECT,
                           out=zw_n[s][1:].reshape(-1, 3), flow=ff, time=t_n, dt=dt,
                           workspace=self._pair_ws(f"ws_roll_{s}", base.shape[0], seg_nm1.shape[0], lats),
                           lattices=lats)
            else:
                K.wake_convect(prev["zeta_w"][s], None, ff, t_n, dt, zw_n[s])
        # ---- system (aero/uvlm.py:944-993)
        grids_b = self._grid_list(snap["zeta_b"], snap["gamma_b"])
        grids_w = self._grid_list(zw_n, gw_n)
        s_b, s_w = self._table_len(grids_b), self._table_len(grids_w)
        seg_n = self._buf("seg_n", (s_b + s_w, L.SEG_STRIDE))
        v_w = None
        if s_w:
            K.segments_from_grids(grids_w, kern.kind, kern.eps, self.mirror, out=seg_n[s_b:])
            lats_w = K.lattices_of(grids_w)
            ws_up = K.pair_workspace_bytes(n_b, s_w, L.PREC_FP64, lats_w)
            v_w = K.pair_sum(seg_n[s_b:], c_all, kern.kind, kern.eps, kern.rcut, L.PREC_FP64, L.OUT_STORE,
                             out=self._buf("v_w", (n_b, 3)), workspace=self._buf("ws_up", (ws_up // 8 + 1,)),
                             lattices=lats_w)
        rhs = K.rhs(c_all, nc_all, cd_all, v_w, ff, t_n, out=self._buf("rhs", (n_b,)))
        gout = snap["gamma_b"][0].view(-1) if single else self._buf("gvec", (n_b,))
        if self._solvers is None:
            gout.copy_(K.lu_solve(K.lu_factor(A, overwrite=True), rhs))
        else:
            solver = self._solvers[False]
            if self._refactor:
                solver.factor(A)
            self._last_system = (A, rhs)
            solver.submit(A, rhs, out=gout)
        if not single:
            for s, gd in enumerate(self.grid_disc):
                snap["gamma_b"][s].copy_(gout[self._b_off[s]:self._b_off[s + 1]].view(gd.m, gd.n))
                snap["c"][s].copy_(c_all[self._b_off[s]:self._b_off[s + 1]].view(gd.m, gd.n, 3))
                snap["nc"][s].copy_(nc_all[self._b_off[s]:self._b_off[s + 1]].view(gd.m, gd.n, 3))
        # ---- forces (aero/uvlm.py:1018-1103)
        if s_b:
            K.segments_from_grids(grids_b, kern.kind, kern.eps, self.mirror, out=seg_n[:s_b])
        lats_n = K.lattices_of(grids_b + grids_w)
        if fuse_next:
            tgt_all[n_mid:].copy_(zw_n[0][:-1].reshape(-1, 3))
            v_mid = self._buf("v_mid", (n_mid, 3))
            t_next = float(case.t[i_ts] + self.dt)  # the next step's time, as case_solve sets it
            K.pair_sum_split(seg_n, tgt_all, n_mid, kern.kind, kern.eps, kern.rcut, v_mid, L.OUT_STORE,
                             case.zeta_w[0][i_ts + 1][1:].reshape(-1, 3), L.OUT_CONVECT, self._prec, flow=ff,
                             time=t_next, dt=dt, lattices=lats_n,
                             workspace=self._pair_ws("ws_fused", tgt_all.shape[0], s_b + s_w, lats_n))
            case._fj_rollup_mark = (id(self), self._design_version, i_ts + 1, hg_n.data_ptr())
        else:
            v_mid = K.pair_sum(seg_n if s_b + s_w else None, mids, kern.kind, kern.eps, kern.rcut, self._prec,
                               L.OUT_STORE, out=self._buf("v_mid", (n_mid, 3)),
                               workspace=self._pair_ws("ws_mid", n_mid, s_b + s_w, lats_n), lattices=lats_n)
        for s, gd in enumerate(self.grid_disc):
            sa, sb = self._s_off[s], self._s_off[s + 1]
            gw0 = gw_n[s][0] if gd.m_star > 0 else None
            K.step_forces(snap["zeta_b"][s], snap["zeta_b_dot"][s], snap["gamma_b"][s], prev["gamma_b"][s],
                          prev["gamma_b_dot"][s], gw0, v_mid[sa:sb], nc_all[self._b_off[s]:self._b_off[s + 1]], ff,
                          t_n, ff.rho, dt, self.gamma_dot_relaxation, snap["f_steady"][s], snap["gamma_b_dot"][s],
                          snap["f_unsteady"][s])
            K.strip_alpha(snap["zeta_b"][s], snap["f_steady"][s], ff, t_n, ff.rho, out=snap["alpha"][s])

    def _store(self, case, i_ts, res, static, horseshoe=False):
        names = ("c", "nc", "gamma_b", "gamma_w", "gamma_b_dot", "zeta_b", "zeta_w", "zeta_b_dot", "f_steady",
                 "f_unsteady", "alpha")
        for name, vals in zip(names, res):
            if vals is None:
                continue
            if static and name in ("gamma_b_dot", "zeta_b_dot", "f_unsteady"):
                continue
            for dst, src in zip(getattr(case, name), vals):
                if dst[i_ts].data_ptr() != src.data_ptr():
                    if name == "zeta_w" and horseshoe:
                        continue
                    dst[i_ts].copy_(src.expand_as(dst[i_ts]) if name == "gamma_w" else src)

    def case_solve(self, case, i_ts, hg_n, hg_nm1, hg_dot_n, static, horseshoe, cs_ang_n=None, cs_ang_nm1=None,
                   cs_vel_n=None):
        """aero/uvlm.py:1119-1215 (in place in ``case``).  With a grid
        function the step takes the general path (the fused one maps a single
        undeflected local grid)."""
        if not static:
            case.t[i_ts] = case.t[i_ts - 1] + self.dt if i_ts else 0.0
        if i_ts > 0 and self.grid_func is None and self._fast_step_ok(static, hg_nm1, hg_dot_n):
            self._run_checked(lambda: self._dynamic_step(case, i_ts, hg_n, hg_nm1, hg_dot_n), False, i_ts)
            return case
        q_nm1 = None
        if not static:
            q_nm1 = {k: [a[i_ts - 1] for a in getattr(case, k)] for k in ("gamma_b", "gamma_w", "gamma_b_dot", "zeta_w")}
        snap = case.snapshot(i_ts)
        cs_n = self.cs_ang0 if cs_ang_n is None else cs_ang_n
        zeta_b_n = self.hg_to_zeta_b(hg_n if hg_n is not None else self.hg_ref, cs_n, out=snap["zeta_b"])
        zeta_b_dot_n = zeta_b_nm1 = None
        if hg_dot_n is not None:
            zeta_b_dot_n = self.hg_dot_to_zeta_b_dot(hg_n, hg_dot_n, cs_n, {} if cs_vel_n is None else cs_vel_n,
                                                     out=snap["zeta_b_dot"])
            zeta_b_nm1 = self.hg_to_zeta_b(hg_nm1, self.cs_ang0 if cs_ang_nm1 is None else cs_ang_nm1)
        out = None if static else snap
        res = self._run_checked(
            lambda: self.base_solve_from_grid(q_nm1, case.t[i_ts], zeta_b_n, zeta_b_nm1, zeta_b_dot_n, static,
                                              horseshoe, out=out), static, i_ts)
        self._store(case, i_ts, res, static, horseshoe)
        if horseshoe:
            self.initialise_wake(zeta_b_n, out=snap["zeta_w"])
        return case

    def solve_static(self, hg=None, t=0.0, horseshoe=False, cs_ang=None):
        """aero/uvlm.py:1286-1333."""
        if horseshoe and any(gd.m_star == 0 for gd in self.grid_disc):
            warn("Horseshoe wake requested but m_star == 0 for at least one surface, skipping horseshoe wake.")
        case = AeroCase.zeros(self.grid_disc, 1, self.hg_ref.device)
        case.t[0] = float(t)
        # settings the AeroCase accessors use (aero/uvlm.py:1300-1326)
        case.kernels = list(self.kernels_b) + list(self.kernels_w)
        case.mirror_point, case.mirror_normal = self.mirror_point, self.mirror_normal
        case.flowfield = self.flowfield
        case.dof_mapping = list(self._dof_host)
        case.static_horseshoe = bool(horseshoe)
        case.free_wake = self.free_wake
        case.gamma_dot_relaxation = self.gamma_dot_relaxation
        case.batch_size = self.batch_size
        cs = self.cs_ang0 if cs_ang is None else cs_ang
        case.cs_ang = {k: to_device(v).reshape(1) for k, v in cs.items()}
        case.cs_vel = {k: torch.zeros_like(v) for k, v in case.cs_ang.items()}
        return self.case_solve(case, 0, hg, None, None, True, horseshoe, cs_ang_n=cs)

    def solve_prescribed_dynamic(self, init_case, hg_t, hg_dot_t, cs_ang_t=None, cs_vel_t=None, callback=None):
        """aero/uvlm.py:1335-1406: one device step per time step.
        cs_ang_t, cs_vel_t: {name: (n_tstep,)} control angle and velocity
        histories (empty when None, as in the reference)."""
        hg_t = to_device(hg_t)
        hg_dot_t = to_device(hg_dot_t)
        if hg_t.dim() != 4 or tuple(hg_t.shape[2:]) != (4, 4):
            raise ValueError("hg_t must have shape (n_tstep, n_nodes, 4, 4)")
        if hg_t.shape != hg_dot_t.shape:
            raise ValueError(f"hg_dot_n must have the same shape as hg_n, got {tuple(hg_dot_t.shape)} vs "
                             f"{tuple(hg_t.shape)}")
        n_tstep = hg_t.shape[0]
        case = init_case.to_dynamic(0, n_tstep)
        ang = {k: to_device(v) for k, v in (cs_ang_t or {}).items()}
        vel = {k: to_device(v) for k, v in (cs_vel_t or {}).items()}
        for k, v in list(ang.items()) + list(vel.items()):
            if v.shape[0] != n_tstep:
                raise ValueError(f"Inconsistent number of time steps for control surface input '{k}'")
        if ang or vel:
            case.cs_ang, case.cs_vel = ang, vel
        for i in range(1, n_tstep):
            if ang or vel or self.grid_func is not None:
                self.case_solve(case, i, hg_t[i], hg_t[i - 1], hg_dot_t[i], False, False,
                                cs_ang_n={k: v[i] for k, v in ang.items()},
                                cs_ang_nm1={k: v[i - 1] for k, v in ang.items()},
                                cs_vel_n={k: v[i] for k, v in vel.items()})
            else:
                self.case_solve(case, i, hg_t[i], hg_t[i - 1], hg_dot_t[i], False, False)
            if callback is not None:
                callback(i, case)
        return case
