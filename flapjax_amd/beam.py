"""Geometrically exact SE(3) beam on the device (structure/beam.py).

``BeamStructure`` keeps the reference's constructor, ``set_design_variables``,
``static_solve`` and ``dynamic_solve`` signatures.  Its Newton iterations run
on the GPU through the C-ABI (``fj_beam_system`` assembles the reduced
system from per-element kernels, ``fj_band_solve`` (banded LU, or the dense
``fj_lu_factor``/``fj_lu_solve`` when the band does not fit) solves it,
``fj_beam_update`` applies the increment and returns the convergence norms);
the host loop reads five norms per iteration and applies the reference's
stop test (``ConvergenceStatus``).  State (frames, velocities, accelerations)
stays in HBM for the whole run.

Scope: the reference's default ``OptionalJacobians`` (no geometric, dead-load
or gravity stiffness, no q_dot term in C_T), no gravity, no lumped masses, no
thrust; those options raise ``NotImplementedError`` instead of running on the
CPU.  Outputs carry the states (hg, varphi, v, v_dot, a, loads); the
post-processed force split of ``resolve_forces`` is not part of this path.
"""

from __future__ import annotations

import ctypes
from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np
import torch

from . import _kernels as K
from . import _lib as L
from .utils import to_device

F64 = torch.float64


@dataclass(frozen=True)
class ConvergenceSettings:
    """utils/data_structures.py:23-28."""

    max_n_iter: int | None
    rel_disp_tol: float | None
    abs_disp_tol: float | None
    rel_force_tol: float | None
    abs_force_tol: float | None


# structure/beam.py:71-77
DEFAULT_STRUCT_CONVERGENCE_SETTINGS = ConvergenceSettings(25, 1e-5, 1e-7, 1e-6, 1e-8)


@dataclass(frozen=True)
class OptionalJacobians:
    """structure/data_structures.py:26-32 (all off by default)."""

    d_f_ext_dead_d_n: bool = False
    d_f_grav_d_n: bool = False
    d_f_gyr_d_q_dot: bool = False
    d_f_int_d_p_d: bool = False


def _ratio(num, den):
    if den == 0.0:
        return np.inf if num != 0.0 else np.nan
    return num / den


def _finite_or_inf(x):
    """jnp.nan_to_num(x, True, inf): NaN -> inf, +inf -> the largest double."""
    if np.isnan(x):
        return np.inf
    if np.isposinf(x):
        return np.finfo(np.float64).max
    return x


class ConvergenceStatus:
    """utils/data_structures.py:32-170, fed with the norms the device
    computes: |delta disp|, |total disp|, |delta force|, max |total force|."""

    def __init__(self, settings: ConvergenceSettings):
        s = settings
        if s.rel_disp_tol is None and s.abs_disp_tol is None and s.rel_force_tol is None and s.abs_force_tol is None:
            if s.max_n_iter is None:
                raise ValueError("No convergence criteria provided, at least one tolerance or maximum iteration count "
                                 "must be specified.")
        self.convergence_settings = settings
        self.reset_status()

    def reset_status(self):
        self.i_iter = 0
        self.abs_disp_val = self.rel_disp_val = self.abs_force_val = self.rel_force_val = 0.0
        self.converged = self.has_nan = self.final_iter = False

    def update_norms(self, abs_disp, total_disp, abs_force, total_force_max, has_nan):
        s = self.convergence_settings
        self.i_iter += 1
        self.abs_disp_val = float(abs_disp)
        self.has_nan = bool(has_nan)
        c_ad = s.abs_disp_tol is not None and self.abs_disp_val < s.abs_disp_tol
        c_rd = False
        if s.rel_disp_tol is not None:
            self.rel_disp_val = _ratio(self.abs_disp_val, float(total_disp))
            c_rd = _finite_or_inf(self.rel_disp_val) < s.rel_disp_tol
        c_af = False
        if s.abs_force_tol is not None:
            self.abs_force_val = float(abs_force)
            c_af = self.abs_force_val < s.abs_force_tol
        c_rf = False
        if s.rel_force_tol is not None:
            self.rel_force_val = _ratio(float(abs_force), float(total_force_max))
            c_rf = _finite_or_inf(self.rel_force_val) < s.rel_force_tol
        self.converged = (c_rd or c_ad or c_rf or c_af) and self.i_iter > 0
        if s.max_n_iter is not None:
            self.final_iter = self.i_iter >= s.max_n_iter

    def get_status(self) -> bool:
        return self.converged or self.has_nan or self.final_iter


class TimeIntegrator:
    """Generalised-alpha parameters (structure/time_integration.py:10-51)."""

    def __init__(self, spectral_radius: float, dt: float):
        r = float(spectral_radius)
        self.spectral_radius = r
        self.dt = float(dt)
        self.alpha_m = (2.0 * r - 1.0) / (r + 1.0)
        self.alpha_f = r / (r + 1.0)
        self.gamma = (3.0 - r) / (2.0 + 2.0 * r)
        self.beta = 1.0 / ((r + 1.0) ** 2)
        self.gamma_prime = self.gamma / (self.beta * self.dt)
        self.beta_prime = (1.0 - self.alpha_m) / (self.beta * self.dt * self.dt * (1.0 - self.alpha_f))

    def predict_alpha(self, v, v_dot, a):
        """predict_q + calculate_q_alpha (structure/time_integration.py:86-150):
        the increment, velocity and acceleration at the alpha point."""
        am, af, g, b, dt = self.alpha_m, self.alpha_f, self.gamma, self.beta, self.dt
        a_n = (af * v_dot - am * a) / (1.0 - am)
        phi_n = dt * v + dt * dt * ((0.5 - b) * a + b * a_n)
        v_n = v + (1.0 - g) * dt * a + g * dt * a_n
        v_dot_n = ((1.0 - am) * a_n + am * a - af * v_dot) / (1.0 - af)
        return ((1.0 - af) * phi_n).contiguous(), ((1.0 - af) * v_n + af * v).contiguous(), \
            ((1.0 - af) * v_dot_n + af * v_dot).contiguous()

    def f_alpha(self, f_nm1, f_n):
        return (1.0 - self.alpha_f) * f_n + self.alpha_f * f_nm1


@dataclass
class StructureCase:
    """Beam states (structure/data_structures.py:36-...): a snapshot holds
    (n_nodes, ...) tensors, a trajectory (n_tstep, n_nodes, ...)."""

    hg: torch.Tensor
    varphi: torch.Tensor
    t: np.ndarray
    v: torch.Tensor | None = None
    v_dot: torch.Tensor | None = None
    a: torch.Tensor | None = None
    f_ext_follower: torch.Tensor | None = None
    f_ext_dead: torch.Tensor | None = None
    f_ext_aero: torch.Tensor | None = None
    prescribed_dofs: tuple = ()
    iters: list | None = None

    @property
    def x(self):
        return self.hg[..., :3, 3]

    @property
    def rmat(self):
        return self.hg[..., :3, :3]

    @property
    def is_dynamic(self):
        return self.v is not None

    def to_dynamic(self):
        """Snapshot with zero velocity and acceleration (structure/data_structures.py:189-230)."""
        if self.is_dynamic:
            return self
        z = torch.zeros(self.hg.shape[:-2] + (6,), dtype=F64, device=self.hg.device)
        return StructureCase(hg=self.hg, varphi=self.varphi, t=np.zeros(1), v=z, v_dot=z.clone(), a=z.clone(),
                             f_ext_follower=self.f_ext_follower, f_ext_dead=self.f_ext_dead,
                             f_ext_aero=self.f_ext_aero, prescribed_dofs=self.prescribed_dofs)


def _rot_euler_zyx(angles):
    """Rotation.from_euler('zyx', angles).as_matrix(): extrinsic rotations
    about z, then the fixed y, then the fixed x axis (R = Rx Ry Rz)."""
    z, y, x = (float(a) for a in angles)
    cz, sz, cy, sy, cx, sx = np.cos(z), np.sin(z), np.cos(y), np.sin(y), np.cos(x), np.sin(x)
    rz = np.array(((cz, -sz, 0.0), (sz, cz, 0.0), (0.0, 0.0, 1.0)))
    ry = np.array(((cy, 0.0, sy), (0.0, 1.0, 0.0), (-sy, 0.0, cy)))
    rx = np.array(((1.0, 0.0, 0.0), (0.0, cx, -sx), (0.0, sx, cx)))
    return rx @ ry @ rz


def _host(a, dtype=np.float64):
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=dtype)


class BeamStructure:
    """structure/beam.py:79-2990 (BaseBeamStructure) on the device."""

    def __init__(self, num_nodes: int, connectivity, y_vector, k_cs_index=None, m_cs_index=None,
                 m_lumped_index=None, gravity=None, thrust_nodes=None, thrust_direction=None,
                 optional_jacobians: OptionalJacobians | None = None, relaxation_factor: float = 1.0,
                 spectral_radius: float = 0.9,
                 struct_convergence_settings: ConvergenceSettings = DEFAULT_STRUCT_CONVERGENCE_SETTINGS):
        if not isinstance(num_nodes, (int, np.integer)):
            raise TypeError("num_nodes must be an int")
        conn = np.asarray(_host(connectivity, np.float64))
        if conn.ndim != 2 or (conn.size and conn.shape[1] != 2):
            raise ValueError(f"connectivity must have shape (n_elem, 2), got {conn.shape}")
        craw = connectivity.detach().cpu().numpy() if isinstance(connectivity, torch.Tensor) else np.asarray(connectivity)
        if craw.size and not np.issubdtype(craw.dtype, np.integer):
            raise TypeError("connectivity must be an integer array")
        conn = craw.reshape(-1, 2).astype(np.int64)
        # structure/utils.py:_check_connectivity
        if num_nodes != 1:
            if set(conn.ravel().tolist()) != set(range(num_nodes)):
                raise ValueError("Connectivity array either contains invalid node indices, or is missing nodes.")
        elif conn.size:
            raise ValueError("Connectivity array must be empty for a structure with a single node.")
        if m_lumped_index is not None:
            raise NotImplementedError("lumped masses are not on the device path")
        if gravity is not None and np.any(_host(gravity)):
            raise NotImplementedError("gravity loads are not on the device path")
        if thrust_nodes or thrust_direction:
            raise NotImplementedError("thrust forces are not on the device path")
        oj = optional_jacobians or OptionalJacobians()
        if oj.d_f_ext_dead_d_n or oj.d_f_grav_d_n or oj.d_f_gyr_d_q_dot or oj.d_f_int_d_p_d:
            raise NotImplementedError("optional tangent Jacobians are not on the device path")
        self.optional_jacobians = oj
        self.n_nodes = int(num_nodes)
        self.n_dof = 6 * self.n_nodes
        self.connectivity = tuple(tuple(int(v) for v in c) for c in conn)
        self.n_elem = conn.shape[0]
        self._conn = conn
        y = _host(y_vector)
        if y.shape == (3,):
            y = y[None]
        if y.shape == (1, 3):
            y = np.broadcast_to(y, (self.n_elem, 3))
        if y.shape != (self.n_elem, 3):
            raise ValueError(f"y_vector must have shape ({self.n_elem}, 3), got {y.shape}")
        self.y_vector_reference = np.array(y)
        self.k_cs_index = np.zeros(self.n_elem, int) if k_cs_index is None else _host(k_cs_index, np.int64)
        self.m_cs_index = np.zeros(self.n_elem, int) if m_cs_index is None else _host(m_cs_index, np.int64)
        for name, idx in (("k_cs_index", self.k_cs_index), ("m_cs_index", self.m_cs_index)):
            if idx.shape != (self.n_elem,):
                raise ValueError(f"{name} must have shape ({self.n_elem},)")
        self.relaxation_factor = float(relaxation_factor)
        self.spectral_radius = float(spectral_radius)
        self.struct_convergence_settings = struct_convergence_settings
        # the dynamic Newton loop's stop test on the device (one status read
        # per batch of iterations); False: the host reads the norms after
        # every iteration
        self.device_newton = True
        self.orientation = np.eye(3)
        self.time_integrator = None
        self._model = None

    # structure/beam.py:322-530
    def set_design_variables(self, coords, k_cs, m_cs, m_lumped=None, orientation_euler=None,
                             thrust_reference=None, *, remove_checks: bool = False):
        if m_lumped is not None:
            raise NotImplementedError("lumped masses are not on the device path")
        if thrust_reference:
            raise NotImplementedError("thrust forces are not on the device path")
        if orientation_euler is not None:
            oe = _host(orientation_euler)
            if oe.shape != (3,):
                raise ValueError("orientation_euler must have shape (3,)")
            self.orientation = _rot_euler_zyx(oe)
        x0r = _host(coords)
        if x0r.shape != (self.n_nodes, 3):
            raise ValueError(f"coords must have shape ({self.n_nodes}, 3), got {x0r.shape}")
        self.x0_reference = x0r
        self.x0 = x0r @ self.orientation.T
        k = _host(k_cs)
        k = k[None] if k.ndim == 2 else k
        if k.ndim != 3 or k.shape[1:] != (6, 6):
            raise ValueError("k_cs must have shape (n, 6, 6)")
        m = np.zeros((1, 6, 6)) if m_cs is None else _host(m_cs)
        m = m[None] if m.ndim == 2 else m
        if m.ndim != 3 or m.shape[1:] != (6, 6):
            raise ValueError("m_cs must have shape (n, 6, 6)")
        self.k_cs, self.m_cs = k, m
        c = self._conn
        dx = x0r[c[:, 1]] - x0r[c[:, 0]]
        if not remove_checks and self.n_elem and np.any(
                np.linalg.norm(np.cross(dx, self.y_vector_reference), axis=-1) < 1e-6):
            raise ValueError("y_vector is collinear with beam element direction for at least one element. "
                             "Please provide a different y_vector.")
        l0 = np.linalg.norm(dx, axis=-1)
        ex = dx / l0[:, None]
        ez = np.cross(ex, self.y_vector_reference)
        ez = ez / np.linalg.norm(ez, axis=-1)[:, None]
        ey = np.cross(ez, ex)
        self.l0 = l0
        self.o0 = np.stack((ex, ey, ez), axis=-1)
        self.d0 = np.zeros((self.n_elem, 6))
        self.d0[:, 0] = l0
        hg0 = np.broadcast_to(np.eye(4), (self.n_nodes, 4, 4)).copy()
        hg0[:, :3, :3] = self.orientation
        hg0[:, :3, 3] = self.x0
        self.hg0 = to_device(hg0)
        # device element tables and the node -> element CSR
        dev = self.hg0.device
        off = np.zeros(self.n_nodes + 1, np.int32)
        lists = [[] for _ in range(self.n_nodes)]
        for e, (a, b) in enumerate(c):
            lists[a].append(2 * e)
            lists[b].append(2 * e + 1)
        off[1:] = np.cumsum([len(x) for x in lists])
        flat = np.array([v for x in lists for v in x], np.int32) if self.n_elem else np.zeros(1, np.int32)
        self._tables = dict(
            conn=torch.as_tensor(c.astype(np.int32).reshape(-1) if self.n_elem else np.zeros(2, np.int32), device=dev),
            o0=to_device(self.o0.reshape(-1) if self.n_elem else np.zeros(9)),
            l0=to_device(l0 if self.n_elem else np.zeros(1)),
            d0=to_device(self.d0.reshape(-1) if self.n_elem else np.zeros(6)),
            k_cs=to_device(k[self.k_cs_index].reshape(-1) if self.n_elem else np.zeros(36)),
            m_cs=to_device(m[self.m_cs_index].reshape(-1) if self.n_elem else np.zeros(36)),
            node_off=torch.as_tensor(off, device=dev),
            node_elem=torch.as_tensor(flat, device=dev),
        )
        t = self._tables
        self._model = L.FjBeamModel(self.n_nodes, self.n_elem, *[t[k].data_ptr() for k in (
            "conn", "o0", "l0", "d0", "k_cs", "m_cs", "node_off", "node_elem")])
        self._ws = torch.empty(max(1, L.lib().fj_beam_workspace_bytes(self.n_elem) // 8), dtype=F64, device=dev)
        self._norms = torch.zeros(8, dtype=F64, device=dev)
        self._norms_host = torch.zeros(8, dtype=F64).pin_memory()

    def set_stiffness_scaling(self, scale):
        """Scale each cross-section stiffness of the k_cs table (one factor
        per table entry) for the following solves; the design's k_cs is
        kept.  The element table is rewritten in place on the device."""
        self._need_design()
        sc = _host(scale).reshape(-1)
        if sc.shape != (self.k_cs.shape[0],):
            raise ValueError(f"scale must have shape ({self.k_cs.shape[0]},)")
        k = self.k_cs * sc[:, None, None]
        self._tables["k_cs"].copy_(to_device(k[self.k_cs_index].reshape(-1) if self.n_elem else np.zeros(36)))
        self.stiffness_scaling = sc.copy()

    # ------------------------------------------------------------------ helpers
    def _need_design(self):
        if self._model is None:
            raise ValueError("set_design_variables has not been called")

    def make_prescribed_dofs_tuple(self, prescribed_dofs) -> tuple:
        """structure/beam.py:1858-1876."""
        if isinstance(prescribed_dofs, slice):
            return tuple(np.arange(self.n_dof)[prescribed_dofs].tolist())
        if prescribed_dofs is None:
            return ()
        if isinstance(prescribed_dofs, (int, np.integer)):
            return (int(prescribed_dofs),)
        if isinstance(prescribed_dofs, (Sequence, np.ndarray, torch.Tensor)):
            return tuple(int(x) for x in np.atleast_1d(_host(prescribed_dofs, np.int64)).tolist())
        raise TypeError("prescribed_dofs must be an int, slice, Sequence[int], or Array")

    def _solve_dofs(self, prescribed):
        pre = set(prescribed)
        sd = np.array(sorted(set(range(self.n_dof)) - pre), np.int32)
        self._bandwidth = self._reduced_bandwidth(sd)
        return torch.as_tensor(sd, device=self.hg0.device), len(sd)

    def _reduced_bandwidth(self, sd):
        """Half-bandwidth of the Newton matrix over the solve DOFs: the 12
        DOFs of an element couple with each other, a node's 6 with each
        other (T(phi)), nothing else."""
        pos = np.full(self.n_dof, -1, np.int64)
        pos[sd] = np.arange(len(sd))
        bw = 0
        groups = [np.arange(6 * a, 6 * a + 6) for a in range(self.n_nodes)]
        groups += [np.concatenate((np.arange(6 * a, 6 * a + 6), np.arange(6 * b, 6 * b + 6))) for a, b in self._conn]
        for g in groups:
            p = pos[g]
            p = p[p >= 0]
            if p.size:
                bw = max(bw, int(p.max() - p.min()))
        return bw

    def _newton_solve(self, A, r, n_s):
        """Newton increment A^-1 r: banded LU when the band fits one
        workgroup's LDS, else the dense LU (same pivot rows either way)."""
        if not n_s:
            return r
        bw = self._bandwidth
        key = (n_s, bw)
        if getattr(self, "_band_key", None) != key:
            self._band_key, self._band_ok = key, K.band_fits(n_s, bw, bw)
        if self._band_ok:
            return K.band_solve(A, r, bw, bw, out=self._sys_bufs[n_s][3])
        return K.lu_solve(K.lu_factor(A, overwrite=True), r)

    def _force(self, f, name, shape):
        if f is None:
            return None
        f = to_device(f)
        if tuple(f.shape) != shape:
            raise ValueError(f"{name} must have shape {shape}, got {tuple(f.shape)}")
        return f.contiguous()

    def _nodes(self, op, hg, x=None, hg_ref=None, hg_out=None, t_out=None, h_out=None):
        n = hg.numel() // 16
        rc = L.lib().fj_beam_nodes(op, n, L.ptr(hg), L.ptr(x), L.ptr(hg_ref), L.ptr(hg_out), L.ptr(t_out),
                                   L.ptr(h_out), L.stream())
        L.check(rc, "fj_beam_nodes")

    def varphi_from_hg(self, hg):
        """hg_to_d(hg0, hg) per node (structure/beam.py:591-597); hg (..., n_nodes, 4, 4)."""
        hg = hg.contiguous()
        ref = self.hg0.expand(hg.shape).contiguous()
        out = torch.empty(hg.shape[:-2] + (6,), dtype=F64, device=hg.device)
        self._nodes(L.BEAM_NODE_REL, ref, hg_ref=hg, h_out=out)
        return out

    def make_hg_dot(self, hg, v):
        """structure/beam.py:1521-1531."""
        out = torch.empty_like(hg)
        self._nodes(L.BEAM_NODE_RATE, hg.contiguous(), x=v.contiguous(), hg_out=out)
        return out

    def _sys_buffers(self, n_s, device):
        """(A, r, f_abs, x) of one Newton system size, reused by every Newton
        iteration (the previous iteration's system is consumed before the
        next is built)."""
        buf = self._sys_bufs.get(n_s) if hasattr(self, "_sys_bufs") else None
        if buf is None or buf[0].device != device:
            buf = (torch.empty((n_s, n_s), dtype=F64, device=device), torch.empty(n_s, dtype=F64, device=device),
                   torch.empty(n_s, dtype=F64, device=device), torch.empty(n_s, dtype=F64, device=device))
            if not hasattr(self, "_sys_bufs"):
                self._sys_bufs = {}
            self._sys_bufs[n_s] = buf
        return buf

    def _system(self, dynamic, hg, v, v_dot, t_phi, f_fol, f_dead, sd, n_s, ti=None):
        A, r, fa, _ = self._sys_buffers(n_s, hg.device)
        bp = ti.beta_prime if ti is not None else 0.0
        gp = ti.gamma_prime if ti is not None else 0.0
        rc = L.lib().fj_beam_system(ctypes.c_void_p(ctypes.addressof(self._model)), int(dynamic), L.ptr(hg), L.ptr(v), L.ptr(v_dot),
                                    L.ptr(t_phi), L.ptr(f_fol), L.ptr(f_dead), bp, gp, L.ptr(sd), n_s, L.ptr(A),
                                    L.ptr(r), L.ptr(fa), L.ptr(self._ws), L.stream())
        L.check(rc, "fj_beam_system")
        return A, r, fa

    def _wait(self):
        """Wait for the work enqueued so far on the current stream (an event
        wait returns sooner after the device finishes than a stream wait)."""
        ev = getattr(self, "_ev", None)
        if ev is None:
            ev = self._ev = torch.cuda.Event()
        ev.record()
        ev.synchronize()

    def _read_norms(self):
        self._norms_host.copy_(self._norms, non_blocking=True)
        self._wait()
        return self._norms_host.numpy()

    # ------------------------------------------------------------------ static
    def reference_configuration(self, use_f_ext_follower=True, use_f_ext_dead=True, use_f_aero=True,
                                use_f_grav=True, prescribed_dofs=()) -> StructureCase:
        """structure/beam.py:532-564."""
        self._need_design()
        z = lambda use: torch.zeros((self.n_nodes, 6), dtype=F64, device=self.hg0.device) if use else None  # noqa: E731
        return StructureCase(hg=self.hg0.clone(), varphi=z(True), t=np.zeros(1), f_ext_follower=z(use_f_ext_follower),
                             f_ext_dead=z(use_f_ext_dead), f_ext_aero=z(use_f_aero),
                             prescribed_dofs=tuple(prescribed_dofs))

    def static_solve(self, prescribed_dofs, f_ext_follower=None, f_ext_dead=None, f_ext_aero=None,
                     load_steps: int = 1) -> StructureCase:
        """structure/beam.py:1878-2092: Newton-Raphson per load step from hg0."""
        self._need_design()
        if load_steps < 1:
            raise ValueError("load_steps must be at least 1")
        if not (0.0 < self.relaxation_factor <= 1.0):
            raise ValueError("struct_relaxation_factor must be in the range (0, 1]")
        shp = (self.n_nodes, 6)
        f_fol = self._force(f_ext_follower, "f_ext_follower", shp)
        f_dead = self._force(f_ext_dead, "f_ext_dead", shp)
        f_aero = self._force(f_ext_aero, "f_ext_aero", shp)
        pre = self.make_prescribed_dofs_tuple(prescribed_dofs)
        sd, n_s = self._solve_dofs(pre)
        w = np.linspace(0.0, 1.0, load_steps + 1)[1:]
        hg = self.hg0.clone()
        hg_b = torch.empty_like(hg)
        dphi = torch.zeros(shp, dtype=F64, device=hg.device)
        h_full = torch.empty(shp, dtype=F64, device=hg.device)
        iters = []
        lib = L.lib()
        for i_ls in range(load_steps):
            ffol = None if f_fol is None else (w[i_ls] * f_fol).contiguous()
            fdead = None
            if f_dead is not None and f_aero is not None:
                fdead = (w[i_ls] * f_dead + w[i_ls] * f_aero).contiguous()
            elif f_dead is not None or f_aero is not None:
                fdead = (w[i_ls] * (f_dead if f_dead is not None else f_aero)).contiguous()
            st = ConvergenceStatus(self.struct_convergence_settings)
            while not st.get_status():
                A, r, fa = self._system(0, hg, None, None, None, ffol, fdead, sd, n_s)
                x = self._newton_solve(A, r, n_s)
                L.check(lib.fj_beam_update(0, self.n_nodes, L.ptr(sd), n_s, L.ptr(x), self.relaxation_factor, 0.0,
                                           0.0, L.ptr(dphi), None, None, L.ptr(r), L.ptr(fa), L.ptr(self._norms),
                                           L.stream()), "fj_beam_update")
                self._nodes(L.BEAM_NODE_UPDATE, hg, x=dphi, hg_ref=self.hg0, hg_out=hg_b, h_out=h_full)
                L.check(lib.fj_beam_subnorm(L.ptr(h_full), L.ptr(sd), n_s, L.ptr(self._norms[1:2]), L.stream()),
                        "fj_beam_subnorm")
                hg, hg_b = hg_b, hg
                nv = self._read_norms()
                st.update_norms(nv[0], nv[1], nv[2], nv[3], nv[4] != 0.0)
            iters.append(st.i_iter)
        return StructureCase(hg=hg, varphi=self.varphi_from_hg(hg), t=np.zeros(1), f_ext_follower=f_fol,
                             f_ext_dead=None if f_dead is None else self._to_local(f_dead, hg),
                             f_ext_aero=None if f_aero is None else self._to_local(f_aero, hg),
                             prescribed_dofs=pre, iters=iters)

    @staticmethod
    def _frame_map(op, f, hg, out=None):
        lead = torch.broadcast_shapes(f.shape[:-1], hg.shape[:-2])
        f = f.expand(lead + (6,)).contiguous()
        hg = hg.expand(lead + (4, 4)).contiguous()
        if out is None or not out.is_contiguous() or tuple(out.shape) != lead + (6,):
            out = torch.empty(lead + (6,), dtype=F64, device=f.device)
        L.check(L.lib().fj_beam_nodes(op, hg.numel() // 16, L.ptr(hg), L.ptr(f), None, None, None, L.ptr(out),
                                      L.stream()), "fj_beam_nodes")
        return out

    @staticmethod
    def _to_local(f, hg, out=None):
        """make_f_dead_ext (structure/beam.py:1387-1395): R^T f per node."""
        return BeamStructure._frame_map(L.BEAM_NODE_LOCAL, f, hg, out)

    @staticmethod
    def _to_global(f, hg):
        """R f per node (the inverse of ``_to_local``)."""
        return BeamStructure._frame_map(L.BEAM_NODE_GLOBAL, f, hg)

    # ----------------------------------------------------------------- dynamic
    #: Newton iterations enqueued before the first status read of a loop
    newton_batch = 3
    #: end a coupled step (one aero surface) with fj_beam_step_end
    fused_step_end = True

    def _conv_struct(self):
        s = self.struct_convergence_settings
        nan = float("nan")
        f = lambda v: nan if v is None else float(v)  # noqa: E731
        return L.FjConvergence(0 if s.max_n_iter is None else int(s.max_n_iter), f(s.rel_disp_tol),
                               f(s.abs_disp_tol), f(s.rel_force_tol), f(s.abs_force_tol))

    def _newton_dynamic_device(self, hg_prev, phi_a, v_a, vd_a, f_fol, f_dead, sd, n_s, ti, bufs, on_batch=None):
        """The Newton loop with the stop test on the device
        (fj_beam_newton_dynamic): iterations are enqueued in batches and the
        host reads one status word per batch; the iterates are those of the
        host-driven loop.  ``on_batch`` enqueues work behind every batch
        that the same synchronisation waits for."""
        hg_u, t_phi = bufs
        A, r, fa, x = self._sys_buffers(n_s, hg_prev.device)
        if not hasattr(self, "_nstatus"):
            self._nstatus = torch.zeros(4, dtype=torch.int32, device=hg_prev.device)
            self._nstatus_host = torch.zeros(4, dtype=torch.int32).pin_memory()
        conv = self._conv_struct()
        bw = self._bandwidth
        lib = L.lib()
        batch, reset = self.newton_batch, 1
        while True:
            rc = lib.fj_beam_newton_dynamic(
                ctypes.c_void_p(ctypes.addressof(self._model)), batch, reset, L.ptr(hg_prev), L.ptr(phi_a),
                L.ptr(v_a), L.ptr(vd_a), L.ptr(f_fol), L.ptr(f_dead), ti.beta_prime, ti.gamma_prime,
                self.relaxation_factor, L.ptr(sd), n_s, bw, bw, ctypes.byref(conv), L.ptr(hg_u), L.ptr(t_phi),
                L.ptr(A), L.ptr(r), L.ptr(fa), L.ptr(x), L.ptr(self._norms), L.ptr(self._nstatus), None,
                L.ptr(self._ws), L.stream())
            L.check(rc, "fj_beam_newton_dynamic")
            reset = 0
            self._nstatus_host.copy_(self._nstatus, non_blocking=True)
            if on_batch is not None:
                on_batch()
            self._wait()
            it, done = int(self._nstatus_host[0]), int(self._nstatus_host[1])
            if done:
                return it
            batch = 2

    def _newton_dynamic(self, hg_prev, phi_a, v_a, vd_a, f_fol, f_dead, sd, n_s, ti, bufs, on_batch=None):
        """The Newton loop of one load step at the alpha point
        (structure/beam.py:2170-2316, 2727-2792); phi_a, v_a, vd_a are updated
        in place.  ``on_batch`` (optional) enqueues work that reads the
        final iterate: on return it has run behind the last iteration and
        the stream has been synchronised."""
        if n_s and self.device_newton and K.band_fits(n_s, self._bandwidth, self._bandwidth):
            return self._newton_dynamic_device(hg_prev, phi_a, v_a, vd_a, f_fol, f_dead, sd, n_s, ti, bufs, on_batch)
        hg_u, t_phi = bufs
        lib = L.lib()
        st = ConvergenceStatus(self.struct_convergence_settings)
        while not st.get_status():
            self._nodes(L.BEAM_NODE_UPDATE, hg_prev, x=phi_a, hg_out=hg_u, t_out=t_phi)
            A, r, fa = self._system(1, hg_u, v_a, vd_a, t_phi, f_fol, f_dead, sd, n_s, ti)
            x = self._newton_solve(A, r, n_s)
            L.check(lib.fj_beam_update(1, self.n_nodes, L.ptr(sd), n_s, L.ptr(x), self.relaxation_factor,
                                       ti.gamma_prime, ti.beta_prime, L.ptr(phi_a), L.ptr(v_a), L.ptr(vd_a), L.ptr(r),
                                       L.ptr(fa), L.ptr(self._norms), L.stream()), "fj_beam_update")
            nv = self._read_norms()
            st.update_norms(nv[0], nv[1], nv[2], nv[3], nv[4] != 0.0)
        if on_batch is not None:
            on_batch()
            self._wait()
        return st.i_iter

    def dynamic_solve(self, init_state, prescribed_dofs, n_tstep: int, dt, f_ext_follower=None, f_ext_dead=None,
                      f_ext_aero=None, thrust_t=None, load_steps: int = 1) -> StructureCase:
        """structure/beam.py:2863-2990 (structure only)."""
        self._need_design()
        if load_steps <= 0:
            raise ValueError("load_steps must be a positive integer")
        if thrust_t:
            raise NotImplementedError("thrust forces are not on the device path")

        def series(f, name):
            if f is None:
                return None
            f = to_device(f)
            if f.dim() == 2:
                f = f.expand(n_tstep, self.n_nodes, 6)
            if tuple(f.shape) != (n_tstep, self.n_nodes, 6):
                raise ValueError(f"{name} must have shape [n_node, 6] or [n_tstep, n_node, 6]")
            return f.contiguous()

        f_fol = series(f_ext_follower, "f_ext_follower")
        f_dead = series(f_ext_dead, "f_ext_dead")
        series(f_ext_aero, "f_ext_aero")  # checked, not applied (as the reference)
        init = (self.reference_configuration() if init_state is None else init_state).to_dynamic()
        t0 = float(np.atleast_1d(init.t)[0]) if init_state is not None and init_state.is_dynamic else 0.0
        pre = self.make_prescribed_dofs_tuple(prescribed_dofs)
        return self.base_dynamic_solve(init, t0 + float(dt) * np.arange(n_tstep), float(dt), pre, load_steps, f_fol,
                                       f_dead)

    def base_dynamic_solve(self, init: StructureCase, t, dt, prescribed, load_steps, f_fol_t=None, f_dead_t=None,
                           aero=None, fsi_settings=None) -> StructureCase:
        """structure/beam.py:2129-2861.  ``aero``: the coupling hooks of
        ``CoupledAeroelastic`` (aero solve and force projection), or None."""
        if not (0.0 < self.relaxation_factor <= 1.0):
            raise ValueError("Relaxation factor must be in range (0, 1]")
        n_tstep = len(t)
        nn = self.n_nodes
        dev = self.hg0.device
        ti = TimeIntegrator(self.spectral_radius, dt)
        self.time_integrator = ti
        sd, n_s = self._solve_dofs(prescribed)
        w = np.linspace(0.0, 1.0, load_steps + 1)[1:]

        def alpha_series(f):
            if f is None:
                return None
            fa = torch.zeros_like(f)
            fa[1:] = ti.f_alpha(f[:-1], f[1:])
            return fa

        ffa, fda = alpha_series(f_fol_t), alpha_series(f_dead_t)
        hg = torch.zeros((n_tstep, nn, 4, 4), dtype=F64, device=dev)
        v = torch.zeros((n_tstep, nn, 6), dtype=F64, device=dev)
        vd = torch.zeros_like(v)
        acc = torch.zeros_like(v)
        hg[0] = to_device(init.hg)
        v[0] = to_device(init.v)
        vd[0] = to_device(init.v_dot)
        acc[0] = vd[0]  # structure/beam.py:2970
        f_aero_g = f_aero_l = None
        if aero is not None:
            f_aero_g = torch.zeros((n_tstep, nn, 6), dtype=F64, device=dev)
            f_aero_g[0] = aero.initial_force(hg[0])
            f_aero_l = torch.zeros_like(f_aero_g)
            f_aero_l[0] = self._to_local(f_aero_g[0], hg[0])
        bufs = (torch.empty((nn, 4, 4), dtype=F64, device=dev), torch.empty((nn, 6, 6), dtype=F64, device=dev))
        struct_iters, fsi_iters = [], []
        af = ti.alpha_f

        def load_step_loop(i, phi_a, v_a, vd_a, f_aero_steps, on_last=None):
            for i_ls in range(load_steps):
                ffol = None if ffa is None else (w[i_ls] * ffa[i]).contiguous()
                fdead = None
                if fda is not None and f_aero_steps is not None:
                    fdead = (w[i_ls] * fda[i] + f_aero_steps[i_ls]).contiguous()
                elif fda is not None:
                    fdead = (w[i_ls] * fda[i]).contiguous()
                elif f_aero_steps is not None:
                    fdead = f_aero_steps[i_ls].contiguous()
                struct_iters.append(self._newton_dynamic(hg[i - 1], phi_a, v_a, vd_a, ffol, fdead, sd, n_s, ti,
                                                         bufs, on_last if i_ls == load_steps - 1 else None))

        lib = L.lib()
        n6 = 6 * nn
        am, g_, b_ = ti.alpha_m, ti.gamma, ti.beta

        def ga(op, ins, outs):
            # the generalised-alpha bookkeeping in one launch (fj_beam_ga)
            p = [L.ptr(x) for x in ins] + [None] * (6 - len(ins))
            o = [L.ptr(x) for x in outs] + [None] * (4 - len(outs))
            L.check(lib.fj_beam_ga(op, n6, am, af, g_, b_, ti.dt, *p, *o, L.stream()), "fj_beam_ga")

        phi_n = torch.empty((nn, 6), dtype=F64, device=dev)
        v_n = torch.empty_like(phi_n)
        phi_old = torch.empty_like(phi_n)
        phi = torch.empty_like(phi_n)
        nrm_dev = torch.empty(5, dtype=F64, device=dev)
        hg_fsi = (torch.empty((nn, 4, 4), dtype=F64, device=dev), torch.empty((nn, 4, 4), dtype=F64, device=dev))
        nrm_host = torch.empty(5, dtype=F64).pin_memory()
        fsi = {}

        def fsi_norms():
            # the FSI norms of this iteration, read at the Newton loop's last
            # status synchronisation
            L.check(lib.fj_beam_fsi_norms(n_s, L.ptr(sd), L.ptr(fsi["phi_old"]), L.ptr(fsi["phi_a"]),
                                          L.ptr(fsi["f_al"]), L.ptr(fsi["f_prev"]), L.ptr(nrm_dev), L.stream()),
                    "fj_beam_fsi_norms")
            nrm_host.copy_(nrm_dev, non_blocking=True)
        predicted = None
        for i in range(1, n_tstep):
            # predict_q + calculate_q_alpha (structure/time_integration.py:86-150)
            if predicted is not None:
                phi_a, v_a, vd_a = predicted  # by the previous step's fj_beam_step_end
                predicted = None
            else:
                phi_a, v_a, vd_a = (torch.empty((nn, 6), dtype=F64, device=dev) for _ in range(3))
                ga(L.GA_PREDICT, (v[i - 1], vd[i - 1], acc[i - 1]), (phi_a, v_a, vd_a))
            if aero is None:
                load_step_loop(i, phi_a, v_a, vd_a, None)
            else:
                # coupled/coupled.py fsi loop (structure/beam.py:2612-2725)
                f_nm1 = f_aero_g[i - 1]
                f_prev = f_nm1
                fst = ConvergenceStatus(fsi_settings)
                while not fst.get_status():
                    # phi_n, v_n, the frames hg_n and their rates in one launch
                    # (the aero step copies what it keeps of them)
                    hg_n, hg_dot = hg_fsi
                    L.check(lib.fj_beam_ga_nodes(L.GA_ALPHA, nn, am, af, L.ptr(phi_a), L.ptr(v_a), None,
                                                 L.ptr(v[i - 1]), None, None, L.ptr(hg[i - 1]), L.ptr(phi_n),
                                                 L.ptr(v_n), None, None, L.ptr(phi_old), L.ptr(hg_n),
                                                 L.ptr(hg_dot), L.stream()), "fj_beam_ga_nodes")
                    f_n = aero.solve(i, hg_n, hg[i - 1], hg_dot).contiguous()
                    f_al = torch.empty((nn, 6), dtype=F64, device=dev)
                    ga(L.GA_FORCE, (f_nm1, f_n), (f_al,))
                    steps = [f_al] if load_steps == 1 else [(wk * f_al).contiguous() for wk in w]
                    fsi.update(phi_old=phi_old, phi_a=phi_a, f_al=f_al, f_prev=f_prev)
                    load_step_loop(i, phi_a, v_a, vd_a, steps, fsi_norms)
                    nrm = nrm_host.numpy()
                    fst.update_norms(*nrm[:4], nrm[4] != 0.0)
                    f_prev = f_al
                fsi_iters.append(fst.i_iter)
            # calculate_q_n_from_q_alpha + calculate_a_n (structure/time_integration.py:53-84)
            # and the step's frames, in one launch; with one aero surface
            # also the force projection, its local copy and the next step's
            # predictor (fj_beam_step_end: the same arithmetic, one launch
            # instead of six)
            se = None
            if aero is not None and self.fused_step_end and nn <= 512 and hasattr(aero, "step_end_inputs"):
                se = aero.step_end_inputs(i)
            if se is not None:
                fa_i, fb_i, x0_i, dof_i = se
                nxt = None
                if i + 1 < n_tstep:
                    nxt = tuple(torch.empty((nn, 6), dtype=F64, device=dev) for _ in range(3))
                m1, n1 = fa_i.shape[0], fa_i.shape[1]
                L.check(lib.fj_beam_step_end(
                    nn, am, af, g_, b_, ti.dt, L.ptr(phi_a), L.ptr(v_a), L.ptr(vd_a), L.ptr(v[i - 1]),
                    L.ptr(vd[i - 1]), L.ptr(acc[i - 1]), L.ptr(hg[i - 1]), m1, n1, L.ptr(fa_i.contiguous()),
                    L.ptr(None if fb_i is None else fb_i.contiguous()), L.ptr(x0_i.contiguous()), L.ptr(dof_i),
                    L.ptr(phi), L.ptr(v[i]), L.ptr(vd[i]), L.ptr(acc[i]), L.ptr(hg[i]), L.ptr(f_aero_g[i]),
                    L.ptr(f_aero_l[i]), *(L.ptr(x) for x in (nxt or (None, None, None))), L.stream()),
                    "fj_beam_step_end")
                predicted = nxt
                continue
            L.check(lib.fj_beam_ga_nodes(L.GA_FINISH, nn, am, af, L.ptr(phi_a), L.ptr(v_a), L.ptr(vd_a),
                                         L.ptr(v[i - 1]), L.ptr(vd[i - 1]), L.ptr(acc[i - 1]), L.ptr(hg[i - 1]),
                                         L.ptr(phi), L.ptr(v[i]), L.ptr(vd[i]), L.ptr(acc[i]), None, L.ptr(hg[i]),
                                         None, L.stream()), "fj_beam_ga_nodes")
            if aero is not None:
                fg = aero.project(i, hg[i], out=f_aero_g[i])
                if fg.data_ptr() != f_aero_g[i].data_ptr():
                    f_aero_g[i] = fg
                fl = self._to_local(f_aero_g[i], hg[i], out=f_aero_l[i])
                if fl.data_ptr() != f_aero_l[i].data_ptr():
                    f_aero_l[i] = fl
        out = StructureCase(hg=hg, varphi=self.varphi_from_hg(hg), t=np.asarray(t, float), v=v, v_dot=vd, a=acc,
                            f_ext_follower=f_fol_t,
                            f_ext_dead=None if f_dead_t is None else self._to_local(f_dead_t, hg),
                            f_ext_aero=f_aero_l, prescribed_dofs=tuple(prescribed),
                            iters=dict(struct=struct_iters, fsi=fsi_iters))
        return out
