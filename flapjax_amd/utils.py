"""Segment kernels, grid stencils, forces and wake convection
(aero/utils.py of the reference), on the HIP path.

The reference's ``KernelFunction`` plug-in (aero/utils.py:19) becomes a
``SegmentKernel`` descriptor: the three kernels the reference ships map to a
kernel kind the HIP code evaluates in bulk.  Calling a descriptor on one
(target, filament) pair evaluates that pair on the device, so the reference's
kernel unit tests read the same here.  Any other Python callable has no
device implementation and raises ``NotImplementedError``.
"""

from __future__ import annotations

from collections.abc import Callable, Sequence

import numpy as np
import torch

from . import _kernels as K
from . import _lib as L

# utils/constants.py:13-19
HORSESHOE_LENGTH = 1e3
EPSILON = 1e-7
R_CUTOFF = 1e-4

F64 = torch.float64


class SegmentKernel:
    """A Biot-Savart segment kernel the device path implements."""

    def __init__(self, kind: int, name: str, eps: float = EPSILON, rcut: float = R_CUTOFF):
        self.kind = int(kind)
        self.name = name
        self.eps = float(eps)
        self.rcut = float(rcut)

    def __repr__(self):
        return f"SegmentKernel({self.name}, eps={self.eps:g}, rcut={self.rcut:g})"

    def __call__(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        """Influence per unit circulation at x (3,) of the filament y (2, 3)."""
        x = torch.as_tensor(x, dtype=F64, device=_device()).reshape(1, 3)
        y = torch.as_tensor(y, dtype=F64, device=_device()).reshape(1, 2, 3)
        seg = K.segments_from_endpoints(y, None, self.kind, self.eps)
        return K.pair_sum(seg, x, self.kind, self.eps, self.rcut).reshape(3)


# aero/utils.py:580-648
biot_savart = SegmentKernel(L.KERNEL_PLAIN, "biot_savart")
biot_savart_epsilon = SegmentKernel(L.KERNEL_EPSILON, "biot_savart_epsilon")
biot_savart_cutoff = SegmentKernel(L.KERNEL_CUTOFF, "biot_savart_cutoff")

KernelFunction = SegmentKernel


def as_segment_kernel(kernel) -> SegmentKernel:
    if kernel is None:
        return biot_savart_epsilon
    if isinstance(kernel, SegmentKernel):
        return kernel
    raise NotImplementedError(
        "only the reference's biot_savart / biot_savart_epsilon / biot_savart_cutoff kernels have a "
        f"device implementation (got {kernel!r})"
    )


_gpu_seen = False


def _device():
    global _gpu_seen
    if not _gpu_seen:
        if not torch.cuda.is_available():
            raise L.FlapjaxLibError("flapjax_amd needs a ROCm GPU; none is visible (there is no CPU fallback)")
        _gpu_seen = True
    return torch.device("cuda", torch.cuda.current_device())


def to_device(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        # already what the kernels take (the common case inside a time loop)
        if x.is_cuda and x.dtype == F64 and _gpu_seen and x.get_device() == torch.cuda.current_device() \
                and x.is_contiguous():
            return x
        return x.to(device=_device(), dtype=F64).contiguous()
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=_device()).contiguous()


# ---------------------------------------------------------------------------
# grids (aero/utils.py:23-36, 87-123, 651-663)
# ---------------------------------------------------------------------------
def make_rectangular_grid(m: int, n: int, chord: float, ea: float) -> torch.Tensor:
    """aero/utils.py:23-36 (host-built geometry, moved to the device)."""
    grid = np.zeros((m + 1, n + 1, 3))
    grid[..., 0] = (np.linspace(0.0, chord, m + 1) - ea * chord)[:, None]
    return to_device(grid)


def compute_surf_c(zeta: torch.Tensor) -> torch.Tensor:
    """Collocation points, aero/utils.py:87-93."""
    return K.colloc_normals(to_device(zeta), want_n=False)[0]


def compute_surf_nc(zeta: torch.Tensor) -> torch.Tensor:
    """Panel normals of length 2 * area, aero/utils.py:96-105."""
    return K.colloc_normals(to_device(zeta), want_c=False)[1]


def compute_c(zetas):
    return [compute_surf_c(z) for z in zetas]


def compute_nc(zetas):
    return [compute_surf_nc(z) for z in zetas]


def mirror_grid(zeta: torch.Tensor, mirror_point, mirror_normal) -> torch.Tensor:
    """aero/utils.py:651-663 (geometry set-up helper)."""
    zeta = to_device(zeta)
    p = to_device(mirror_point)
    nrm = to_device(mirror_normal)
    dn = ((zeta - p) * nrm).sum(-1, keepdim=True)
    return zeta - 2.0 * dn * nrm


# ---------------------------------------------------------------------------
# control surfaces: building blocks of a UVLM ``grid_func``
# ---------------------------------------------------------------------------
HINGE_AXIS_DEFAULT = (0.0, 1.0, 0.0)  # aero/utils.py:39


def _index_rows(index, n_entries: int) -> torch.Tensor:
    """utils/utils.py:93-113 (index_to_arr) on the host."""
    if isinstance(index, slice):
        return torch.arange(n_entries)[index]
    if isinstance(index, torch.Tensor):
        return index.detach().cpu().long()
    if isinstance(index, np.ndarray):
        return torch.as_tensor(index, dtype=torch.long)
    if isinstance(index, Sequence):
        return torch.as_tensor(list(index), dtype=torch.long)
    if isinstance(index, (int, np.integer)):
        return torch.tensor([int(index)])
    if index is None:
        return torch.arange(n_entries)
    raise TypeError("index must be a slices, sequence of ints, or Array")


def add_control_surface(grid: torch.Tensor, angle, m_slice, n_slice, hinge_axis=HINGE_AXIS_DEFAULT) -> torch.Tensor:
    """aero/utils.py:42-84: rotate the vertices ``grid[m_slice, n_slice]``
    by ``angle`` about ``hinge_axis`` through each spanwise column's hinge
    point (its first chordwise vertex of the slice).  Differentiable in
    ``angle`` (torch autograd and torch.func.jvp), so a grid function built
    from it gives the control-surface grid velocities of
    ``UVLM.hg_dot_to_zeta_b_dot``."""
    from .beam_adjoint import exp_so3

    grid = to_device(grid)
    ms = _index_rows(m_slice, grid.shape[0]).to(grid.device)
    ns = _index_rows(n_slice, grid.shape[1]).to(grid.device)
    ang = angle if isinstance(angle, torch.Tensor) else torch.as_tensor(float(angle), dtype=grid.dtype)
    ang = ang.to(device=grid.device, dtype=grid.dtype).reshape(())
    axis = torch.as_tensor(hinge_axis, dtype=grid.dtype, device=grid.device)
    rmat = exp_so3(axis * ang)
    sub = grid.index_select(0, ms).index_select(1, ns)  # (ms, ns, 3)
    hinge = sub[:1]  # (1, ns, 3)
    rot = torch.einsum("ij,hnj->hni", rmat, sub - hinge) + hinge
    out = grid.clone()
    out[ms[:, None], ns[None, :]] = rot
    return out


def cs_ang_to_cs_vel(cs_ang_t: dict, dt: float) -> dict:
    """aero/utils.py:700-720: control velocities from the angle histories by
    first-order backward differences (zero at the first step, which has no
    predecessor)."""
    out = {}
    for k, v in cs_ang_t.items():
        v = to_device(v)
        d = torch.zeros_like(v)
        if v.shape[0] > 1:
            d[1:] = (v[1:] - v[:-1]) / float(dt)
        out[k] = d
    return out


# ---------------------------------------------------------------------------
# velocity closures the fused device path recognises
# ---------------------------------------------------------------------------
class FlowAt:
    """x -> U(x, t): the freestream closure of aero/uvlm.py:1045-1046."""

    def __init__(self, flowfield, t):
        self.flowfield = flowfield
        self.t = float(t)

    def __call__(self, x):
        return self.flowfield.vmap_call(x, self.t)


class InducedVelocity:
    """x -> U(x, t) + sum of filament influences: the closures of
    aero/uvlm.py:917-929 (free-wake convection) and :1023-1033 (forces).
    ``segments`` maps (kernel kind, eps, rcut) -> filament table
    (aic.segment_tables)."""

    def __init__(self, flowfield, t, segments, kernel: SegmentKernel, precision=L.PREC_FP64):
        self.flowfield = flowfield
        self.t = float(t)
        self.segments = segments
        self.kernel = kernel
        self.precision = precision

    def induced(self, x, out=None, mode=L.OUT_STORE, flow=None, dt=0.0):
        x = to_device(x)
        flat = x.reshape(-1, 3)
        if mode == L.OUT_CONVECT and len(self.segments) > 1:
            raise NotImplementedError("fused convection needs one kernel kind for all surfaces")
        first = True
        for (kind, eps, rcut), seg in self.segments.items():
            m = mode if first else L.OUT_ADD
            out = K.pair_sum(seg, flat, kind, eps, rcut, self.precision, m, out=out, flow=flow, time=self.t, dt=dt)
            first = False
        if first:
            out = K.pair_sum(None, flat, self.kernel.kind, self.kernel.eps, self.kernel.rcut, self.precision, mode,
                             out=out, flow=flow, time=self.t, dt=dt)
        return out

    def __call__(self, x):
        x = to_device(x)
        v = self.induced(x)
        return (v + self.flowfield.vmap_call(x.reshape(-1, 3), self.t)).reshape(x.shape)


# ---------------------------------------------------------------------------
# forces (aero/utils.py:126-273)
# ---------------------------------------------------------------------------
def calculate_steady_forcing(zeta_b, zeta_dot_b, gamma_b, gamma_w, rho, v_func: Callable, v_inputs=None):
    """aero/utils.py:126-223.  ``v_func`` gives the total velocity at points;
    an ``InducedVelocity`` closure is evaluated by the fused kernels."""
    if v_inputs is not None:
        raise NotImplementedError("v_inputs (linearised UVLM perturbations) are not on the device path")
    out = []
    zdots = zeta_dot_b if zeta_dot_b is not None else [None] * len(zeta_b)
    for zb, zd, gb, gw in zip(zeta_b, zdots, gamma_b, gamma_w):
        zb = to_device(zb)
        mid = K.filament_midpoints(zb)
        gw0 = to_device(gw)[0].contiguous() if gw.shape[0] > 0 else None
        if isinstance(v_func, InducedVelocity):
            v = v_func.induced(mid)
            f = K.steady_forces(zb, to_device(zd) if zd is not None else None, to_device(gb), gw0, v,
                                v_func.flowfield, v_func.t, rho)
        else:
            v = to_device(v_func(mid)).reshape(-1, 3).contiguous()
            f = K.steady_forces(zb, to_device(zd) if zd is not None else None, to_device(gb), gw0, v, None, 0.0, rho)
        out.append(f)
    return out


def strip_alpha(zeta_b, f_steady, v_func: FlowAt, rho):
    """aero/utils.py:226-273 (v_func must be the freestream closure)."""
    if not isinstance(v_func, FlowAt):
        raise NotImplementedError("strip_alpha evaluates the freestream on the device; pass utils.FlowAt")
    return [K.strip_alpha(to_device(z), to_device(f), v_func.flowfield, v_func.t, rho) for z, f in zip(zeta_b, f_steady)]


# ---------------------------------------------------------------------------
# wake convection (aero/utils.py:400-577)
# ---------------------------------------------------------------------------
def propagate_surf_wake(gamma_b_nm1, gamma_w_nm1, zeta_b_n, zeta_w_nm1, delta_w, v_func, dt, frozen_wake,
                        linearise_variable_wake=False):
    """aero/utils.py:400-500 (uniform and variable discretisation)."""
    if delta_w is not None:
        return _propagate_variable(to_device(gamma_b_nm1), to_device(gamma_w_nm1), to_device(zeta_b_n),
                                   to_device(zeta_w_nm1), delta_w, v_func, dt, frozen_wake)
    gamma_b_nm1 = to_device(gamma_b_nm1)
    gamma_w_nm1 = to_device(gamma_w_nm1)
    zeta_b_n = to_device(zeta_b_n)
    zeta_w_nm1 = to_device(zeta_w_nm1)
    m_star = zeta_w_nm1.shape[0] - 1
    zeta_w_n = torch.empty_like(zeta_w_nm1)
    gamma_w_n = torch.empty_like(gamma_w_nm1)
    K.wake_shift(gamma_b_nm1, gamma_w_nm1, zeta_b_n, m_star, gamma_w_n, zeta_w_n)
    if m_star > 0:
        base = zeta_w_nm1[:-1].reshape(-1, 3)
        if isinstance(v_func, InducedVelocity):
            v_func.induced(base, out=zeta_w_n[1:].reshape(-1, 3), mode=L.OUT_CONVECT, flow=v_func.flowfield, dt=dt)
        elif isinstance(v_func, FlowAt):
            K.wake_convect(zeta_w_nm1, None, v_func.flowfield, v_func.t, dt, zeta_w_n)
        else:
            vel = to_device(v_func(zeta_w_nm1[:-1])).reshape(-1, 3).contiguous()
            K.wake_convect(zeta_w_nm1, vel, None, 0.0, dt, zeta_w_n)
    if frozen_wake:
        return None, gamma_w_n
    return zeta_w_n, gamma_w_n


def _convect_rows(zeta_rows, v_func, dt, dst):
    """dst = zeta_rows + dt * v_func(zeta_rows) on the device (fused for the
    package's velocity closures)."""
    flat = zeta_rows.reshape(-1, 3)
    out = dst.reshape(-1, 3)
    if isinstance(v_func, InducedVelocity):
        v_func.induced(flat, out=out, mode=L.OUT_CONVECT, flow=v_func.flowfield, dt=dt)
    else:
        # fj_wake_convect maps rows 0..r-1 of a (r+1)-row source onto rows
        # 1..r of a separate destination
        n_rows, ncol = zeta_rows.shape[0], zeta_rows.shape[1] - 1
        src = torch.empty((n_rows + 1, ncol + 1, 3), dtype=F64, device=zeta_rows.device)
        src[:-1] = zeta_rows
        res = torch.empty_like(src)
        if isinstance(v_func, FlowAt):
            K.wake_convect(src, None, v_func.flowfield, v_func.t, dt, res)
        else:
            vel = to_device(v_func(zeta_rows)).reshape(-1, 3).contiguous()
            K.wake_convect(src, vel, None, 0.0, dt, res)
        out.copy_(res[1:].reshape(-1, 3))


def _propagate_variable(gamma_b_nm1, gamma_w_nm1, zeta_b_n, zeta_w_nm1, delta_w, v_func, dt, frozen_wake):
    """Variable discretisation (aero/utils.py:431-491): all m_star+1 rows are
    convected, then every column is re-sampled by arc length."""
    m_star, n = zeta_w_nm1.shape[0] - 1, zeta_w_nm1.shape[1] - 1
    dw = np.asarray(delta_w.detach().cpu() if isinstance(delta_w, torch.Tensor) else delta_w, dtype=np.float64)
    if dw.shape != (m_star,):
        raise ValueError(f"delta_w must have shape ({m_star},)")
    s_disc = to_device(np.concatenate((np.zeros(1), np.cumsum(dw))))
    zeta_pre = torch.empty((m_star + 2, n + 1, 3), dtype=F64, device=zeta_w_nm1.device)
    zeta_pre[0] = zeta_b_n[-1]
    _convect_rows(zeta_w_nm1, v_func, dt, zeta_pre[1:])
    gamma_pre = torch.cat((gamma_b_nm1[-1:], gamma_w_nm1), dim=0).contiguous()
    zeta_w_n, gamma_w_n = K.wake_rediscretise(zeta_pre, gamma_pre, s_disc)
    if m_star == 0:
        gamma_w_n = torch.zeros_like(gamma_w_nm1)
    if frozen_wake:
        return None, gamma_w_n
    return zeta_w_n, gamma_w_n


def propagate_wake(gamma_b_nm1, gamma_w_nm1, zeta_b_n, zeta_w_nm1, delta_w: Sequence, v_func, dt, frozen_wake,
                   linearise_variable_wake=False):
    """aero/utils.py:531-577."""
    zs = None if frozen_wake else []
    gs = []
    for i in range(len(gamma_b_nm1)):
        z, g = propagate_surf_wake(gamma_b_nm1[i], gamma_w_nm1[i], zeta_b_n[i], zeta_w_nm1[i], delta_w[i], v_func, dt,
                                   frozen_wake, linearise_variable_wake)
        if zs is not None:
            zs.append(z)
        gs.append(g)
    return zs, gs
