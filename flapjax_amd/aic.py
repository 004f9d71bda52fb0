"""Pair sums and AIC assembly (aero/aic.py of the reference) on the HIP path.

Signatures follow the reference; ``batch_size`` is accepted and ignored (the
device kernels tile targets x filaments themselves).  Arrays are torch
tensors on the GPU; an ``ArrayList`` is a Python list of tensors.
"""

from __future__ import annotations

from collections.abc import Sequence

import torch

from . import _kernels as K
from . import _lib as L
from .utils import SegmentKernel, as_segment_kernel, to_device

F64 = torch.float64


def _mirror(mirror_point, mirror_normal):
    if mirror_point is None or mirror_normal is None:
        return None
    p = mirror_point.detach().cpu().numpy() if isinstance(mirror_point, torch.Tensor) else mirror_point
    nrm = mirror_normal.detach().cpu().numpy() if isinstance(mirror_normal, torch.Tensor) else mirror_normal
    import numpy as np

    nrm = np.asarray(nrm, dtype=np.float64)
    return (np.asarray(p, dtype=np.float64), nrm / np.linalg.norm(nrm))


def segment_tables(zetas, gammas, kernels, mirror=None) -> dict:
    """De-duplicated filament tables of all source surfaces (and their mirror
    images, with negated circulation), grouped by kernel kind
    (aero/aic.py:276-308)."""
    if isinstance(kernels, SegmentKernel) or kernels is None:
        kernels = [kernels] * len(zetas)
    groups: dict = {}
    lats: dict = {}
    for zeta, gamma, kern in zip(zetas, gammas, kernels):
        kern = as_segment_kernel(kern)
        zeta = to_device(zeta)
        gamma = to_device(gamma)
        if zeta.numel() == 0 or gamma.numel() == 0:
            continue
        tabs = [K.segments_from_grid(zeta, gamma, kern.kind, kern.eps)]
        if mirror is not None:
            tabs.append(K.segments_from_grid(zeta, gamma, kern.kind, kern.eps, mirror=mirror))
        key = (kern.kind, kern.eps, kern.rcut)
        m, n = zeta.shape[0] - 1, zeta.shape[1] - 1
        for t in tabs:
            off = sum(x.shape[0] for x in groups.get(key, []))
            lats.setdefault(key, []).append((off, m, n))
            groups.setdefault(key, []).append(t)
    out = {}
    for k, v in groups.items():
        t = torch.cat(v, dim=0)
        # the table holds whole lattices: the FP64 pair sums run the lattice
        # kernel on it (_kernels.pair_sum)
        t.fj_lattices = lats[k]
        out[k] = t
    return out


def _pair_sum_tables(tables, targets, precision=L.PREC_FP64):
    flat = to_device(targets).reshape(-1, 3)
    out = None
    for (kind, eps, rcut), seg in tables.items():
        out = K.pair_sum(seg, flat, kind, eps, rcut, precision, L.OUT_STORE if out is None else L.OUT_ADD, out=out)
    if out is None:
        out = torch.zeros_like(flat)
    return out


def v_ind_vmap(c, zeta, gamma, kernel, batch_size=None):
    """aero/aic.py:208-245: c (c_m, c_n, 3), filament grid zeta (zm, zn, 2, 3),
    circulations gamma (zm, zn) -> (c_m, c_n, 3)."""
    kern = as_segment_kernel(kernel)
    c = to_device(c)
    zeta = to_device(zeta)
    if zeta.numel() == 0:
        return torch.zeros_like(c)
    seg = K.segments_from_endpoints(zeta.reshape(-1, 2, 3), to_device(gamma).reshape(-1), kern.kind, kern.eps)
    return K.pair_sum(seg, c.reshape(-1, 3), kern.kind, kern.eps, kern.rcut).reshape(c.shape)


def compute_v_ind(cs, zetas, gammas, kernels, mirror_point=None, mirror_normal=None, batch_size=None,
                  precision: int = L.PREC_FP64):
    """aero/aic.py:248-310: induced velocity of all source surfaces at one
    target grid (tensor) or several (list), returned in the same structure."""
    tables = segment_tables(zetas, gammas, kernels, _mirror(mirror_point, mirror_normal))
    if isinstance(cs, torch.Tensor):
        return _pair_sum_tables(tables, cs, precision).reshape(cs.shape)
    return [_pair_sum_tables(tables, c, precision).reshape(to_device(c).shape) for c in cs]


def compute_aic_grid(c, n, zeta, kernel, batch_size=None, mirror=None):
    """aero/aic.py:11-63: (c_m, c_n, m, n) normal-projected ring influence."""
    if n is None:
        raise NotImplementedError("the device AIC is always normal-projected")
    kern = as_segment_kernel(kernel)
    c = to_device(c)
    n = to_device(n)
    zeta = to_device(zeta)
    c_m, c_n = c.shape[:2]
    m, nn = zeta.shape[0] - 1, zeta.shape[1] - 1
    out = torch.zeros((c_m * c_n, m * nn), dtype=F64, device=c.device)
    if c_m * c_n and m * nn:
        K.aic_assemble(c.reshape(-1, 3), n.reshape(-1, 3), zeta, kern.kind, kern.eps, kern.rcut, out, 0,
                       mirror=mirror)
    return out.reshape(c_m, c_n, m, nn)


def compute_aic_sys(zetas, cs, ns, kernels, batch_size=None, mirror_point=None, mirror_normal=None):
    """aero/aic.py:66-112: nested list [target][source] of AIC blocks; a mirror
    image is subtracted inside the same kernel."""
    if isinstance(kernels, SegmentKernel) or kernels is None:
        kernels = [kernels] * len(zetas)
    mirror = _mirror(mirror_point, mirror_normal)
    return [[compute_aic_grid(c, n, z, k, mirror=mirror) for z, k in zip(zetas, kernels)] for c, n in zip(cs, ns)]


def reshape_aic_sys(aic_mat):
    """aero/aic.py:115-122."""
    s = aic_mat.shape
    return aic_mat.reshape(s[0] * s[1], s[2] * s[3])


def assemble_aic_sys(aic_mats):
    """aero/aic.py:125-134."""
    return torch.cat([torch.cat([reshape_aic_sys(a) for a in row], dim=1) for row in aic_mats], dim=0)


def add_wake_influence(aic_bs, aic_ws):
    """aero/aic.py:190-205 on separately assembled blocks (compute_aic_solve
    fuses the same lumping into the assembly kernel)."""
    out = []
    for rb, rw in zip(aic_bs, aic_ws):
        row = []
        for ab, aw in zip(rb, rw):
            ab = ab.clone()
            ab[:, :, -1, :] += aw.sum(dim=2)
            row.append(ab)
        out.append(row)
    return out


def compute_aic_solve(cs, ns, zetas_b, zetas_w, kernels_b, kernels_w, batch_size=None, mirror_point=None,
                      mirror_normal=None, out=None):
    """aero/aic.py:137-187: the square system matrix (c_tot, N_b), written
    block by block by the assembly kernel (static wake lumped onto the TE
    row when zetas_w is given)."""
    if zetas_w is not None and kernels_w is None:
        raise ValueError("kernels_w must not be None")
    if isinstance(kernels_b, SegmentKernel) or kernels_b is None:
        kernels_b = [kernels_b] * len(zetas_b)
    mirror = _mirror(mirror_point, mirror_normal)
    cs = [to_device(c) for c in cs]
    ns = [to_device(n) for n in ns]
    zb = [to_device(z) for z in zetas_b]
    zw = [to_device(z) for z in zetas_w] if zetas_w is not None else None
    rows = [c.shape[0] * c.shape[1] for c in cs]
    cols = [(z.shape[0] - 1) * (z.shape[1] - 1) for z in zb]
    if out is None:
        out = torch.empty((sum(rows), sum(cols)), dtype=F64, device=cs[0].device)
    r0 = 0
    for c, n, nr in zip(cs, ns, rows):
        blk = out[r0:r0 + nr]
        col0 = 0
        for j, (z, k) in enumerate(zip(zb, kernels_b)):
            kern = as_segment_kernel(k)
            K.aic_assemble(c.reshape(-1, 3), n.reshape(-1, 3), z, kern.kind, kern.eps, kern.rcut, blk, col0,
                           zeta_w=zw[j] if zw is not None else None, mirror=mirror)
            col0 += cols[j]
        r0 += nr
    return out


def solve_circulation(aic, rhs):
    """The circulation solve jnp.linalg.solve(aic, rhs) of aero/uvlm.py:985:
    FP64 LU with partial pivoting on the device."""
    from . import linalg

    return linalg.solve(aic, rhs)
