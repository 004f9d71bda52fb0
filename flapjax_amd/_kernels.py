"""Tensor-level wrappers of the C-ABI (device tensors in, device tensors out).

Every function validates shapes on the host, allocates its outputs and any
workspace through torch's caching allocator, and enqueues the HIP kernels on
torch's current stream.  Nothing here computes on the host.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L

F64 = torch.float64


def _dev():
    return torch.device("cuda", torch.cuda.current_device())


def _c64(t, shape_tail=None, name="array"):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.dtype != F64:
        raise TypeError(f"{name} must be float64, got {t.dtype}")
    if shape_tail is not None and tuple(t.shape[-len(shape_tail):]) != tuple(shape_tail):
        raise ValueError(f"{name} must end in shape {shape_tail}, got {tuple(t.shape)}")
    return t.contiguous()


def flow_desc(flow):
    """fj_flowfield from a flapjax_amd.flowfields object."""
    return flow.descriptor()


# ---------------------------------------------------------------------------
# K1
# ---------------------------------------------------------------------------
def n_segments(m: int, n: int) -> int:
    return m * (n + 1) + (m + 1) * n


def segments_from_grid(zeta, gamma, kernel_kind, eps, out=None, mirror=None):
    """Filament table of one vertex grid (m+1, n+1, 3) with ring strengths (m, n)."""
    zeta = _c64(zeta, (3,), "zeta")
    if zeta.dim() != 3:
        raise ValueError("zeta must be (m+1, n+1, 3)")
    m, n = zeta.shape[0] - 1, zeta.shape[1] - 1
    ns = n_segments(m, n)
    if gamma is not None:
        gamma = _c64(gamma, None, "gamma")
        if tuple(gamma.shape) != (m, n):
            raise ValueError(f"gamma must be ({m}, {n}), got {tuple(gamma.shape)}")
    if out is None:
        out = torch.empty((ns, L.SEG_STRIDE), dtype=F64, device=zeta.device)
    if ns == 0:
        return out
    mp = mn = None
    keep = []
    if mirror is not None:
        mp, a = L.host3(mirror[0])
        mn, b = L.host3(mirror[1])
        keep += [a, b]
    rc = L.lib().fj_segments_from_grid(L.ptr(zeta), L.ptr(gamma) if gamma is not None else None, m, n, mp, mn,
                                       int(kernel_kind), float(eps), L.ptr(out), L.stream())
    L.check(rc, "fj_segments_from_grid")
    return out


def segments_from_grids(grids, kernel_kind, eps, mirror=None, out=None):
    """One table for several grids: ``grids`` is a list of (zeta, gamma,
    mirrored) with zeta (m+1, n+1, 3) and gamma (m, n) or None; records are
    concatenated in list order (fj_segments_from_grids)."""
    refs = (L.FjGridRef * max(1, len(grids)))()
    total = 0
    for k, (zeta, gamma, mirrored) in enumerate(grids):
        if zeta.dim() != 3 or zeta.shape[-1] != 3 or zeta.dtype != F64 or not zeta.is_contiguous():
            raise ValueError("each grid must be a contiguous float64 (m+1, n+1, 3) tensor")
        m, n = zeta.shape[0] - 1, zeta.shape[1] - 1
        if gamma is not None and (tuple(gamma.shape) != (m, n) or gamma.dtype != F64 or not gamma.is_contiguous()):
            raise ValueError(f"gamma must be a contiguous float64 ({m}, {n}) tensor")
        refs[k].zeta = zeta.data_ptr()
        refs[k].gamma = gamma.data_ptr() if gamma is not None else None
        refs[k].m, refs[k].n, refs[k].mirror = m, n, int(bool(mirrored))
        total += n_segments(m, n)
    if out is None:
        dev = grids[0][0].device if grids else _dev()
        out = torch.empty((total, L.SEG_STRIDE), dtype=F64, device=dev)
    elif out.shape[0] != total or out.dtype != F64 or not out.is_contiguous():
        raise ValueError(f"table output must be a contiguous float64 ({total}, {L.SEG_STRIDE}) tensor")
    mp = mn = None
    keep = []
    if mirror is not None:
        mp, a = L.host3(mirror[0])
        mn, b = L.host3(mirror[1])
        keep += [a, b]
    if total:
        rc = L.lib().fj_segments_from_grids(len(grids), refs, mp, mn, int(kernel_kind), float(eps), L.ptr(out),
                                            L.stream())
        L.check(rc, "fj_segments_from_grids")
    return out


def step_geometry(hg_n, hg_nm1, hg_dot_n, dof, zeta0, zeta_n=None, zeta_nm1=None, zeta_dot_n=None, c=None,
                  nc=None, c_dot=None, mid=None):
    """fj_step_geometry: every output is optional (pass preallocated tensors)."""
    m, n = zeta0.shape[0] - 1, zeta0.shape[1] - 1
    p = lambda t: L.ptr(t) if t is not None else None  # noqa: E731
    rc = L.lib().fj_step_geometry(L.ptr(hg_n), p(hg_nm1), p(hg_dot_n), L.ptr(dof), L.ptr(zeta0), m, n, p(zeta_n),
                                  p(zeta_nm1), p(zeta_dot_n), p(c), p(nc), p(c_dot), p(mid), L.stream())
    L.check(rc, "fj_step_geometry")


def step_forces(zeta, zeta_dot, gamma_b, gamma_b_nm1, gamma_dot_nm1, gamma_w0, v_mid, nc, flow, time, rho, dt,
                relax, f_steady, gamma_dot_n=None, f_unsteady=None):
    """fj_step_forces into the given output tensors."""
    m, n = gamma_b.shape
    p = lambda t: L.ptr(t) if t is not None else None  # noqa: E731
    fd = flow_desc(flow) if flow is not None else None
    rc = L.lib().fj_step_forces(L.ptr(zeta), p(zeta_dot), L.ptr(gamma_b), p(gamma_b_nm1), p(gamma_dot_nm1),
                                p(gamma_w0), L.ptr(v_mid), p(nc), m, n, ctypes.byref(fd) if fd is not None else None,
                                float(time), float(rho), float(dt), float(relax), L.ptr(f_steady), p(gamma_dot_n),
                                p(f_unsteady), L.stream())
    L.check(rc, "fj_step_forces")


def segments_from_endpoints(ends, gamma, kernel_kind, eps):
    """Filament table from end points (S, 2, 3) and circulations (S,)."""
    ends = _c64(ends, (2, 3), "filaments").reshape(-1, 2, 3)
    ns = ends.shape[0]
    if gamma is not None:
        gamma = _c64(gamma, None, "gamma").reshape(-1)
        if gamma.shape[0] != ns:
            raise ValueError("one circulation per filament expected")
    out = torch.empty((ns, L.SEG_STRIDE), dtype=F64, device=ends.device)
    if ns:
        rc = L.lib().fj_segments_from_endpoints(L.ptr(ends), L.ptr(gamma) if gamma is not None else None, ns,
                                                int(kernel_kind), float(eps), L.ptr(out), L.stream())
        L.check(rc, "fj_segments_from_endpoints")
    return out


def _lattice_path(seg, precision, lattices):
    """The lattices of ``seg`` when the FP64 lattice kernel applies to it:
    the explicit list, else the list a table built from whole grids carries
    (``fj_lattices``, set by aic.segment_tables)."""
    if seg is None or int(precision) != L.PREC_FP64:
        return None
    lats = lattices if lattices is not None else getattr(seg, "fj_lattices", None)
    if not lats or len(lats) > MAX_LATTICES:
        return None
    return lats


def pair_workspace_bytes(n_tgt, n_seg, precision=L.PREC_FP64, lattices=None):
    """Workspace bytes of the pair sum that ``pair_sum`` runs for this size,
    precision and lattice list."""
    if int(precision) == L.PREC_FP64 and lattices and len(lattices) <= MAX_LATTICES and n_tgt > 0:
        return lattice_workspace_bytes(n_tgt, lattices)
    return L.lib().fj_pair_sum_workspace_bytes(n_tgt, n_seg, int(precision))


def pair_sum(seg, targets, kernel_kind, eps, rcut, precision=L.PREC_FP64, mode=L.OUT_STORE, out=None, flow=None,
             time=0.0, dt=0.0, workspace=None, lattices=None):
    """Induced velocity of the filament table `seg` at `targets` (T, 3).
    ``workspace`` (float64 device tensor of at least pair_workspace_bytes)
    avoids an allocation per call.  When the table holds whole lattices
    (``lattices``: (seg_off, m, n) per grid, or the ``fj_lattices`` a table
    from aic.segment_tables carries) the FP64 form runs the lattice kernel
    (fj_pair_sum_lattices), which shares the per-vertex terms between
    filaments."""
    lats = _lattice_path(seg, precision, lattices)
    if lats is not None:
        return pair_sum_lattices(seg, lats, targets, kernel_kind, eps, rcut, mode, out, flow, time, dt, workspace)
    targets = _c64(targets, (3,), "targets")
    flat = targets.reshape(-1, 3)
    nt = flat.shape[0]
    ns = 0 if seg is None else seg.shape[0]
    if out is None:
        if mode == L.OUT_ADD:
            raise ValueError("OUT_ADD needs an output tensor")
        out = torch.empty_like(flat)
    if nt == 0:
        return out
    ws_bytes = L.lib().fj_pair_sum_workspace_bytes(nt, ns, int(precision))
    if workspace is not None and workspace.numel() * 8 >= ws_bytes:
        ws = workspace
    else:
        ws = torch.empty(ws_bytes // 8 + 1, dtype=F64, device=flat.device)
    fd = flow_desc(flow) if flow is not None else None
    rc = L.lib().fj_pair_sum(L.ptr(seg) if ns else None, ns, L.ptr(flat), nt, L.ptr(out), int(kernel_kind),
                             float(eps), float(rcut), int(precision), int(mode),
                             ctypes.byref(fd) if fd is not None else None, float(time), float(dt), L.ptr(ws),
                             L.stream())
    L.check(rc, "fj_pair_sum")
    return out


MAX_LATTICES = 16


def lattices_of(grids, base=0):
    """(seg_off, m, n) of each grid of a table built by segments_from_grids
    from ``grids`` [(zeta, gamma, mirrored), ...], starting at row ``base``."""
    out, off = [], base
    for zeta, _, _ in grids:
        m, n = zeta.shape[0] - 1, zeta.shape[1] - 1
        out.append((off, m, n))
        off += n_segments(m, n)
    return out


_REFS: dict = {}
_LAT_WS: dict = {}


def lattice_refs(lattices):
    """ctypes array of fj_lattice_ref from (seg_off, m, n) tuples (kept per
    list: a step builds the same few lists every time, and the callee only
    reads the array)."""
    key = tuple((int(o), int(m), int(n)) for o, m, n in lattices)
    arr = _REFS.get(key)
    if arr is None:
        if len(_REFS) > 1024:
            _REFS.clear()
        arr = (L.FjLatticeRef * max(1, len(key)))()
        for k, (off, m, n) in enumerate(key):
            arr[k].seg_off, arr[k].m, arr[k].n = off, m, n
        _REFS[key] = arr
    return arr


def lattice_workspace_bytes(n_tgt, lattices):
    """fj_pair_sum_lattices_workspace_bytes, kept per (targets, lattices)."""
    key = (int(n_tgt), tuple((int(o), int(m), int(n)) for o, m, n in lattices))
    nb = _LAT_WS.get(key)
    if nb is None:
        if len(_LAT_WS) > 1024:
            _LAT_WS.clear()
        nb = L.lib().fj_pair_sum_lattices_workspace_bytes(key[0], len(key[1]), lattice_refs(key[1]))
        _LAT_WS[key] = nb
    return nb


def pair_sum_lattices(seg, lattices, targets, kernel_kind, eps, rcut, mode=L.OUT_STORE, out=None, flow=None,
                      time=0.0, dt=0.0, workspace=None):
    """FP64 induced velocity at `targets` of the lattices (seg_off, m, n) of
    the filament table `seg` (at most MAX_LATTICES), with the per-vertex part
    of the kernel shared between filaments."""
    if len(lattices) > MAX_LATTICES:
        raise ValueError(f"at most {MAX_LATTICES} lattices per call")
    targets = _c64(targets, (3,), "targets")
    flat = targets.reshape(-1, 3)
    nt = flat.shape[0]
    if out is None:
        if mode == L.OUT_ADD:
            raise ValueError("OUT_ADD needs an output tensor")
        out = torch.empty_like(flat)
    if nt == 0:
        return out
    refs = lattice_refs(lattices)
    nl = len(lattices)
    ws_bytes = lattice_workspace_bytes(nt, lattices)
    if workspace is not None and workspace.numel() * 8 >= ws_bytes:
        ws = workspace
    else:
        ws = torch.empty(ws_bytes // 8 + 1, dtype=F64, device=flat.device)
    fd = flow_desc(flow) if flow is not None else None
    rc = L.lib().fj_pair_sum_lattices(L.ptr(seg) if seg is not None else None, nl, refs, L.ptr(flat), nt,
                                      L.ptr(out), int(kernel_kind), float(eps), float(rcut), int(mode),
                                      ctypes.byref(fd) if fd is not None else None, float(time), float(dt),
                                      L.ptr(ws), L.stream())
    L.check(rc, "fj_pair_sum_lattices")
    return out


def pair_sum_split(seg, targets, n_a, kernel_kind, eps, rcut, out_a, mode_a, out_b, mode_b,
                   precision=L.PREC_FP64, flow=None, time=0.0, dt=0.0, workspace=None, lattices=None):
    """pair_sum over two target ranges in one pair-kernel launch: rows
    [0, n_a) of ``targets`` get ``mode_a`` into ``out_a`` and the rest
    ``mode_b`` into ``out_b``; the lattice kernel when the FP64 table holds
    whole lattices (fj_pair_sum_lattices_split), else the filament-table
    kernels of ``precision`` (fj_pair_sum_split)."""
    lats = _lattice_path(seg, precision, lattices)
    if lats is not None:
        return pair_sum_lattices_split(seg, lats, targets, n_a, kernel_kind, eps, rcut, out_a, mode_a, out_b,
                                       mode_b, flow, time, dt, workspace)
    flat = _c64(targets, (3,), "targets").reshape(-1, 3)
    nt = flat.shape[0]
    _check_split_outputs(nt, n_a, out_a, out_b)
    ns = 0 if seg is None else seg.shape[0]
    ws_bytes = L.lib().fj_pair_sum_workspace_bytes(nt, ns, int(precision))
    if workspace is not None and workspace.numel() * 8 >= ws_bytes:
        ws = workspace
    else:
        ws = torch.empty(ws_bytes // 8 + 1, dtype=F64, device=flat.device)
    fd = flow_desc(flow) if flow is not None else None
    rc = L.lib().fj_pair_sum_split(L.ptr(seg) if ns else None, ns, L.ptr(flat), nt, int(n_a),
                                   L.ptr(out_a) if n_a else None, int(mode_a), L.ptr(out_b) if nt - n_a else None,
                                   int(mode_b), int(kernel_kind), float(eps), float(rcut), int(precision),
                                   ctypes.byref(fd) if fd is not None else None, float(time), float(dt), L.ptr(ws),
                                   L.stream())
    L.check(rc, "fj_pair_sum_split")


def _check_split_outputs(nt, n_a, out_a, out_b):
    if not 0 <= n_a <= nt:
        raise ValueError("n_a must lie in [0, number of targets]")
    for o, k in ((out_a, n_a), (out_b, nt - n_a)):
        if k and (o is None or not o.is_contiguous() or o.numel() != 3 * k or o.dtype != F64):
            raise ValueError("each output must be a contiguous float64 tensor of its range's (T, 3)")


def pair_sum_lattices_split(seg, lattices, targets, n_a, kernel_kind, eps, rcut, out_a, mode_a, out_b, mode_b,
                            flow=None, time=0.0, dt=0.0, workspace=None):
    """pair_sum_lattices over two target ranges in one pair-kernel launch:
    rows [0, n_a) of ``targets`` get ``mode_a`` into ``out_a`` and the rest
    ``mode_b`` into ``out_b`` (fj_pair_sum_lattices_split)."""
    if len(lattices) > MAX_LATTICES:
        raise ValueError(f"at most {MAX_LATTICES} lattices per call")
    flat = _c64(targets, (3,), "targets").reshape(-1, 3)
    nt = flat.shape[0]
    _check_split_outputs(nt, n_a, out_a, out_b)
    refs = lattice_refs(lattices)
    nl = len(lattices)
    ws_bytes = lattice_workspace_bytes(nt, lattices)
    if workspace is not None and workspace.numel() * 8 >= ws_bytes:
        ws = workspace
    else:
        ws = torch.empty(ws_bytes // 8 + 1, dtype=F64, device=flat.device)
    fd = flow_desc(flow) if flow is not None else None
    rc = L.lib().fj_pair_sum_lattices_split(L.ptr(seg) if seg is not None else None, nl, refs, L.ptr(flat), nt,
                                            int(n_a), L.ptr(out_a) if n_a else None, int(mode_a),
                                            L.ptr(out_b) if nt - n_a else None, int(mode_b), int(kernel_kind),
                                            float(eps), float(rcut), ctypes.byref(fd) if fd is not None else None,
                                            float(time), float(dt), L.ptr(ws), L.stream())
    L.check(rc, "fj_pair_sum_lattices_split")


def pair_sum_lattices_noalloc(seg, lattices, targets, kernel_kind, eps, rcut, mode, out, flow=None, time=0.0,
                              dt=0.0):
    """pair_sum_lattices into a caller-given `out`, with no device allocation
    at all (fj_pair_sum_lattices_noalloc: the library's fixed scratch array of
    NOALLOC_SCRATCH_BYTES per GPU).  Calls overlapping on one GPU race."""
    if len(lattices) > MAX_LATTICES:
        raise ValueError(f"at most {MAX_LATTICES} lattices per call")
    flat = _c64(targets, (3,), "targets").reshape(-1, 3)
    if out is None or out.dtype != F64 or not out.is_contiguous() or out.numel() != flat.numel():
        raise ValueError("a contiguous float64 output of the targets' shape is required")
    refs = lattice_refs(lattices)
    fd = flow_desc(flow) if flow is not None else None
    rc = L.lib().fj_pair_sum_lattices_noalloc(L.ptr(seg) if seg is not None else None, len(lattices), refs,
                                              L.ptr(flat), flat.shape[0], L.ptr(out), int(kernel_kind), float(eps),
                                              float(rcut), int(mode), ctypes.byref(fd) if fd is not None else None,
                                              float(time), float(dt), L.stream())
    L.check(rc, "fj_pair_sum_lattices_noalloc")
    return out


# ---------------------------------------------------------------------------
# K2
# ---------------------------------------------------------------------------
def aic_assemble(c, nc, zeta, kernel_kind, eps, rcut, out, col0=0, zeta_w=None, mirror=None, accumulate=False):
    """Write the (T, m*n) block of grid `zeta` into `out` (T, ld) at column col0."""
    c = _c64(c, (3,), "c").reshape(-1, 3)
    nc = _c64(nc, (3,), "normals").reshape(-1, 3)
    zeta = _c64(zeta, (3,), "zeta")
    m, n = zeta.shape[0] - 1, zeta.shape[1] - 1
    m_star = 0
    if zeta_w is not None:
        zeta_w = _c64(zeta_w, (3,), "zeta_w")
        if zeta_w.shape[1] != n + 1:
            raise ValueError("wake grid must have the bound grid's spanwise size")
        m_star = zeta_w.shape[0] - 1
    if out.dtype != F64 or not out.is_contiguous() or out.dim() != 2 or out.shape[0] != c.shape[0]:
        raise ValueError("AIC output must be a contiguous float64 (T, ld) tensor")
    mp = mn = None
    keep = []
    if mirror is not None:
        mp, a = L.host3(mirror[0])
        mn, b = L.host3(mirror[1])
        keep += [a, b]
    rc = L.lib().fj_aic_assemble(L.ptr(c), L.ptr(nc), c.shape[0], L.ptr(zeta), m, n,
                                 L.ptr(zeta_w) if zeta_w is not None else None, m_star, mp, mn, int(kernel_kind),
                                 float(eps), float(rcut), L.ptr(out), out.shape[1], int(col0), int(accumulate),
                                 L.stream())
    L.check(rc, "fj_aic_assemble")
    return out


# ---------------------------------------------------------------------------
# K3
# ---------------------------------------------------------------------------
class Factor(tuple):
    """(lu, piv, aux, info) of fj_lu_factor / fj_lu_factor_f32: the in-place
    factor, 2n int32 pivots + permutation, diagonal-block inverses, status."""

    __slots__ = ()

    def __new__(cls, lu, piv, aux, info):
        return tuple.__new__(cls, (lu, piv, aux, info))

    lu = property(lambda self: self[0])
    piv = property(lambda self: self[1])
    aux = property(lambda self: self[2])
    info = property(lambda self: self[3])


def _factor(A, fn, dtype):
    n = A.shape[0]
    if A.dim() != 2 or A.shape[1] != n:
        raise ValueError("A must be square")
    piv = torch.empty(2 * n, dtype=torch.int32, device=A.device)
    info = torch.zeros(1, dtype=torch.int32, device=A.device)
    aux = torch.empty(L.lib().fj_lu_aux_bytes(n) // 8 + 1, dtype=F64, device=A.device)
    rc = fn(L.ptr(A), n, n, L.ptr(piv), L.ptr(info), L.ptr(aux), L.stream())
    L.check(rc, "LU factorisation")
    return Factor(A, piv, aux, info)


def lu_factor(A, overwrite=False):
    """LU with partial pivoting (FP64) -> Factor."""
    A = _c64(A, None, "A")
    return _factor(A if overwrite else A.clone(), L.lib().fj_lu_factor, F64)


def lu_factor_f32(A32):
    """FP32 LU (matrix cores) -> Factor."""
    if A32.dtype != torch.float32 or not A32.is_cuda:
        raise TypeError("A32 must be a float32 device tensor")
    return _factor(A32.contiguous(), L.lib().fj_lu_factor_f32, torch.float32)


def lu_solve(fac, b, trans=False):
    """Solve with an FP64 factor; b (n,) or (n, nrhs) -> new tensor."""
    LU, piv, aux = fac[0], fac[1], fac[2]
    if LU.dtype != F64:
        raise TypeError("lu_solve takes an FP64 factor (use solve_refined for FP32 factors)")
    n = LU.shape[0]
    b = _c64(b, None, "b")
    x = b.reshape(n, -1).t().contiguous().clone() if b.dim() == 2 else b.reshape(1, n).clone()
    rc = L.lib().fj_lu_solve(L.ptr(LU), n, n, L.ptr(piv), L.ptr(aux), L.ptr(x), x.shape[0], int(bool(trans)),
                             L.stream())
    L.check(rc, "fj_lu_solve")
    return x.t().contiguous() if b.dim() == 2 else x.reshape(n)


def band_fits(n, kl, ku):
    """Whether fj_band_solve takes an n x n system of bandwidths (kl, ku)."""
    return 0 <= kl <= 63 and ku >= 0 and n * (2 * kl + ku + 2) <= L.lib().fj_band_solve_capacity()


def band_solve(A, b, kl, ku, out=None):
    """x = A^-1 b for a banded A stored dense (LAPACK dgbtf2/dgbtrs pivoting;
    same pivot rows as lu_factor / lu_solve)."""
    A = _c64(A, None, "A")
    n = A.shape[0]
    b = _c64(b, (n,), "b")
    x = torch.empty_like(b) if out is None else out
    L.check(L.lib().fj_band_solve(L.ptr(A), n, n, int(kl), int(ku), L.ptr(b), L.ptr(x), None, L.stream()),
            "fj_band_solve")
    return x


def _refine_ws(n, device):
    return torch.empty(L.lib().fj_refine_workspace_bytes(n) // 8 + 1, dtype=F64, device=device)


def solve_refined(A, fac, b, max_iter=8, tol=None, out=None, workspace=None, iters=None):
    """FP64-refined solution of A x = b from the LU factor (FP32 or FP64) of A
    or of a nearby matrix.  Returns (x, iters tensor: corrections made,
    negative if the stop test ||r|| <= tol ||A|| ||x|| never passed).  tol None
    is dsgesv's sqrt(n) eps."""
    A = _c64(A, None, "A")
    LU, piv, aux = fac[0], fac[1], fac[2]
    n = A.shape[0]
    b = _c64(b, None, "b").reshape(n)
    x = torch.empty_like(b) if out is None else out
    iters = torch.zeros(1, dtype=torch.int32, device=A.device) if iters is None else iters
    ws = _refine_ws(n, A.device) if workspace is None else workspace
    fn = L.lib().fj_solve_refined if LU.dtype == torch.float32 else L.lib().fj_solve_refined_f64
    rc = fn(L.ptr(A), n, L.ptr(LU), n, L.ptr(piv), L.ptr(aux), n, L.ptr(b), L.ptr(x), int(max_iter),
            float(tol or 0.0), L.ptr(iters), L.ptr(ws), L.stream())
    L.check(rc, "fj_solve_refined")
    return x, iters


def lu_inverse(fac):
    """Explicit inverse (same dtype as the factor) from a Factor."""
    LU, piv, aux = fac[0], fac[1], fac[2]
    n = LU.shape[0]
    X = torch.empty_like(LU)
    fn = L.lib().fj_lu_inverse_f32 if LU.dtype == torch.float32 else L.lib().fj_lu_inverse
    L.check(fn(L.ptr(LU), n, n, L.ptr(piv), L.ptr(aux), L.ptr(X), n, L.stream()), "fj_lu_inverse")
    return X


def solve_refined_inv(A, Ainv, b, max_iter=4, tol=None, out=None, workspace=None, iters=None):
    """FP64-refined solution of A x = b with M = Ainv (FP64 or FP32), an
    inverse of A or of a nearby matrix; same return and stop test as
    solve_refined."""
    A = _c64(A, None, "A")
    n = A.shape[0]
    if Ainv.shape != (n, n) or not Ainv.is_contiguous():
        raise ValueError("Ainv must be a contiguous (n, n) tensor")
    b = _c64(b, None, "b").reshape(n)
    x = torch.empty_like(b) if out is None else out
    iters = torch.zeros(1, dtype=torch.int32, device=A.device) if iters is None else iters
    ws = _refine_ws(n, A.device) if workspace is None else workspace
    fn = L.lib().fj_solve_refined_inv_f32 if Ainv.dtype == torch.float32 else L.lib().fj_solve_refined_inv
    rc = fn(L.ptr(A), n, L.ptr(Ainv), n, n, L.ptr(b), L.ptr(x), int(max_iter), float(tol or 0.0), L.ptr(iters),
            L.ptr(ws), L.stream())
    L.check(rc, "fj_solve_refined_inv")
    return x, iters


# ---------------------------------------------------------------------------
# K4 - K8
# ---------------------------------------------------------------------------
def grid_from_hg(hg, hg_dot, dof, zeta0, zeta=None, zeta_dot=None):
    hg = _c64(hg, (4, 4), "hg")
    zeta0 = _c64(zeta0, (3,), "zeta0")
    m, n = zeta0.shape[0] - 1, zeta0.shape[1] - 1
    if dof.dtype != torch.int32 or dof.numel() != n + 1:
        raise ValueError("dof mapping must be int32 (n+1,)")
    if zeta is None:
        zeta = torch.empty_like(zeta0)
    if hg_dot is not None:
        hg_dot = _c64(hg_dot, (4, 4), "hg_dot")
        if zeta_dot is None:
            zeta_dot = torch.empty_like(zeta0)
    rc = L.lib().fj_grid_from_hg(L.ptr(hg), L.ptr(hg_dot) if hg_dot is not None else None, L.ptr(dof),
                                 L.ptr(zeta0), m, n, L.ptr(zeta), L.ptr(zeta_dot) if hg_dot is not None else None,
                                 L.stream())
    L.check(rc, "fj_grid_from_hg")
    return zeta, (zeta_dot if hg_dot is not None else None)


def colloc_normals(zeta, zeta_dot=None, want_c=True, want_n=True):
    zeta = _c64(zeta, (3,), "zeta")
    m, n = zeta.shape[0] - 1, zeta.shape[1] - 1
    c = torch.empty((m, n, 3), dtype=F64, device=zeta.device) if want_c else None
    nc = torch.empty((m, n, 3), dtype=F64, device=zeta.device) if want_n else None
    cd = torch.empty((m, n, 3), dtype=F64, device=zeta.device) if zeta_dot is not None else None
    rc = L.lib().fj_colloc_normals(L.ptr(zeta), L.ptr(zeta_dot) if zeta_dot is not None else None, m, n,
                                   L.ptr(c) if c is not None else None, L.ptr(nc) if nc is not None else None,
                                   L.ptr(cd) if cd is not None else None, L.stream())
    L.check(rc, "fj_colloc_normals")
    return c, nc, cd


def filament_midpoints(zeta):
    zeta = _c64(zeta, (3,), "zeta")
    m, n = zeta.shape[0] - 1, zeta.shape[1] - 1
    mid = torch.empty((n_segments(m, n), 3), dtype=F64, device=zeta.device)
    rc = L.lib().fj_filament_midpoints(L.ptr(zeta), m, n, L.ptr(mid), L.stream())
    L.check(rc, "fj_filament_midpoints")
    return mid


def rhs(c, nc, c_dot, v_w, flow, time, out=None):
    c = _c64(c, (3,), "c").reshape(-1, 3)
    nc = _c64(nc, (3,), "nc").reshape(-1, 3)
    nt = c.shape[0]
    if out is None:
        out = torch.empty(nt, dtype=F64, device=c.device)
    fd = flow_desc(flow)
    rc = L.lib().fj_rhs(L.ptr(c), L.ptr(nc), L.ptr(c_dot.reshape(-1, 3).contiguous()) if c_dot is not None else None,
                        L.ptr(v_w.reshape(-1, 3).contiguous()) if v_w is not None else None, nt, ctypes.byref(fd),
                        float(time), L.ptr(out), L.stream())
    L.check(rc, "fj_rhs")
    return out


def wake_shift(gamma_b_nm1, gamma_w_nm1, zeta_b_n, m_star, gamma_w_n, zeta_w_n):
    m, n = gamma_b_nm1.shape
    rc = L.lib().fj_wake_shift(L.ptr(gamma_b_nm1), L.ptr(gamma_w_nm1) if m_star else None, L.ptr(zeta_b_n), m, n,
                               m_star, L.ptr(gamma_w_n) if m_star else None, L.ptr(zeta_w_n), L.stream())
    L.check(rc, "fj_wake_shift")


def wake_convect(zeta_w_nm1, vel, flow, time, dt, zeta_w_n):
    m_star, n = zeta_w_nm1.shape[0] - 1, zeta_w_nm1.shape[1] - 1
    fd = flow_desc(flow) if flow is not None else None
    rc = L.lib().fj_wake_convect(L.ptr(zeta_w_nm1), L.ptr(vel) if vel is not None else None, m_star, n,
                                 ctypes.byref(fd) if fd is not None else None, float(time), float(dt),
                                 L.ptr(zeta_w_n), L.stream())
    L.check(rc, "fj_wake_convect")


def wake_initialise(zeta_b, s, direction, zeta_w=None):
    zeta_b = _c64(zeta_b, (3,), "zeta_b")
    m, n = zeta_b.shape[0] - 1, zeta_b.shape[1] - 1
    s = _c64(s, None, "s")
    m_star = s.shape[0] - 1
    if zeta_w is None:
        zeta_w = torch.empty((m_star + 1, n + 1, 3), dtype=F64, device=zeta_b.device)
    dp, keep = L.host3(direction)
    rc = L.lib().fj_wake_initialise(L.ptr(zeta_b), m, n, L.ptr(s), m_star, dp, L.ptr(zeta_w), L.stream())
    L.check(rc, "fj_wake_initialise")
    return zeta_w


def static_wake_gamma(gamma_b, rows, out=None):
    m, n = gamma_b.shape
    if out is None:
        out = torch.empty((rows, n), dtype=F64, device=gamma_b.device)
    rc = L.lib().fj_static_wake_gamma(L.ptr(gamma_b), m, n, rows, L.ptr(out) if rows * n else None, L.stream())
    L.check(rc, "fj_static_wake_gamma")
    return out


def steady_forces(zeta, zeta_dot, gamma_b, gamma_w0, v_mid, flow, time, rho, out=None):
    zeta = _c64(zeta, (3,), "zeta")
    m, n = zeta.shape[0] - 1, zeta.shape[1] - 1
    if out is None:
        out = torch.empty_like(zeta)
    fd = flow_desc(flow) if flow is not None else None
    rc = L.lib().fj_steady_forces(L.ptr(zeta), L.ptr(zeta_dot) if zeta_dot is not None else None, L.ptr(gamma_b),
                                  L.ptr(gamma_w0) if gamma_w0 is not None else None, L.ptr(v_mid), m, n,
                                  ctypes.byref(fd) if fd is not None else None, float(time), float(rho), L.ptr(out),
                                  L.stream())
    L.check(rc, "fj_steady_forces")
    return out


def gamma_dot_unsteady(gamma_n, gamma_nm1, gamma_dot_nm1, nc, dt, relax, rho, gamma_dot=None, f_uns=None):
    m, n = gamma_n.shape
    if gamma_dot is None:
        gamma_dot = torch.empty_like(gamma_n)
    if f_uns is None and nc is not None:
        f_uns = torch.empty((m + 1, n + 1, 3), dtype=F64, device=gamma_n.device)
    rc = L.lib().fj_gamma_dot_unsteady(L.ptr(gamma_n), L.ptr(gamma_nm1), L.ptr(gamma_dot_nm1),
                                       L.ptr(nc) if nc is not None else None, m, n, float(dt), float(relax),
                                       float(rho), L.ptr(gamma_dot), L.ptr(f_uns) if f_uns is not None else None,
                                       L.stream())
    L.check(rc, "fj_gamma_dot_unsteady")
    return gamma_dot, f_uns


def strip_alpha(zeta, f_steady, flow, time, rho, out=None):
    m, n = zeta.shape[0] - 1, zeta.shape[1] - 1
    if out is None:
        out = torch.empty(n, dtype=F64, device=zeta.device)
    fd = flow_desc(flow)
    rc = L.lib().fj_strip_alpha(L.ptr(zeta), L.ptr(f_steady), m, n, ctypes.byref(fd), float(time), float(rho),
                                L.ptr(out), L.stream())
    L.check(rc, "fj_strip_alpha")
    return out


def flow_eval(flow, x, time):
    x = _c64(x, (3,), "x")
    flat = x.reshape(-1, 3)
    out = torch.empty_like(flat)
    fd = flow_desc(flow)
    rc = L.lib().fj_flow_eval(ctypes.byref(fd), L.ptr(flat), flat.shape[0], float(time), L.ptr(out), L.stream())
    L.check(rc, "fj_flow_eval")
    return out.reshape(x.shape)


def wake_rediscretise(zeta_pre, gamma_pre, s_disc, zeta_w=None, gamma_w=None):
    """Arc-length re-sampling of the convected wake (variable discretisation)."""
    m_star = zeta_pre.shape[0] - 2
    n = zeta_pre.shape[1] - 1
    dev = zeta_pre.device
    if zeta_w is None:
        zeta_w = torch.empty((m_star + 1, n + 1, 3), dtype=F64, device=dev)
    if gamma_w is None:
        gamma_w = torch.empty((m_star, n), dtype=F64, device=dev)
    ws = torch.empty(L.lib().fj_wake_rediscretise_workspace_bytes(m_star, n) // 8 + 1, dtype=F64, device=dev)
    rc = L.lib().fj_wake_rediscretise(L.ptr(zeta_pre), L.ptr(gamma_pre) if m_star > 0 else None, m_star, n,
                                      L.ptr(s_disc), L.ptr(zeta_w), L.ptr(gamma_w) if m_star > 0 else None,
                                      L.ptr(ws), L.stream())
    L.check(rc, "fj_wake_rediscretise")
    return zeta_w, gamma_w
