"""FP64 CPU oracle for the UVLM aerodynamic hot path (test infrastructure only).

This module is the checker that the HIP implementation in ``flapjax_amd`` is
compared against.  It is a plain NumPy FP64 restatement of the reference's
algorithms, written from the reference's source and equations; every function
cites the reference file:line it follows (paths are relative to
``src/flapjax/`` of ben-l-p/flapjax v1.2.1).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker or as the timed
CPU baseline.  The product path (``flapjax_amd``) never calls into it.

Pinning status (how each function is checked, see tests/test_oracle_*.py):
  * kernels (epsilon / cutoff / plain): the reference's own kernel tests
    (tests/algebra/test_kernels.py:9-98), closed forms for a straight segment
    (perpendicular bisector, infinite-line limit), an independent quadrature of
    the Biot-Savart line integral, and the exact algebraic relation between
    the epsilon and the plain kernel.
  * pair sums / compute_v_ind: brute-force per-ring 4-segment sums without
    de-duplication, square-ring centre closed form, superposition, mirror
    equivalence.
  * AIC: A @ gamma == compute_v_ind(c) . n for random gamma; static wake
    lumping == explicit wake with gamma_w = gamma_TE.
  * circulation solve: numpy.linalg.solve + backward-error bound.
  * wake convection: closed form for the prescribed wake, free wake ==
    brute-force pair sums at the vertices, TE-row and shift invariants.
  * forces: zero net force without wake (tests/aero/test_static_no_wake.py:49-55),
    rotation invariance (ibid.), mirrored == full wing (tests/aero/test_mirrored_wing.py).
  * full steps: composition of the above plus equilibrium persistence.
The reference's exact rounding of the self-segment midpoint term, XLA's
reduction order and LAPACK's pivot bits are the only parts an import of the
reference could pin beyond this; they sit well below the 1e-10 tolerance.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Constants, utils/constants.py:13-19
HORSESHOE_LENGTH = 1e3
EPSILON = 1e-7
R_CUTOFF = 1e-4
FOUR_PI = 4.0 * np.pi

KERNELS = ("epsilon", "cutoff", "plain")


# ---------------------------------------------------------------------------
# Segment kernels (per unit circulation), vectorised over leading axes.
# x: (..., 3) target; y0, y1: (..., 3) filament start / end.
# ---------------------------------------------------------------------------
def _dot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def _cross(a, b):
    return np.stack(
        (
            a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
            a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
            a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0],
        ),
        axis=-1,
    )


def bs_epsilon(x, y0, y1, eps=EPSILON):
    """Regularised Biot-Savart kernel, aero/utils.py:606-623.

    K = c (r0 . d) / (4 pi (|c|^2 + eps |r0|^4)), c = r1 x r2,
    d = r1 / sqrt(|r1|^2 + eps^2) - r2 / sqrt(|r2|^2 + eps^2),
    r0 = y1 - y0, r1 = x - y0, r2 = x - y1.  eps enters unsquared in the
    denominator and squared inside the square roots, as in the reference.
    """
    r0 = y1 - y0
    r1 = x - y0
    r2 = x - y1
    inv1 = 1.0 / np.sqrt(_dot(r1, r1) + eps * eps)
    inv2 = 1.0 / np.sqrt(_dot(r2, r2) + eps * eps)
    d = r1 * inv1[..., None] - r2 * inv2[..., None]
    c = _cross(r1, r2)
    r0_sq = _dot(r0, r0)
    denom = _dot(c, c) + eps * r0_sq * r0_sq
    return c * (_dot(r0, d) / (FOUR_PI * denom))[..., None]


def bs_plain(x, y0, y1):
    """Un-regularised Katz-Plotkin segment kernel, aero/utils.py:580-592."""
    r0 = y1 - y0
    r1 = x - y0
    r2 = x - y1
    c = _cross(r1, r2)
    d = r1 / np.sqrt(_dot(r1, r1))[..., None] - r2 / np.sqrt(_dot(r2, r2))[..., None]
    with np.errstate(divide="ignore", invalid="ignore"):
        return c / (_dot(c, c) * FOUR_PI)[..., None] * _dot(r0, d)[..., None]


def bs_cutoff(x, y0, y1, eps=EPSILON, rcut=R_CUTOFF):
    """Cut-off kernel, aero/utils.py:626-648 (make_unit_epsilon :595-603).

    Zero when the distance from x to the (infinite) filament line is <= rcut,
    else the plain form with unit vectors r / sqrt(|r|^2 + eps^2).
    """
    r0 = y1 - y0
    r1 = x - y0
    r2 = x - y1
    sm = _dot(r0, r1) / _dot(r0, r0)
    foot = y0 + sm[..., None] * r0
    rad = np.sqrt(_dot(x - foot, x - foot))
    c = _cross(r1, r2)
    d = r1 / np.sqrt(_dot(r1, r1) + eps * eps)[..., None] - r2 / np.sqrt(
        _dot(r2, r2) + eps * eps
    )[..., None]
    with np.errstate(divide="ignore", invalid="ignore"):
        val = c / _dot(c, c)[..., None] / FOUR_PI * _dot(r0, d)[..., None]
    return np.where((rad > rcut)[..., None], val, 0.0)


def kernel_fn(name):
    return {"epsilon": bs_epsilon, "cutoff": bs_cutoff, "plain": bs_plain}[name]


# ---------------------------------------------------------------------------
# Grid helpers
# ---------------------------------------------------------------------------
def make_rectangular_grid(m, n, chord, ea):
    """aero/utils.py:23-36: x = linspace(0, chord, m+1) - ea*chord, y = z = 0."""
    grid = np.zeros((m + 1, n + 1, 3))
    grid[..., 0] = (np.linspace(0.0, chord, m + 1) - ea * chord)[:, None]
    return grid


def neighbour_average(arr, axes):
    """algebra/array_utils.py:106-135."""
    for ax in (axes,) if isinstance(axes, int) else axes:
        a = np.moveaxis(arr, ax, 0)
        arr = np.moveaxis(0.5 * (a[:-1] + a[1:]), 0, ax)
    return arr


def split_to_vertex(arr, axes):
    """algebra/array_utils.py:331-366: half of each entry to either neighbour."""
    for ax in (axes,) if isinstance(axes, int) else axes:
        a = np.moveaxis(arr, ax, 0)
        out = np.zeros((a.shape[0] + 1,) + a.shape[1:])
        out[:-1] += 0.5 * a
        out[1:] += 0.5 * a
        arr = np.moveaxis(out, 0, ax)
    return arr


def compute_surf_c(zeta):
    """aero/utils.py:87-93: collocation point = 4-vertex average."""
    return neighbour_average(zeta, (-3, -2))


def compute_surf_nc(zeta):
    """aero/utils.py:96-105: diag1 x diag2, |n| = 2 * panel area (not unit)."""
    d1 = zeta[..., 1:, 1:, :] - zeta[..., :-1, :-1, :]
    d2 = zeta[..., 1:, :-1, :] - zeta[..., :-1, 1:, :]
    return _cross(d1, d2)


def mirror_grid(zeta, mirror_point, mirror_normal):
    """aero/utils.py:651-663: reflection across the plane (point, unit normal)."""
    dn = _dot(zeta - mirror_point, np.broadcast_to(mirror_normal, zeta.shape))
    return zeta - 2.0 * dn[..., None] * mirror_normal


def grid_segments(zeta):
    """Unique filaments of a vertex grid, aero/aic.py:277-282 (also :31-36).

    Returns (chordwise y0, y1) of shape (m, n+1, 3) and (spanwise y0, y1) of
    shape (m+1, n, 3).
    """
    return (zeta[:-1], zeta[1:]), (zeta[:, :-1], zeta[:, 1:])


def effective_gamma(gamma):
    """Filament circulations from ring circulations, aero/aic.py:284-287.

    chordwise (m, n+1): g[i, j] - g[i, j-1] (zero padded);
    spanwise (m+1, n): g[i-1, j] - g[i, j] (zero padded).
    """
    gm = np.diff(np.pad(gamma, ((0, 0), (1, 1))), axis=1)
    gn = -np.diff(np.pad(gamma, ((1, 1), (0, 0))), axis=0)
    return gm, gn


# ---------------------------------------------------------------------------
# Pair sums
# ---------------------------------------------------------------------------
def pair_sum(targets, y0, y1, g, kernel="epsilon", chunk=256):
    """sum_s g_s K(x_t; y0_s, y1_s) for flat arrays (the body of v_ind_vmap,
    aero/aic.py:226-245).  targets (T,3), y0/y1 (S,3), g (S,) -> (T,3)."""
    targets = np.asarray(targets, dtype=np.float64).reshape(-1, 3)
    y0 = np.asarray(y0, dtype=np.float64).reshape(-1, 3)
    y1 = np.asarray(y1, dtype=np.float64).reshape(-1, 3)
    g = np.asarray(g, dtype=np.float64).reshape(-1)
    out = np.zeros_like(targets)
    if y0.shape[0] == 0 or targets.shape[0] == 0:
        return out
    kf = kernel_fn(kernel)
    chunk = max(1, int(chunk * 8192 // max(y0.shape[0], 1)) or 1)
    for a in range(0, targets.shape[0], chunk):
        x = targets[a : a + chunk, None, :]
        k = kf(x, y0[None], y1[None])
        out[a : a + chunk] = np.einsum("tsk,s->tk", k, g)
    return out


# optional accelerated pair sum (the C restatement, same formula); tests may
# install it to make larger oracle runs affordable.
_PAIR_SUM = pair_sum


def set_pair_sum(fn):
    global _PAIR_SUM
    _PAIR_SUM = fn if fn is not None else pair_sum


def surface_segments(zeta, gamma):
    """Flat de-duplicated segments of one surface: y0, y1 (S,3), g (S,)."""
    (cy0, cy1), (sy0, sy1) = grid_segments(zeta)
    gm, gn = effective_gamma(gamma)
    y0 = np.concatenate((cy0.reshape(-1, 3), sy0.reshape(-1, 3)))
    y1 = np.concatenate((cy1.reshape(-1, 3), sy1.reshape(-1, 3)))
    g = np.concatenate((gm.ravel(), gn.ravel()))
    return y0, y1, g


def compute_v_ind(cs, zetas, gammas, kernels, mirror_point=None, mirror_normal=None):
    """aero/aic.py:248-310.  cs: array (...,3) or list of arrays."""
    single = isinstance(cs, np.ndarray)
    cs_ = [cs] if single else list(cs)
    if isinstance(kernels, str):
        kernels = [kernels] * len(zetas)
    outs = []
    for c in cs_:
        flat = c.reshape(-1, 3)
        v = np.zeros_like(flat)
        for zeta, gamma, kern in zip(zetas, gammas, kernels):
            if zeta.size == 0 or gamma.size == 0:
                continue
            y0, y1, g = surface_segments(zeta, gamma)
            v += _PAIR_SUM(flat, y0, y1, g, kern)
            if mirror_point is not None and mirror_normal is not None:
                zm = mirror_grid(zeta, mirror_point, mirror_normal)
                y0m, y1m, _ = surface_segments(zm, gamma)
                v -= _PAIR_SUM(flat, y0m, y1m, g, kern)
        outs.append(v.reshape(c.shape))
    return outs[0] if single else outs


def compute_aic_grid(c, n, zeta, kernel="epsilon"):
    """aero/aic.py:11-63: normal-projected ring influence matrix
    (c_m, c_n, m, n).  A = -(M[:, 1:] - M[:, :-1]) + (N[1:] - N[:-1])."""
    c_m, c_n = c.shape[:2]
    m, nn = zeta.shape[0] - 1, zeta.shape[1] - 1
    if not c_m or not c_n or not m or not nn:
        return np.zeros((c_m, c_n, m, nn))
    kf = kernel_fn(kernel)
    (cy0, cy1), (sy0, sy1) = grid_segments(zeta)
    cf = c.reshape(-1, 3)
    nf = n.reshape(-1, 3)
    out = np.empty((cf.shape[0], m, nn))
    step = 64
    for a in range(0, cf.shape[0], step):
        x = cf[a : a + step, None, None, :]
        nv = nf[a : a + step, None, None, :]
        mi = _dot(kf(x, cy0[None], cy1[None]), nv)  # (B, m, n+1)
        ni = _dot(kf(x, sy0[None], sy1[None]), nv)  # (B, m+1, n)
        out[a : a + step] = -np.diff(mi, axis=2) + np.diff(ni, axis=1)
    return out.reshape(c_m, c_n, m, nn)


def compute_aic_sys(zetas, cs, ns, kernels, mirror_point=None, mirror_normal=None):
    """aero/aic.py:66-112."""
    if isinstance(kernels, str):
        kernels = [kernels] * len(zetas)
    mats = []
    for c, n in zip(cs, ns):
        row = []
        for zeta, kern in zip(zetas, kernels):
            a = compute_aic_grid(c, n, zeta, kern)
            if mirror_point is not None and mirror_normal is not None:
                a = a - compute_aic_grid(c, n, mirror_grid(zeta, mirror_point, mirror_normal), kern)
            row.append(a)
        mats.append(row)
    return mats


def add_wake_influence(aic_bs, aic_ws):
    """aero/aic.py:190-205: sum wake rows onto the TE row (steady Kutta)."""
    out = []
    for rb, rw in zip(aic_bs, aic_ws):
        row = []
        for ab, aw in zip(rb, rw):
            ab = ab.copy()
            ab[:, :, -1, :] += aw.sum(axis=2)
            row.append(ab)
        out.append(row)
    return out


def assemble_aic_sys(mats):
    """aero/aic.py:115-134: flatten (c_m*c_n, m*n) blocks and concatenate."""
    rows = []
    for r in mats:
        rows.append(
            np.concatenate([a.reshape(a.shape[0] * a.shape[1], a.shape[2] * a.shape[3]) for a in r], axis=1)
        )
    return np.concatenate(rows, axis=0)


def compute_aic_solve(cs, ns, zetas_b, zetas_w, kernels_b, kernels_w, mirror_point=None, mirror_normal=None):
    """aero/aic.py:137-187."""
    mb = compute_aic_sys(zetas_b, cs, ns, kernels_b, mirror_point, mirror_normal)
    if zetas_w is not None:
        mw = compute_aic_sys(zetas_w, cs, ns, kernels_w, mirror_point, mirror_normal)
        mb = add_wake_influence(mb, mw)
    return assemble_aic_sys(mb)


# ---------------------------------------------------------------------------
# Flow fields, aero/flowfields.py:13-200
# ---------------------------------------------------------------------------
@dataclass
class Constant:
    """aero/flowfields.py:88-97."""

    u_inf: np.ndarray
    rho: float
    relative_motion: bool = True

    def __post_init__(self):
        self.u_inf = np.asarray(self.u_inf, dtype=np.float64)
        self.u_inf_mag = float(np.linalg.norm(self.u_inf))
        self.u_inf_dir = self.u_inf / self.u_inf_mag

    def __call__(self, x, t):
        base = self.u_inf if self.relative_motion else np.zeros(3)
        return np.broadcast_to(base, np.shape(x)).copy()


@dataclass
class OneMinusCosine(Constant):
    """aero/flowfields.py:101-181: 0.5*A*(1-cos(pi*s/L)) for 0 < s < 2L."""

    gust_length: float = 1.0
    gust_amplitude: float = 0.0
    gust_travel_direction: np.ndarray | None = None
    gust_amplitude_direction: np.ndarray | None = None
    gust_x0: np.ndarray | None = None

    def __post_init__(self):
        super().__post_init__()
        td = self.u_inf_dir if self.gust_travel_direction is None else np.asarray(self.gust_travel_direction, float)
        self.gust_travel_direction = td / np.linalg.norm(td)
        ad = np.array((0.0, 0.0, 1.0)) if self.gust_amplitude_direction is None else np.asarray(
            self.gust_amplitude_direction, float
        )
        self.gust_amplitude_direction = ad / np.linalg.norm(ad)
        self.gust_x0 = np.zeros(3) if self.gust_x0 is None else np.asarray(self.gust_x0, float)

    def __call__(self, x, t):
        x = np.asarray(x, dtype=np.float64)
        rel = x - self.gust_x0
        if self.relative_motion:
            rel = rel - self.u_inf * t
        s = _dot(rel, np.broadcast_to(self.gust_travel_direction, rel.shape))
        amp = self.gust_amplitude * 0.5 * (1.0 - np.cos(np.pi * s / self.gust_length))
        inside = (s > 0) & (s < 2.0 * self.gust_length)
        u = np.where(inside[..., None], self.gust_amplitude_direction * amp[..., None], 0.0)
        if self.relative_motion:
            u = u + self.u_inf
        return u


# ---------------------------------------------------------------------------
# Forces, aero/utils.py:126-273
# ---------------------------------------------------------------------------
def calculate_steady_forcing(zeta_b, zeta_dot_b, gamma_b, gamma_w, rho, v_func):
    """aero/utils.py:126-223: rho * Gamma_fil * (v_rel x r) at filament
    midpoints, split half/half to the two end vertices."""
    out = []
    for zb, zd, gb, gw in zip(zeta_b, zeta_dot_b, gamma_b, gamma_w):
        mp_c = neighbour_average(zb, 0)
        mp_s = neighbour_average(zb, 1)
        v_c = v_func(mp_c) - neighbour_average(zd, 0)
        v_s = v_func(mp_s) - neighbour_average(zd, 1)
        g_c = np.zeros(v_c.shape[:-1])
        g_c[:, :-1] = gb
        g_c[:, 1:] -= gb
        g_s = np.zeros(v_s.shape[:-1])
        g_s[:-1, :] = -gb
        g_s[1:, :] += gb
        if gw.shape[0] > 0:
            g_s[-1, :] -= gw[0, :]
        r_c = zb[1:] - zb[:-1]
        r_s = zb[:, 1:] - zb[:, :-1]
        f_c = rho * g_c[..., None] * _cross(v_c, r_c)
        f_s = rho * g_s[..., None] * _cross(v_s, r_s)
        out.append(split_to_vertex(f_c, 0) + split_to_vertex(f_s, 1))
    return out


def strip_alpha(zeta_b, f_steady, v_func, rho):
    """aero/utils.py:226-273: per-strip lift coefficient / (2 pi)."""
    out = []
    for zs, fs in zip(zeta_b, f_steady):
        le, te = zs[0], zs[-1]
        le_j = 0.5 * (le[:-1] + le[1:])
        te_j = 0.5 * (te[:-1] + te[1:])
        c_len = np.linalg.norm(te_j - le_j, axis=-1)
        span = 0.5 * ((le[1:] + te[1:]) - (le[:-1] + te[:-1]))
        b_len = np.linalg.norm(span, axis=-1)
        e_s = span / b_len[:, None]
        v_ref = v_func(0.5 * (le_j + te_j))
        v2 = _dot(v_ref, v_ref)
        e_l = _cross(v_ref, e_s)
        e_l = e_l / np.linalg.norm(e_l, axis=-1, keepdims=True)
        f_strip = neighbour_average(fs, 1).sum(axis=0)
        q = 0.5 * rho * v2
        out.append(_dot(f_strip, e_l) / (q * c_len * b_len) / (2.0 * np.pi))
    return out


def compute_gamma_dot(gamma_b_n, gamma_b_nm1, gamma_b_dot_nm1, dt, relaxation):
    """aero/uvlm.py:691-715: filtered time derivative of the ring strengths,
    g (gamma_n - gamma_nm1) / dt + (1 - g) gamma_dot_nm1, per surface."""
    return [(gn - gp) / dt * relaxation + gdp * (1.0 - relaxation)
            for gn, gp, gdp in zip(gamma_b_n, gamma_b_nm1, gamma_b_dot_nm1)]


def unsteady_forcing(gamma_b_dot, nc, rho):
    """aero/uvlm.py:1093-1103: split_to_vertex(rho * gamma_dot * n) over both
    grid axes, with the panel normal n of compute_surf_nc (|n| = 2 * area)."""
    return [split_to_vertex(rho * gd[..., None] * n, (0, 1)) for gd, n in zip(gamma_b_dot, nc)]


# ---------------------------------------------------------------------------
# Wake convection, aero/utils.py:400-577
# ---------------------------------------------------------------------------
def propagate_surf_wake(gamma_b_nm1, gamma_w_nm1, zeta_b_n, zeta_w_nm1, delta_w, v_func, dt):
    """aero/utils.py:400-500 (uniform and variable discretisation)."""
    zeta_te = zeta_b_n[-1]
    gamma_te = gamma_b_nm1[-1]
    if delta_w is not None:
        z_base, g_base = zeta_w_nm1, gamma_w_nm1
    else:
        z_base, g_base = zeta_w_nm1[:-1], gamma_w_nm1[:-1]
    gamma_w_n = np.concatenate((gamma_te[None], g_base), axis=0)
    v = v_func(z_base)
    zeta_w_n = np.concatenate((zeta_te[None], z_base + dt * v), axis=0)
    if delta_w is not None:
        seg = np.linalg.norm(zeta_w_n[1:] - zeta_w_n[:-1], axis=-1)
        s_z = np.concatenate((np.zeros((1, zeta_te.shape[0])), np.cumsum(seg, axis=0)), axis=0)
        s_g = neighbour_average(s_z, (0, 1))
        s_zd = np.concatenate((np.zeros(1), np.cumsum(delta_w)))
        s_gd = neighbour_average(s_zd, (0,))
        zn = np.empty((s_zd.shape[0], zeta_w_n.shape[1], 3))
        for j in range(zeta_w_n.shape[1]):
            for k in range(3):
                zn[:, j, k] = np.interp(s_zd, s_z[:, j], zeta_w_n[:, j, k])
        gn = np.empty((s_gd.shape[0], gamma_w_n.shape[1]))
        for j in range(gamma_w_n.shape[1]):
            gn[:, j] = np.interp(s_gd, s_g[:, j], gamma_w_n[:, j])
        zeta_w_n, gamma_w_n = zn, gn
    if gamma_w_nm1.size == 0:
        gamma_w_n = np.zeros_like(gamma_w_nm1)
    return zeta_w_n, gamma_w_n


def propagate_wake(gamma_b_nm1, gamma_w_nm1, zeta_b_n, zeta_w_nm1, delta_w, v_func, dt):
    """aero/utils.py:531-577."""
    zs, gs = [], []
    for i in range(len(gamma_b_nm1)):
        z, g = propagate_surf_wake(gamma_b_nm1[i], gamma_w_nm1[i], zeta_b_n[i], zeta_w_nm1[i], delta_w[i], v_func, dt)
        zs.append(z)
        gs.append(g)
    return zs, gs


# ---------------------------------------------------------------------------
# UVLM orchestration, aero/uvlm.py
# ---------------------------------------------------------------------------
def vect_product(hg, x):
    """algebra/se3.py:209-216 broadcast over the spanwise node axis:
    hg (n+1, 4, 4), x (m+1, n+1, 3) -> R_j x_ij + t_j."""
    return np.einsum("jab,ijb->ija", hg[:, :3, :3], x) + hg[None, :, :3, 3]


SMALL_ANG_THRESH = 1e-10  # utils/constants.py:7


def exp_so3(w):
    """algebra/so3.py:168-189: Rodrigues above the switch on |w|^2, exp_sum of
    order 2 below it.  Complex arguments are allowed (w.w without conjugate),
    so complex-step derivatives pass through."""
    s = np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])
    n2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2]
    if np.real(n2) > SMALL_ANG_THRESH:
        a = np.sqrt(n2)
        return np.eye(3) + np.sin(a) / a * s + (1.0 - np.cos(a)) / (a * a) * (s @ s)
    return np.eye(3) + s + (s @ s) / 2.0


def add_control_surface(grid, angle, m_slice, n_slice, hinge_axis=(0.0, 1.0, 0.0)):
    """aero/utils.py:42-84: the vertices grid[m_slice, n_slice] rotated by
    angle about hinge_axis through each column's hinge point (the column's
    first vertex of the slice)."""
    grid = np.asarray(grid)
    ms = np.arange(grid.shape[0])[m_slice] if isinstance(m_slice, slice) else np.atleast_1d(m_slice)
    ns = np.arange(grid.shape[1])[n_slice] if isinstance(n_slice, slice) else np.atleast_1d(n_slice)
    rmat = exp_so3(np.asarray(hinge_axis, float) * angle)
    out = grid.astype(np.result_type(grid, angle, float)).copy()
    for j in ns:
        hinge = grid[ms[0], j]
        out[ms, j] = (grid[ms, j] - hinge) @ rmat.T + hinge
    return out


def complex_step_jvp(fn, angles, velocities, h=1e-30):
    """Forward derivative of fn(**angles) (a list of arrays) along
    velocities, by one complex-step evaluation: Im f(a + i h v) / h, exact to
    rounding for a function analytic in the angles."""
    pert = {k: angles[k] + 1j * h * velocities[k] for k in angles}
    return [np.imag(z) / h for z in fn(**pert)]


@dataclass
class OracleCase:
    """Snapshot history, the fields of AeroCase (aero/data_structures.py:43-615)."""

    t: np.ndarray
    zeta_b: list
    zeta_b_dot: list
    zeta_w: list
    gamma_b: list
    gamma_b_dot: list
    gamma_w: list
    f_steady: list
    f_unsteady: list
    alpha: list
    c: list
    nc: list
    aic: list = field(default_factory=list)


class UVLMOracle:
    """Restatement of UVLM (aero/uvlm.py:166-1406) for the settings the
    hot path covers: no polars; control surfaces through grid_func, whose
    forward derivative is taken by complex step (the grid function must then
    accept complex angles, as add_control_surface does)."""

    def __init__(self, grid_shapes, dof_mapping, kernel="epsilon", free_wake=False,
                 gamma_dot_relaxation=0.7, mirror_point=None, mirror_normal=None,
                 variable_wake_disc=False, grid_func=None):
        self.grid_disc = [tuple(int(v) for v in g) for g in grid_shapes]
        self.n_surf = len(self.grid_disc)
        dm = dof_mapping if isinstance(dof_mapping, (list, tuple)) else [dof_mapping]
        self.dof_mapping = [np.asarray(d, dtype=np.int64) for d in dm]
        self.kernel = kernel
        self.free_wake = free_wake
        self.gamma_dot_relaxation = gamma_dot_relaxation
        if (mirror_point is None) != (mirror_normal is None):
            raise ValueError("Both mirror_point and mirror_normal must be provided")
        self.mirror_point = None if mirror_point is None else np.asarray(mirror_point, float)
        self.mirror_normal = None if mirror_normal is None else np.asarray(mirror_normal, float) / np.linalg.norm(mirror_normal)
        self.keep_aic = False
        self.grid_func = grid_func
        self.cs_ang0 = {}

    # aero/uvlm.py:387-472
    def set_design_variables(self, dt, flowfield, zeta_b0, hg0, delta_w=None, reference_cs_angles=None):
        self.dt = float(dt)
        self.flowfield = flowfield
        self.zeta_b0 = [np.asarray(z, float) for z in (zeta_b0 if isinstance(zeta_b0, (list, tuple)) else [zeta_b0])]
        self.hg_ref = np.asarray(hg0, float)
        self.delta_w = list(delta_w) if isinstance(delta_w, (list, tuple)) else [delta_w] * self.n_surf
        if reference_cs_angles is not None:
            self.cs_ang0 = {k: float(v) for k, v in reference_cs_angles.items()}
        self.zeta_b_ref = self.hg_to_zeta_b(self.hg_ref, self.cs_ang0)
        self.zeta_w_ref = self.initialise_wake(self.zeta_b_ref)

    def _cs_grids(self, cs_ang):
        if self.grid_func is None:
            return self.zeta_b0
        return list(self.grid_func(self.zeta_b0, **(cs_ang or {})))

    # aero/uvlm.py:537-560
    def hg_to_zeta_b(self, hg, cs_ang=None):
        z0 = self._cs_grids(self.cs_ang0 if cs_ang is None else cs_ang)
        return [vect_product(hg[self.dof_mapping[i]], z0[i]) for i in range(self.n_surf)]

    # aero/uvlm.py:562-621: hg_dot [zeta0_cs; 1] + R d(zeta0_cs)/dt, the last
    # term the forward derivative of grid_func along the control velocities
    def hg_dot_to_zeta_b_dot(self, hg_dot, hg=None, cs_ang=None, cs_vel=None):
        cs_ang = self.cs_ang0 if cs_ang is None else cs_ang
        cs_vel = {} if cs_vel is None else cs_vel
        for key in set(cs_ang) | set(cs_vel):
            if key not in cs_vel or key not in cs_ang:
                raise ValueError(f"Missing pair of control angles and velocities for control surface key '{key}'")
        z0 = self._cs_grids(cs_ang)
        out = [vect_product(hg_dot[self.dof_mapping[i]], z0[i]) for i in range(self.n_surf)]
        if self.grid_func is not None and cs_ang:
            dz = complex_step_jvp(lambda **a: self.grid_func(self.zeta_b0, **a), cs_ang, cs_vel)
            for i in range(self.n_surf):
                rot = hg[self.dof_mapping[i], :3, :3]
                out[i] = out[i] + np.einsum("njk,mnk->mnj", rot, dz[i])
        return out

    # aero/uvlm.py:662-689
    def initialise_wake(self, zeta_b):
        out = []
        for i in range(self.n_surf):
            te = zeta_b[i][-1]
            m_star = self.grid_disc[i][2]
            dw = self.delta_w[i]
            if dw is None:
                dw = np.ones(m_star) * self.dt * self.flowfield.u_inf_mag
            s = np.concatenate((np.zeros(1), np.cumsum(dw)))
            out.append(te[None] + np.outer(s, self.flowfield.u_inf_dir)[:, None, :])
        return out

    def _horseshoe_wake(self, zeta_b, i):
        """aero/uvlm.py:642-660."""
        te = zeta_b[-1]
        if self.grid_disc[i][2] == 0:
            return te[None]
        return np.stack((te, te + self.flowfield.u_inf_dir * HORSESHOE_LENGTH), axis=0)

    def _v_ind(self, x, zetas, gammas):
        return compute_v_ind(x, zetas, gammas, self.kernel, self.mirror_point, self.mirror_normal)

    # aero/uvlm.py:824-1117
    def base_solve_from_grid(self, q_nm1, t_n, zeta_b_n, zeta_b_nm1, zeta_b_dot_n, static, horseshoe):
        if not static:
            horseshoe = False
        c_n = [compute_surf_c(z) for z in zeta_b_n]
        nc_n = [compute_surf_nc(z) for z in zeta_b_n]
        c_dot_n = None if zeta_b_dot_n is None else [neighbour_average(zd, (0, 1)) for zd in zeta_b_dot_n]
        ff = self.flowfield
        if static:
            if horseshoe:
                zeta_w_n = [self._horseshoe_wake(zeta_b_n[i], i) for i in range(self.n_surf)]
            else:
                zeta_w_n = self.initialise_wake(zeta_b_n)
            gamma_w_n = None
        else:
            zeta_full = list(zeta_b_nm1) + list(q_nm1["zeta_w"])
            gamma_full = list(q_nm1["gamma_b"]) + list(q_nm1["gamma_w"])

            def v_wake_prop(x):
                v = ff(x, t_n)
                if self.free_wake:
                    v = v + self._v_ind(x, zeta_full, gamma_full)
                return v

            zeta_w_n, gamma_w_n = propagate_wake(
                q_nm1["gamma_b"], q_nm1["gamma_w"], zeta_b_n, q_nm1["zeta_w"], self.delta_w, v_wake_prop, self.dt
            )
        aic = compute_aic_solve(
            c_n, nc_n, zeta_b_n, zeta_w_n if static else None, self.kernel, self.kernel,
            self.mirror_point, self.mirror_normal,
        )
        v_bc = [ff(c, t_n) for c in c_n]
        if not static:
            v_bc = [v - cd for v, cd in zip(v_bc, c_dot_n)]
            vw = self._v_ind(c_n, zeta_w_n, gamma_w_n)
            v_bc = [v + w for v, w in zip(v_bc, vw)]
        rhs = np.concatenate([_dot(v, n).ravel() for v, n in zip(v_bc, nc_n)])
        gvec = np.linalg.solve(aic, -rhs)
        gamma_b_n = []
        off = 0
        for (m, n, _ms) in self.grid_disc:
            gamma_b_n.append(gvec[off : off + m * n].reshape(m, n))
            off += m * n
        if static:
            gamma_w_n = [
                np.broadcast_to(gb[[-1]], (1 if (horseshoe and gd[2] != 0) else gd[2], gd[1])).copy()
                for gb, gd in zip(gamma_b_n, self.grid_disc)
            ]
        zeta_dot_f = zeta_b_dot_n if zeta_b_dot_n is not None else [np.zeros_like(z) for z in zeta_b_n]
        zeta_all = list(zeta_b_n) + list(zeta_w_n)
        gamma_all = list(gamma_b_n) + list(gamma_w_n)

        def v_total(x):
            return ff(x, t_n) + self._v_ind(x, zeta_all, gamma_all)

        f_steady = calculate_steady_forcing(zeta_b_n, zeta_dot_f, gamma_b_n, gamma_w_n, ff.rho, v_total)
        alpha = strip_alpha(zeta_b_n, f_steady, lambda x: ff(x, t_n), ff.rho)
        if static:
            gamma_b_dot_n, f_unsteady = None, None
        else:
            gamma_b_dot_n = compute_gamma_dot(gamma_b_n, q_nm1["gamma_b"], q_nm1["gamma_b_dot"], self.dt,
                                              self.gamma_dot_relaxation)
            f_unsteady = unsteady_forcing(gamma_b_dot_n, nc_n, ff.rho)
        return dict(c=c_n, nc=nc_n, gamma_b=gamma_b_n, gamma_w=gamma_w_n, gamma_b_dot=gamma_b_dot_n,
                    zeta_b=zeta_b_n, zeta_w=zeta_w_n, zeta_b_dot=zeta_b_dot_n, f_steady=f_steady,
                    f_unsteady=f_unsteady, alpha=alpha, aic=aic)

    def _empty_case(self, n_tstep):
        gd = self.grid_disc
        z = lambda *s: [np.zeros((n_tstep,) + tuple(x)) for x in s]  # noqa: E731
        return OracleCase(
            t=np.zeros(n_tstep),
            zeta_b=z(*[(m + 1, n + 1, 3) for m, n, _ in gd]),
            zeta_b_dot=z(*[(m + 1, n + 1, 3) for m, n, _ in gd]),
            zeta_w=z(*[(ms + 1, n + 1, 3) for m, n, ms in gd]),
            gamma_b=z(*[(m, n) for m, n, _ in gd]),
            gamma_b_dot=z(*[(m, n) for m, n, _ in gd]),
            gamma_w=z(*[(ms, n) for m, n, ms in gd]),
            f_steady=z(*[(m + 1, n + 1, 3) for m, n, _ in gd]),
            f_unsteady=z(*[(m + 1, n + 1, 3) for m, n, _ in gd]),
            alpha=z(*[(n,) for m, n, _ in gd]),
            c=z(*[(m, n, 3) for m, n, _ in gd]),
            nc=z(*[(m, n, 3) for m, n, _ in gd]),
        )

    def _store(self, case, i, out, static):
        for key in ("c", "nc", "gamma_b", "gamma_w", "zeta_b", "f_steady", "alpha", "zeta_w"):
            for s in range(self.n_surf):
                getattr(case, key)[s][i] = out[key][s]
        if not static:
            for key in ("gamma_b_dot", "zeta_b_dot", "f_unsteady"):
                for s in range(self.n_surf):
                    getattr(case, key)[s][i] = out[key][s]
        if self.keep_aic:
            case.aic.append(out["aic"])

    # aero/uvlm.py:1286-1333
    def solve_static(self, hg=None, t=0.0, horseshoe=False, cs_ang=None):
        case = self._empty_case(1)
        case.t[0] = t
        zeta_b = self.hg_to_zeta_b(self.hg_ref if hg is None else np.asarray(hg, float), cs_ang)
        out = self.base_solve_from_grid(None, t, zeta_b, None, None, True, horseshoe)
        if horseshoe:
            # the stored wake is the regular one (aero/uvlm.py:1329-1331)
            out["zeta_w"] = self.initialise_wake(zeta_b)
        self._store(case, 0, out, True)
        return case

    # aero/uvlm.py:1335-1406 with case_solve :1119-1215
    def solve_prescribed_dynamic(self, init_case, hg_t, hg_dot_t, callback=None, cs_ang_t=None, cs_vel_t=None):
        hg_t = np.asarray(hg_t, float)
        hg_dot_t = np.asarray(hg_dot_t, float)
        n_tstep = hg_t.shape[0]
        case = self._empty_case(n_tstep)
        for key in ("zeta_b", "zeta_b_dot", "zeta_w", "gamma_b", "gamma_b_dot", "gamma_w", "f_steady",
                    "f_unsteady", "alpha", "c", "nc"):
            for s in range(self.n_surf):
                getattr(case, key)[s][0] = getattr(init_case, key)[s][0]
        case.t[0] = init_case.t[0]
        for i in range(1, n_tstep):
            case.t[i] = case.t[i - 1] + self.dt
            q = dict(
                gamma_b=[g[i - 1] for g in case.gamma_b],
                gamma_w=[g[i - 1] for g in case.gamma_w],
                gamma_b_dot=[g[i - 1] for g in case.gamma_b_dot],
                zeta_w=[z[i - 1] for z in case.zeta_w],
            )
            if cs_ang_t is None and cs_vel_t is None:
                zb = self.hg_to_zeta_b(hg_t[i])
                zbd = self.hg_dot_to_zeta_b_dot(hg_dot_t[i], hg_t[i])
                zbm1 = self.hg_to_zeta_b(hg_t[i - 1])
            else:
                a_n = {k: float(v[i]) for k, v in (cs_ang_t or {}).items()}
                a_nm1 = {k: float(v[i - 1]) for k, v in (cs_ang_t or {}).items()}
                v_n = {k: float(v[i]) for k, v in (cs_vel_t or {}).items()}
                zb = self.hg_to_zeta_b(hg_t[i], a_n)
                zbd = self.hg_dot_to_zeta_b_dot(hg_dot_t[i], hg_t[i], a_n, v_n)
                zbm1 = self.hg_to_zeta_b(hg_t[i - 1], a_nm1)
            out = self.base_solve_from_grid(q, case.t[i], zb, zbm1, zbd, False, False)
            self._store(case, i, out, False)
            if callback is not None:
                callback(i, out)
        return case


# ---------------------------------------------------------------------------
# Workload builders (seeded geometry used by tests and bench)
# ---------------------------------------------------------------------------
def rigid_wing_hg(n, span, alpha_rad=0.0, z=0.0):
    """Beam poses for a straight rigid wing along +y: (n+1, 4, 4)."""
    ca, sa = np.cos(alpha_rad), np.sin(alpha_rad)
    rot = np.array([[ca, 0.0, sa], [0.0, 1.0, 0.0], [-sa, 0.0, ca]])
    hg = np.zeros((n + 1, 4, 4))
    hg[:, :3, :3] = rot
    hg[:, 0, 3] = 0.0
    hg[:, 1, 3] = np.linspace(0.0, span, n + 1)
    hg[:, 2, 3] = z
    hg[:, 3, 3] = 1.0
    return hg


def pitch_plunge_motion(n, span, n_tstep, dt, alpha0_rad, pitch_amp_rad, plunge_amp, omega, skew_rate=False,
                        t0=0.0):
    """Prescribed pitch-plunge history about the beam axis (y):
    theta(t) = alpha0 + A_p sin(w t), h(t) = A_h sin(w t).

    hg_dot is the exact time derivative of hg (aero/uvlm.py:562-621 maps it
    onto the local grid as hg_dot @ [zeta0; 1], the global grid velocity):
    dR/dt = theta_dot dR/dtheta for the rotation block and h_dot in [2, 3].
    skew_rate=True gives the small-angle form of the reference's
    prescribed-motion tests instead (tests/aero/test_linear_planar_wing.py:168-170):
    a skew rate block [0,2] = theta_dot, [2,0] = -theta_dot, which equals the
    exact derivative only at theta = 0.  Times are t0 + i dt.
    """
    t = t0 + np.arange(n_tstep) * dt
    th = alpha0_rad + pitch_amp_rad * np.sin(omega * t)
    thd = pitch_amp_rad * omega * np.cos(omega * t)
    h = plunge_amp * np.sin(omega * t)
    hd = plunge_amp * omega * np.cos(omega * t)
    hg_t = np.stack([rigid_wing_hg(n, span, th[i], h[i]) for i in range(n_tstep)])
    hg_dot_t = np.zeros_like(hg_t)
    if skew_rate:
        hg_dot_t[:, :, 0, 2] = thd[:, None]
        hg_dot_t[:, :, 2, 0] = -thd[:, None]
    else:
        # R = [[c, 0, s], [0, 1, 0], [-s, 0, c]] -> dR/dtheta = [[-s, 0, c], [0, 0, 0], [-c, 0, -s]]
        s_, c_ = np.sin(th), np.cos(th)
        hg_dot_t[:, :, 0, 0] = (-s_ * thd)[:, None]
        hg_dot_t[:, :, 0, 2] = (c_ * thd)[:, None]
        hg_dot_t[:, :, 2, 0] = (-c_ * thd)[:, None]
        hg_dot_t[:, :, 2, 2] = (-s_ * thd)[:, None]
    hg_dot_t[:, :, 2, 3] = hd[:, None]
    return hg_t, hg_dot_t
