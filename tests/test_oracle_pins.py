"""Pins of the oracle functions that the GPU parity tests use as their checker
and that no closed form elsewhere fixes (CPU):

  * strip_alpha (aero/utils.py:226-273): the reference's pseudo-infinite wing
    (tests/aero/test_infinite_wing_polar.py:53-71, inboard strips within 2 %
    of the geometric angle), and a synthetic uniform force field whose strip
    lift coefficient is known in closed form (lift direction, chord, strip
    width, dynamic pressure);
  * the Gamma-dot filter (aero/uvlm.py:691-715): a linear ramp gamma = a t
    from rest gives gamma_dot_k = a (1 - (1 - g)^k) exactly;
  * the unsteady force (aero/uvlm.py:1093-1103): the panel normal has
    |n| = 2 * area (shoelace formula), each vertex receives a quarter of a
    panel's rho gamma_dot n, and the vertex forces sum to rho sum gamma_dot n;
  * the variable wake discretisation (aero/utils.py:452-491): with uniform
    delta_w and a trailing edge at rest it reproduces the uniform wake to
    rounding, and under heave it matches the uniform wake within the
    reference's 1e-3 (tests/aero/test_variable_wake_disc.py:12-89);
  * the horseshoe wake (aero/uvlm.py:642-660): equal to an explicit flat wake
    of the same total length (collinear filaments add exactly for the plain
    and cut-off kernels);
  * hg_dot_to_zeta_b_dot (aero/uvlm.py:562-621): a central difference of
    hg_to_zeta_b under the exact pitch-plunge derivative.
"""

import numpy as np
import pytest

from oracle import uvlm_oracle as O


# ---------------------------------------------------------------------------
# strip alpha
# ---------------------------------------------------------------------------
def test_strip_alpha_pseudo_infinite_wing():
    """tests/aero/test_infinite_wing_polar.py:10-71: mirrored 4 x 12 wing of
    span 48 at 3 deg, horseshoe static solve; inboard strips 1 .. n/4 - 1 within
    2 % of the geometric angle of attack (the 2 pi thin-aerofoil slope)."""
    m, n, chord, span, u_mag, alpha = 4, 12, 1.0, 48.0, 20.0, np.deg2rad(3.0)
    ff = O.Constant(np.array((u_mag * np.cos(alpha), 0.0, u_mag * np.sin(alpha))), 1.225, True)
    hg = np.zeros((n + 1, 4, 4))
    hg[:, :3, :3] = np.eye(3)
    hg[:, 1, 3] = np.linspace(0.0, span, n + 1)
    hg[:, 3, 3] = 1.0
    o = O.UVLMOracle([(m, n, 2)], np.arange(n + 1), mirror_point=np.zeros(3), mirror_normal=np.array((0.0, 1.0, 0.0)))
    o.set_design_variables(chord / (u_mag * m), ff, O.make_rectangular_grid(m, n, chord, 0.25), hg)
    a = o.solve_static(horseshoe=True).alpha[0][0]
    interior = a[1:n // 4]
    assert np.max(np.abs(interior - alpha) / alpha) < 0.02
    # the outboard (tip) strips carry less lift: the measure is not a constant
    assert a[-1] < 0.9 * alpha


def test_strip_alpha_closed_form_uniform_force():
    """A flat grid with uneven strips and a uniform vertex force F: every
    strip's force is (m+1) F, the lift direction is v x e_span / |.|, so
    alpha_j = (m+1) F . e_l / (rho |v|^2 / 2 * c * b_j) / (2 pi)."""
    m, n, chord = 3, 5, 1.7
    y = np.array([0.0, 0.4, 1.5, 1.9, 3.2, 4.0])
    zeta = np.zeros((m + 1, n + 1, 3))
    zeta[..., 0] = np.linspace(0.0, chord, m + 1)[:, None] - 0.3
    zeta[..., 1] = y[None, :]
    f = np.broadcast_to(np.array((0.3, -0.7, 2.1)), zeta.shape).copy()
    beta, u, rho = np.deg2rad(7.0), 12.0, 1.1
    v = np.array((u * np.cos(beta), 0.0, u * np.sin(beta)))
    a = O.strip_alpha([zeta], [f], lambda x: np.broadcast_to(v, x.shape), rho)[0]
    e_l = np.array((-np.sin(beta), 0.0, np.cos(beta)))
    b = np.diff(y)
    expect = (m + 1) * (f[0, 0] @ e_l) / (0.5 * rho * u * u * chord * b) / (2.0 * np.pi)
    assert np.allclose(a, expect, rtol=1e-13, atol=0)


# ---------------------------------------------------------------------------
# Gamma-dot filter and unsteady force
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("g", [0.7, 0.3, 1.0])
def test_gamma_dot_filter_linear_ramp(g):
    """gamma = a t from rest: the filter of aero/uvlm.py:691-715 gives
    gamma_dot_k = a (1 - (1 - g)^k) exactly (a swapped g and 1 - g gives
    a (1 - g^k) instead)."""
    rng = np.random.default_rng(11)
    a = rng.normal(size=(3, 4))
    dt = 0.05
    gd = [np.zeros((3, 4))]
    gam = [np.zeros((3, 4))]
    for k in range(1, 25):
        gam_k = [a * (k * dt)]
        gd = O.compute_gamma_dot(gam_k, gam, gd, dt, g)
        gam = gam_k
        assert np.allclose(gd[0], a * (1.0 - (1.0 - g) ** k), rtol=1e-12, atol=1e-14)


def _shoelace_area(quad):
    x, y = quad[:, 0], quad[:, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def test_unsteady_force_normal_area_and_split():
    """|n| = 2 * panel area for planar quads of any shape (n along -z for a
    grid with i along +x and j along +y), each vertex receives a quarter of
    rho gamma_dot n of each panel it bounds, and the sum is rho sum gamma_dot n
    (not the unit-normal force rho gamma_dot S n_hat of docs/theory/uvlm.md)."""
    rng = np.random.default_rng(12)
    m, n = 3, 4
    zeta = np.zeros((m + 1, n + 1, 3))
    zeta[..., 0] = np.arange(m + 1)[:, None] + 0.2 * rng.normal(size=(m + 1, n + 1))
    zeta[..., 1] = np.arange(n + 1)[None, :] * 1.3 + 0.2 * rng.normal(size=(m + 1, n + 1))
    nc = O.compute_surf_nc(zeta)
    for i in range(m):
        for j in range(n):
            quad = np.array([zeta[i, j], zeta[i + 1, j], zeta[i + 1, j + 1], zeta[i, j + 1]])
            assert np.isclose(-nc[i, j, 2], 2.0 * _shoelace_area(quad), rtol=1e-13)
    assert np.allclose(nc[..., :2], 0.0, atol=0)
    gd = rng.normal(size=(m, n))
    rho = 1.225
    f = O.unsteady_forcing([gd], [nc], rho)[0]
    assert np.allclose(f.sum(axis=(0, 1)), rho * (gd[..., None] * nc).sum(axis=(0, 1)), rtol=1e-13)
    # single panel: a quarter at each corner
    one = np.zeros((m, n))
    one[1, 2] = 1.0
    fo = O.unsteady_forcing([one], [nc], rho)[0]
    q = 0.25 * rho * nc[1, 2]
    for (i, j) in ((1, 2), (2, 2), (1, 3), (2, 3)):
        assert np.allclose(fo[i, j], q, rtol=1e-15)
    assert np.count_nonzero(np.abs(fo).sum(-1)) == 4


# ---------------------------------------------------------------------------
# variable wake discretisation
# ---------------------------------------------------------------------------
def _te_pitching_case(delta_w, steps=12, m=4, n=6, m_star=8):
    """A wing pitching about its trailing edge (ea = 1: the TE sits on the beam
    axis, so it never moves), prescribed wake, U = (2, 0, 0)."""
    u = 2.0
    dt = 1.0 / (m * u)
    zeta0 = O.make_rectangular_grid(m, n, 1.0, 1.0)
    hg_t, hgd_t = O.pitch_plunge_motion(n, 3.0, steps, dt, np.deg2rad(4.0), np.deg2rad(3.0), 0.0, 3.0)
    o = O.UVLMOracle([(m, n, m_star)], np.arange(n + 1))
    dw = None if delta_w is None else np.full(m_star, dt * u) if delta_w == "uniform" else delta_w
    o.set_design_variables(dt, O.Constant(np.array((u, 0.0, 0.0)), 1.225), zeta0, hg_t[0], delta_w=dw)
    return o.solve_prescribed_dynamic(o.solve_static(hg=hg_t[0]), hg_t, hgd_t)


def test_variable_wake_uniform_spacing_reproduces_uniform_wake():
    base = _te_pitching_case(None)
    var = _te_pitching_case("uniform")
    assert np.max(np.abs(base.gamma_b[0][1:] - base.gamma_b[0][0])) > 1e-3  # circulation varies
    for key in ("gamma_b", "gamma_w", "zeta_w", "f_steady"):
        a, b = getattr(var, key)[0], getattr(base, key)[0]
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b)), key


def test_variable_wake_heave_matches_uniform():
    """tests/aero/test_variable_wake_disc.py:12-89 (6 x 12 wing, 40 wake rows,
    heave 0.1 (1 - cos(6 pi t)) / 2 for 3 s at U = 10, gamma_dot relaxation 1):
    variable and uniform bound circulations agree to atol 1e-3."""
    u, m, n, m_star, b_ref = 10.0, 6, 12, 40, 5.0
    dt = 1.0 / (m * u)
    n_t = int(np.ceil(3.0 / dt))
    hg = O.rigid_wing_hg(n, b_ref, 0.0)
    t = np.arange(n_t) * dt
    om = 6.0 * np.pi
    hg_t = np.broadcast_to(hg, (n_t, n + 1, 4, 4)).copy()
    hg_t[:, :, 2, 3] += (0.1 * 0.5 * (1.0 - np.cos(om * t)))[:, None]
    hgd_t = np.zeros_like(hg_t)
    hgd_t[:, :, 2, 3] = (0.1 * om * 0.5 * np.sin(om * t))[:, None]
    gam = []
    from oracle import c_oracle

    O.set_pair_sum(c_oracle.oracle_pair_sum)
    try:
        for dw in (np.ones(m_star) * dt * u, None):
            o = O.UVLMOracle([(m, n, m_star)], np.arange(n + 1), gamma_dot_relaxation=1.0)
            o.set_design_variables(dt, O.Constant(np.array((u, 0.0, 0.0)), 1.225),
                                   O.make_rectangular_grid(m, n, 1.0, 0.0), hg, delta_w=dw)
            gam.append(o.solve_prescribed_dynamic(o.solve_static(), hg_t, hgd_t).gamma_b[0])
    finally:
        O.set_pair_sum(None)
    assert np.allclose(gam[0], gam[1], atol=1e-3)
    # and the two wakes do differ (the test is not vacuous)
    assert np.max(np.abs(gam[0] - gam[1])) > 1e-9


# ---------------------------------------------------------------------------
# horseshoe wake
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kernel", ["plain", "cutoff"])
def test_horseshoe_equals_long_flat_wake(kernel):
    """A horseshoe of length HORSESHOE_LENGTH (aero/uvlm.py:642-660) and a
    flat static wake of m* rows with the same total length give the same
    static solution: the wake rings all carry gamma_TE, so their interior
    spanwise filaments cancel and the chordwise ones are collinear pieces of
    the horseshoe legs, whose velocities add exactly."""
    m, n, m_star = 4, 8, 10
    zeta0 = O.make_rectangular_grid(m, n, 1.0, 0.25)
    hg = O.rigid_wing_hg(n, 6.0, np.deg2rad(4.0))
    ff = O.Constant(np.array((1.0, 0.0, 0.0)), 1.225)
    o = O.UVLMOracle([(m, n, m_star)], np.arange(n + 1), kernel=kernel)
    o.set_design_variables(O.HORSESHOE_LENGTH / m_star, ff, zeta0, hg)
    hs = o.solve_static(horseshoe=True)
    flat = o.solve_static(horseshoe=False)
    # the plain kernel is singular on a filament's own line (its midpoint
    # force is 0/0 in the reference too), so only the cut-off kernel has forces
    for key in ("gamma_b",) if kernel == "plain" else ("gamma_b", "f_steady", "alpha"):
        a, b = getattr(hs, key)[0][0], getattr(flat, key)[0][0]
        assert np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(b)), key
    # a short wake is measurably different (the comparison is not vacuous)
    o2 = O.UVLMOracle([(m, n, 2)], np.arange(n + 1), kernel=kernel)
    o2.set_design_variables(0.5, ff, zeta0, hg)
    short = o2.solve_static()
    assert np.max(np.abs(short.gamma_b[0][0] - hs.gamma_b[0][0])) > 1e-3 * np.max(np.abs(hs.gamma_b[0][0]))


# ---------------------------------------------------------------------------
# hg_dot -> zeta_dot
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("t0", [0.0, 2.3, 5.1])
def test_zeta_dot_is_time_derivative_of_zeta(t0):
    n, span, omega, h = 8, 4.0, 0.9, 1e-5
    zeta0 = O.make_rectangular_grid(5, n, 1.0, 0.25)
    args = (np.deg2rad(5.0), np.deg2rad(3.0), 0.05, omega)
    hg_t, hgd_t = O.pitch_plunge_motion(n, span, 3, h, *args, t0=t0 - h)
    o = O.UVLMOracle([(5, n, 2)], np.arange(n + 1))
    o.set_design_variables(0.1, O.Constant(np.array((1.0, 0.0, 0.0)), 1.225), zeta0, hg_t[0])
    fd = (o.hg_to_zeta_b(hg_t[2])[0] - o.hg_to_zeta_b(hg_t[0])[0]) / (2.0 * h)
    zd = o.hg_dot_to_zeta_b_dot(hgd_t[1])[0]
    assert np.max(np.abs(zd - fd)) < 1e-8 * np.max(np.abs(fd))
    # the small-angle skew form is not the derivative at 5 +- 3 deg pitch
    _, skew = O.pitch_plunge_motion(n, span, 3, h, *args, skew_rate=True, t0=t0 - h)
    zs = o.hg_dot_to_zeta_b_dot(skew[1])[0]
    assert np.max(np.abs(zs - fd)) > 1e-3 * np.max(np.abs(fd))
