"""Pins of the oracle's AIC, solve, wake convection, forces and full steps (CPU).

  * A @ gamma == v_ind(c) . n, and the static wake lumping == an explicit wake
    with gamma_w = gamma_TE (aero/aic.py:190-205);
  * the static solution satisfies no penetration at every collocation point;
  * circulation solve vs LAPACK (scipy) with a backward-error bound;
  * prescribed wake: closed-form shift + U dt, TE row, circulation shift;
  * free wake: convection velocity == brute-force pair sum at the vertices;
  * the reference's aero tests: rotation invariance and zero net force without
    wake (tests/aero/test_static_no_wake.py), mirrored == full wing
    (tests/aero/test_mirrored_wing.py);
  * equilibrium persistence: a steady wing with no motion keeps its
    circulation (tests/coupled/dynamic/test_dynamic_equilibrium.py:59-69).
"""

import numpy as np
import pytest
import scipy.linalg

from oracle import uvlm_oracle as O


def _wing(m, n, m_star, alpha_deg=5.0, span=10.0, dt=None):
    zeta0 = O.make_rectangular_grid(m, n, 1.0, 0.25)
    hg = O.rigid_wing_hg(n, span, np.deg2rad(alpha_deg))
    zb = O.vect_product(hg, zeta0)
    dt = 1.0 / m if dt is None else dt
    s = np.concatenate(([0.0], np.cumsum(np.ones(m_star) * dt)))
    zw = zb[-1][None] + np.outer(s, [1.0, 0.0, 0.0])[:, None, :]
    return zeta0, hg, zb, zw


def test_aic_matches_pair_sum():
    _, _, zb, zw = _wing(4, 8, 6)
    c, nc = O.compute_surf_c(zb), O.compute_surf_nc(zb)
    A = O.compute_aic_solve([c], [nc], [zb], None, "epsilon", "epsilon")
    g = np.random.default_rng(0).normal(size=(4, 8))
    v = O.compute_v_ind(c, [zb], [g], "epsilon")
    assert np.allclose(A @ g.ravel(), (v * nc).sum(-1).ravel(), rtol=1e-12, atol=1e-15)


def test_static_wake_lumping_equals_explicit_wake():
    _, _, zb, zw = _wing(4, 8, 6)
    c, nc = O.compute_surf_c(zb), O.compute_surf_nc(zb)
    A = O.compute_aic_solve([c], [nc], [zb], [zw], "epsilon", "epsilon")
    g = np.random.default_rng(1).normal(size=(4, 8))
    gw = np.broadcast_to(g[-1], (6, 8))
    v = O.compute_v_ind(c, [zb, zw], [g, gw], "epsilon")
    assert np.allclose(A @ g.ravel(), (v * nc).sum(-1).ravel(), rtol=1e-12, atol=1e-15)


def test_static_no_penetration_and_solve():
    zeta0, hg, zb, _ = _wing(8, 32, 20)
    o = O.UVLMOracle([(8, 32, 20)], np.arange(33))
    o.keep_aic = True
    ff = O.Constant(np.array([1.0, 0.0, 0.0]), 1.225)
    o.set_design_variables(0.125, ff, zeta0, hg)
    case = o.solve_static()
    c, nc = case.c[0][0], case.nc[0][0]
    v = ff(c, 0.0) + O.compute_v_ind(c, [case.zeta_b[0][0], case.zeta_w[0][0]],
                                     [case.gamma_b[0][0], case.gamma_w[0][0]], "epsilon")
    vn = (v * nc).sum(-1)
    assert np.max(np.abs(vn)) < 1e-12 * np.max(np.abs((ff(c, 0.0) * nc).sum(-1)))
    # LAPACK: same solution and a backward-stable residual
    A = case.aic[0]
    rhs = -(ff(c, 0.0) * nc).sum(-1).ravel()
    x = scipy.linalg.lu_solve(scipy.linalg.lu_factor(A), rhs)
    assert np.allclose(case.gamma_b[0][0].ravel(), x, rtol=1e-12, atol=1e-15)
    # a lifting wing at +5 deg: lift coefficient between the 2D value and a
    # lifting-line estimate for AR 10 with a short wake
    lift = case.f_steady[0][0].sum(axis=(0, 1))[2]
    cl = lift / (0.5 * 1.225 * 1.0 * 10.0)
    assert 0.3 < cl < 2 * np.pi * np.deg2rad(5.0)


def test_prescribed_wake_convection_closed_form():
    _, _, zb, zw = _wing(4, 6, 5)
    gb = np.random.default_rng(2).normal(size=(4, 6))
    gw = np.random.default_rng(3).normal(size=(5, 6))
    u = np.array([1.0, 0.1, -0.2])
    zb_n = zb + np.array([0.0, 0.0, 0.01])
    zn, gn = O.propagate_surf_wake(gb, gw, zb_n, zw, None, lambda x: np.broadcast_to(u, x.shape), 0.25)
    assert np.array_equal(zn[0], zb_n[-1])
    assert np.allclose(zn[1:], zw[:-1] + 0.25 * u, rtol=0, atol=1e-15)
    assert np.array_equal(gn[0], gb[-1]) and np.array_equal(gn[1:], gw[:-1])


def test_free_wake_convection_is_pair_sum():
    zeta0, hg, zb, zw = _wing(4, 6, 5)
    rng = np.random.default_rng(4)
    gb, gw = rng.normal(size=(4, 6)) * 0.1, rng.normal(size=(5, 6)) * 0.1
    u = np.array([1.0, 0.0, 0.0])

    def v_func(x):
        return u + O.compute_v_ind(x, [zb, zw], [gb, gw], "epsilon")

    zn, _ = O.propagate_surf_wake(gb, gw, zb, zw, None, v_func, 0.1)
    base = zw[:-1].reshape(-1, 3)
    ref = base + 0.1 * (u + O.pair_sum(base, *np.concatenate(
        [np.stack(O.surface_segments(zb, gb)[:2]), np.stack(O.surface_segments(zw, gw)[:2])], axis=1),
        np.concatenate([O.surface_segments(zb, gb)[2], O.surface_segments(zw, gw)[2]])))
    assert np.allclose(zn[1:].reshape(-1, 3), ref, rtol=1e-14, atol=1e-15)


def _square_no_wake(u_inf):
    mn, width = 5, 1.0
    hg = np.zeros((mn + 1, 4, 4))
    hg[:, :3, :3] = np.eye(3)
    hg[:, 1, 3] = np.linspace(0.0, width, mn + 1)
    hg[:, 3, 3] = 1.0
    o = O.UVLMOracle([(mn, mn, 0)], np.arange(mn + 1))
    o.set_design_variables(1.0, O.Constant(np.asarray(u_inf, float), 1.225), O.make_rectangular_grid(mn, mn, width, 0.0), hg)
    return o.solve_static()


def test_static_no_wake_rotation_invariance_and_zero_force():
    """tests/aero/test_static_no_wake.py:11-65."""
    c0 = _square_no_wake([0.0, 10.0, 3.0])
    c1 = _square_no_wake([10.0, 0.0, 3.0])
    assert np.allclose(c0.gamma_b[0][0], c1.gamma_b[0][0])
    assert np.allclose(np.sum(c0.f_steady[0][0]), 0.0, atol=1e-5, rtol=1e-4)
    assert np.allclose(c0.f_steady[0][0], np.transpose(c1.f_steady[0][0], (1, 0, 2))[..., (1, 0, 2)], atol=1e-5,
                       rtol=1e-4)


def test_mirrored_wing_equals_full_wing():
    """tests/aero/test_mirrored_wing.py:11-78."""
    m, n, m_star, c_ref, b_ref = 5, 10, 20, 1.0, 5.0
    u_inf = np.array((10.0, 0.0, 1.0))
    dt = c_ref / (np.linalg.norm(u_inf) * m)
    ff = O.Constant(u_inf, 1.225)

    def hg_for(y):
        hg = np.zeros((len(y), 4, 4))
        hg[:, :3, :3] = np.eye(3)
        hg[:, 1, 3] = y
        hg[:, 3, 3] = 1.0
        return hg

    om = O.UVLMOracle([(m, n, m_star)], np.arange(n + 1), mirror_point=np.zeros(3), mirror_normal=np.array([0, 1.0, 0]))
    om.set_design_variables(dt, ff, O.make_rectangular_grid(m, n, c_ref, 0.0), hg_for(np.linspace(0, b_ref, n + 1)))
    of = O.UVLMOracle([(m, 2 * n, m_star)], np.arange(2 * n + 1))
    of.set_design_variables(dt, ff, O.make_rectangular_grid(m, 2 * n, c_ref, 0.0),
                            hg_for(np.linspace(-b_ref, b_ref, 2 * n + 1)))
    sm, sf = om.solve_static(), of.solve_static()
    assert np.allclose(sm.gamma_b[0][0], sf.gamma_b[0][0][:, n:], rtol=1e-10, atol=1e-12)
    assert np.allclose(sm.gamma_w[0][0], sf.gamma_w[0][0][:, n:], rtol=1e-10, atol=1e-12)


def test_equilibrium_persistence():
    """A steady rigid wing without motion keeps its circulation (prescribed wake)."""
    m, n, m_star, steps = 4, 8, 10, 6
    zeta0, hg, _, _ = _wing(m, n, m_star, alpha_deg=4.0, span=4.0)
    o = O.UVLMOracle([(m, n, m_star)], np.arange(n + 1))
    o.set_design_variables(1.0 / m, O.Constant(np.array([1.0, 0.0, 0.0]), 1.225), zeta0, hg)
    s = o.solve_static()
    hg_t = np.stack([hg] * steps)
    d = o.solve_prescribed_dynamic(s, hg_t, np.zeros_like(hg_t))
    g = d.gamma_b[0]
    assert np.max(np.abs(g - g[0])) < 1e-12 * np.max(np.abs(g[0]))
    assert np.max(np.abs(d.gamma_b_dot[0][1:])) < 1e-9
    assert np.allclose(d.zeta_w[0][-1], s.zeta_w[0][0], rtol=0, atol=1e-13)


def test_gust_flowfield_profile():
    ff = O.OneMinusCosine(np.array([2.0, 0.0, 0.0]), 1.225, True, gust_length=3.0, gust_amplitude=0.5)
    x = np.array([[1.5, 0.0, 0.0], [-1.0, 0, 0], [6.5, 0, 0]])
    u = ff(x, 0.0)
    assert np.allclose(u[0], [2.0, 0.0, 0.25])  # 0.5 * 0.5 * (1 - cos(pi/2))
    assert np.allclose(u[1], [2.0, 0, 0]) and np.allclose(u[2], [2.0, 0, 0])
    # relative motion: the gust travels with the flow
    assert np.allclose(ff(x[:1] + np.array([2.0, 0, 0]), 1.0), u[:1])
