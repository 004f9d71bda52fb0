"""Pins of the oracle's control-surface path (CPU):

  * add_control_surface (aero/utils.py:42-84): a rotation about the hinge
    line in closed form (R_y(theta) applied to each vertex's offset from its
    column's hinge point), vertices outside the slices untouched, lengths to
    the hinge kept, zero angle the identity;
  * exp_so3 (algebra/so3.py:168-189): orthogonal, det 1, and Rodrigues'
    rotation of a vector perpendicular to the axis;
  * the complex-step forward derivative of a grid function: the closed form
    d/dtheta R(theta)(x - h) = a x R(theta)(x - h) and a central difference;
  * hg_dot_to_zeta_b_dot with control surfaces (aero/uvlm.py:562-621): a
    central difference in time of hg_to_zeta_b under pitch-plunge motion and
    a moving aileron;
  * a missing angle/velocity pair raises the reference's ValueError.
"""

import numpy as np
import pytest

from oracle import uvlm_oracle as O

M, N = 6, 12
FLAP_M, FLAP_N = slice(4, None), slice(0, 4)


def _aileron(x0, *, aileron):
    return [O.add_control_surface(x0[0], aileron, FLAP_M, FLAP_N)]


def _ry(th):
    c, s = np.cos(th), np.sin(th)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


@pytest.mark.parametrize("theta", [0.0, 0.3, -1.2])
def test_add_control_surface_closed_form(theta):
    grid = O.make_rectangular_grid(M, N, 1.0, 0.25)
    grid[..., 2] += 0.01 * np.sin(np.arange(N + 1))[None, :]  # not flat: z offsets take part
    out = O.add_control_surface(grid, theta, FLAP_M, FLAP_N)
    rows, cols = np.arange(M + 1)[FLAP_M], np.arange(N + 1)[FLAP_N]
    exp = grid.copy()
    for j in cols:
        h = grid[rows[0], j]
        exp[rows, j] = (grid[rows, j] - h) @ _ry(theta).T + h
    assert np.max(np.abs(out - exp)) <= 1e-15
    mask = np.ones((M + 1, N + 1), bool)
    mask[np.ix_(rows, cols)] = False
    assert np.array_equal(out[mask], grid[mask])
    d0 = np.linalg.norm(grid[rows][:, cols] - grid[rows[0]][None, cols], axis=-1)
    d1 = np.linalg.norm(out[rows][:, cols] - out[rows[0]][None, cols], axis=-1)
    assert np.max(np.abs(d1 - d0)) <= 1e-15
    if theta == 0.0:
        assert np.array_equal(out, grid)


def test_exp_so3_rotation():
    rng = np.random.default_rng(1)
    for w in (rng.normal(size=3), 1e-7 * rng.normal(size=3), np.zeros(3)):
        r = O.exp_so3(w)
        assert np.max(np.abs(r @ r.T - np.eye(3))) < 1e-14
        assert abs(np.linalg.det(r) - 1.0) < 1e-14
    a = np.array((0.0, 0.0, 1.0))
    r = O.exp_so3(0.7 * a)
    assert np.allclose(r @ np.array((1.0, 0.0, 0.0)), (np.cos(0.7), np.sin(0.7), 0.0), atol=1e-15)


@pytest.mark.parametrize("theta", [0.0, 0.2, -0.5])
def test_complex_step_jvp_of_grid_function(theta):
    grid = O.make_rectangular_grid(M, N, 1.0, 0.25)
    vel = 1.7
    fn = lambda aileron: _aileron([grid], aileron=aileron)  # noqa: E731
    dz = O.complex_step_jvp(lambda **a: fn(**a), {"aileron": theta}, {"aileron": vel})[0]
    rows, cols = np.arange(M + 1)[FLAP_M], np.arange(N + 1)[FLAP_N]
    rot = fn(theta)[0]
    exp = np.zeros_like(grid)
    y = np.array((0.0, 1.0, 0.0))
    for j in cols:
        exp[rows, j] = vel * np.cross(y, rot[rows, j] - rot[rows[0], j])
    assert np.max(np.abs(dz - exp)) <= 1e-15 * max(1.0, np.max(np.abs(exp)))
    h = 1e-6
    fd = vel * (fn(theta + h)[0] - fn(theta - h)[0]) / (2.0 * h)
    assert np.max(np.abs(dz - fd)) <= 1e-8 * np.max(np.abs(fd))


@pytest.mark.parametrize("t0", [0.3, 2.1])
def test_zeta_dot_with_aileron_is_time_derivative(t0):
    span, omega, h = 4.0, 0.9, 1e-5
    zeta0 = O.make_rectangular_grid(M, N, 1.0, 0.25)
    args = (np.deg2rad(5.0), np.deg2rad(3.0), 0.05, omega)
    hg_t, hgd_t = O.pitch_plunge_motion(N, span, 3, h, *args, t0=t0 - h)
    t = t0 + h * np.arange(-1, 2)
    ang = np.deg2rad(4.0) * np.sin(2.3 * t) + np.deg2rad(1.0)
    vel = np.deg2rad(4.0) * 2.3 * np.cos(2.3 * t)
    o = O.UVLMOracle([(M, N, 2)], np.arange(N + 1), grid_func=_aileron)
    o.set_design_variables(0.1, O.Constant(np.array((1.0, 0.0, 0.0)), 1.225), zeta0, hg_t[0],
                           reference_cs_angles={"aileron": 0.0})
    fd = (o.hg_to_zeta_b(hg_t[2], {"aileron": ang[2]})[0] - o.hg_to_zeta_b(hg_t[0], {"aileron": ang[0]})[0]) / (2 * h)
    zd = o.hg_dot_to_zeta_b_dot(hgd_t[1], hg_t[1], {"aileron": ang[1]}, {"aileron": vel[1]})[0]
    assert np.max(np.abs(zd - fd)) < 1e-8 * np.max(np.abs(fd))
    # without the control-velocity term the derivative is measurably wrong
    z_no = o.hg_dot_to_zeta_b_dot(hgd_t[1], hg_t[1], {"aileron": ang[1]}, {"aileron": 0.0})[0]
    assert np.max(np.abs(z_no - fd)) > 1e-3 * np.max(np.abs(fd))


def test_missing_control_pair_raises():
    zeta0 = O.make_rectangular_grid(M, N, 1.0, 0.25)
    hg = O.rigid_wing_hg(N, 4.0)
    o = O.UVLMOracle([(M, N, 2)], np.arange(N + 1), grid_func=_aileron)
    o.set_design_variables(0.1, O.Constant(np.array((1.0, 0.0, 0.0)), 1.225), zeta0, hg,
                           reference_cs_angles={"aileron": 0.0})
    with pytest.raises(ValueError, match="Missing pair of control angles and velocities"):
        o.hg_dot_to_zeta_b_dot(np.zeros_like(hg), hg, {"aileron": 0.1}, {})
