"""Pins of the beam / coupling oracle (oracle/beam_oracle.py) against the
reference's own structural tests, closed forms and finite differences."""

from math import factorial

import numpy as np
import pytest

from oracle import beam_oracle as B

RNG = np.random.default_rng(7)


def _series_exp(x, order=30):
    out = np.eye(x.shape[0])
    for i in range(1, order + 1):
        out = out + np.linalg.matrix_power(x, i) / factorial(i)
    return out


@pytest.mark.parametrize("scale", [0.7, 2e-6, 0.0])
def test_exp_log_against_series(scale):
    """exp_se3 (both branches) against the matrix-exponential series; log is
    its inverse."""
    for _ in range(5):
        h = np.concatenate((RNG.normal(size=3), scale * RNG.normal(size=3)))
        g = B.exp_se3(h)
        ref = _series_exp(B.tilde4(h))
        # below the small-angle switch the reference truncates the series
        # after the quadratic term: the cubic term bounds the difference
        w = h[3:]
        trunc = np.linalg.norm(w) ** 2 * np.linalg.norm(h[:3]) if w @ w <= B.SMALL_ANG_THRESH else 0.0
        assert np.allclose(g, ref, rtol=0, atol=1e-13 + trunc)
        assert np.allclose(B.log_se3(g), h, rtol=0, atol=1e-12 + trunc)
        assert np.allclose(B.exp_so3(w), _series_exp(B.skew(w)), atol=1e-14)


@pytest.mark.parametrize("scale,tol", [(0.9, 1e-13), (1e-3, 1e-9), (0.0, 1e-13)])
def test_tangent_operator_against_differences(scale, tol):
    """exp(h + e dh) = exp(h) exp(e T(h) dh + O(e^2)); T^-1 T = I.  Above the
    small-angle switch the reference's closed forms cancel: their rounding
    error grows like eps / |omega|^2 (about 2e-10 at |omega| = 1e-3)."""
    for _ in range(4):
        h = np.concatenate((RNG.normal(size=3), scale * RNG.normal(size=3)))
        dh = RNG.normal(size=6)
        e = 1e-6
        num = (B.log_se3(np.linalg.inv(B.exp_se3(h)) @ B.exp_se3(h + e * dh))
               - B.log_se3(np.linalg.inv(B.exp_se3(h)) @ B.exp_se3(h - e * dh))) / (2 * e)
        assert np.allclose(num, B.t_se3(h) @ dh, atol=2e-8)
        assert np.allclose(B.t_inv_se3(h) @ B.t_se3(h), np.eye(6), atol=tol)
        w = h[3:]
        assert np.allclose(B.t_inv_so3(w) @ B.t_so3(w), np.eye(3), atol=tol)


def _bent_beam(n_nodes=4):
    b = B.BeamOracle(n_nodes, B.chain(n_nodes), np.array((0.0, 0.0, 1.0)))
    coords = np.zeros((n_nodes, 3))
    coords[:, 1] = np.linspace(0.0, 2.0, n_nodes)
    coords[:, 0] = 0.1 * np.arange(n_nodes) ** 2
    k = np.diag((3e4, 2e4, 2e4, 50.0, 80.0, 120.0))
    m = np.diag((1.3, 1.3, 1.3, 0.2, 0.1, 0.15))
    b.set_design_variables(coords, k, m)
    return b


def _deformed(b, amp=0.2):
    phi = amp * RNG.normal(size=(b.n_nodes, 6))
    return b.update_hg(b.hg0, phi)


def test_p_operator_against_differences():
    """P(d) = dd/dh_AB under right perturbations of the node frames."""
    b = _bent_beam()
    hg = _deformed(b)
    d = b.make_d(hg)
    dh = RNG.normal(size=(b.n_nodes, 6))
    e = 1e-6
    dp = b.make_d(b.update_hg(hg, e * dh))
    dm = b.make_d(b.update_hg(hg, -e * dh))
    for k in range(b.n_elem):
        p = B.p_mat(d[k], b.ad_inv[k])
        assert np.allclose((dp[k] - dm[k]) / (2 * e), p @ b.elem_vec(dh, k), atol=1e-7)


def test_element_strain_forces_reference():
    """tests/structure/static_primal/test_multi_element_beam.py: axial and
    torsional strain of a 10-node beam of length 2."""
    n = 10
    b = B.BeamOracle(n, B.chain(n), np.array((0.0, 1.0, 0.0)))
    coords = np.zeros((n, 3))
    coords[:, 0] = np.linspace(0, 2.0, n)
    for comp in (0, 3):
        k = np.full(6, 1e5)
        k[comp] = 4.56
        b.set_design_variables(coords, np.diag(k), None)
        d = np.zeros((n - 1, 6))
        d[:, 0] = 2.0 / (n - 1)
        eps0 = b.make_eps(d)
        assert np.allclose(eps0, 0.0)
        dx = 0.1
        d[:, comp] += dx / (n - 1)
        eps = b.make_eps(d)
        expect = np.zeros(6)
        expect[comp] = dx / 2.0
        assert np.allclose(eps, expect[None])
        f = b.assemble_vec([B.p_mat(d[e], b.ad_inv[e]).T @ -(b.kmat(e) @ eps[e]) for e in range(n - 1)]).reshape(-1, 6)
        rot = np.einsum("ij,kj->ki", B.chi(b.o0[0].T), f)
        ef = 4.56 * dx / 2.0
        assert np.isclose(rot[0, comp], ef) and np.isclose(rot[-1, comp], -ef)
        rest = np.delete(rot.ravel(), [comp, 6 * (n - 1) + comp])
        assert np.allclose(rest, 0.0)  # the reference test's default tolerance


@pytest.mark.parametrize("load,z_tip,rot", [(600.0, -2.6819e-3, -8.025680e-4), (600000.0, -2.157409, -6.721341e-1)])
def test_geradin_reference_values(load, z_tip, rot):
    """tests/structure/static_primal/test_geradin_beam.py (1 % tolerance)."""
    b = B.geradin_beam(20)
    f = np.zeros((b.n_nodes, 6))
    f[-1, 2] = -load
    out = b.static_solve(np.arange(6), f_ext_dead=f, load_steps=3)
    zt = out["hg"][-1, 2, 3]
    r = -B.log_so3(out["hg"][-1, :3, :3])[1]
    assert np.isclose(zt, z_tip, rtol=1e-2)
    assert np.isclose(r, rot, rtol=1e-2)


def test_small_tip_load_closed_form():
    """Cantilever tip rotation P L^2 / 2EI and deflection under a small
    follower tip load.  Constant-curvature elements take each element's mean
    moment, which leaves nodal rotations exact and integrates the rotation
    with the trapezoidal rule: deflection P L^3 / 3EI (1 - 1 / 4n^2) + P L / GA
    for n elements."""
    b = B.geradin_beam(11)
    p = 50.0
    f = np.zeros((b.n_nodes, 6))
    f[-1, 2] = -p
    out = b.static_solve(np.arange(6), f_ext_follower=f)
    ei, ga, l = 9.345e6, 3.231e8, 5.0
    n = b.n_elem
    assert np.isclose(out["hg"][-1, 2, 3], -(p * l ** 3 / (3 * ei) * (1 - 1 / (4 * n * n)) + p * l / ga), rtol=1e-6)
    assert np.isclose(-B.log_so3(out["hg"][-1, :3, :3])[1], -p * l ** 2 / (2 * ei), rtol=1e-6)


def test_mass_matrix_rigid_kinetic_energy():
    """1/2 v^T M v of a rigid motion equals the closed-form kinetic energy of
    a uniform straight bar (Simpson's rule is exact for it)."""
    n = 5
    b = B.BeamOracle(n, B.chain(n), np.array((0.0, 1.0, 0.0)))
    coords = np.zeros((n, 3))
    coords[:, 0] = np.linspace(0.0, 3.0, n)
    mu, jz = 2.5, 0.4
    b.set_design_variables(coords, np.diag(np.full(6, 1e3)), np.diag((mu, mu, mu, 0.3, 0.2, jz)))
    d = b.make_d(b.hg0)
    mt = b.assemble_mat([b.element(e, d[e], np.zeros((n, 6)), np.zeros((n, 6)))["m_t"] for e in range(n - 1)])
    u = np.array((0.3, -1.1, 0.7))
    v = np.zeros((n, 6))
    v[:, :3] = u
    assert np.isclose(0.5 * v.ravel() @ mt @ v.ravel(), 0.5 * mu * 3.0 * u @ u, rtol=1e-13)
    om = 1.7
    v = np.zeros((n, 6))
    v[:, 1] = om * coords[:, 0]
    v[:, 5] = om
    ke = 0.5 * mu * om ** 2 * 3.0 ** 3 / 3.0 + 0.5 * jz * om ** 2 * 3.0
    assert np.isclose(0.5 * v.ravel() @ mt @ v.ravel(), ke, rtol=1e-13)


def test_q_dot_complex_step_against_differences():
    b = _bent_beam()
    d = b.make_d(_deformed(b, 0.3))
    for k in range(b.n_elem):
        dd = RNG.normal(size=6)
        for s in B.LEGENDRE3:
            e = 1e-6
            num = (B.q_mat(s, d[k] + e * dd, b.ad_inv[k]) - B.q_mat(s, d[k] - e * dd, b.ad_inv[k])) / (2 * e)
            assert np.allclose(B.q_dot_mat(s, d[k], dd, b.ad_inv[k]), num, atol=1e-7)


def test_rigid_rotation_about_beam_axis():
    """tests/structure/dynamic_primal/test_const_initial_rotational_velocity.py,
    beam along x spinning about x: angle = omega t, no deformation."""
    omega, dt, n_t = 10.0, 0.001, 60
    b = B.BeamOracle(3, np.array([[0, 1], [1, 2]]), np.array((0.0, 1.0, 0.0)), spectral_radius=1.0)
    coords = np.zeros((3, 3))
    coords[:, 0] = (-4.56, 0.0, 4.56)
    b.set_design_variables(coords, np.diag(np.full(6, 1e4)), np.eye(6))
    v0 = np.zeros((3, 6))
    v0[:, 3] = omega
    out = b.dynamic_solve(dict(hg=b.hg0, v=v0), None, n_t, dt)
    th = omega * dt * np.arange(n_t)
    assert np.allclose(out["hg"][:, 0, :3, 3], coords[0], atol=1e-12)
    ang = np.array([B.log_so3(h[1, :3, :3])[0] for h in out["hg"]])
    assert np.allclose(ang, np.arctan2(np.sin(th), np.cos(th)), atol=1e-10)
    assert np.allclose(out["v"][:, :, 3], omega, atol=1e-10)
    assert np.allclose(out["v_dot"], 0.0, atol=1e-8)


def test_rigid_translation():
    b = _bent_beam()
    v0 = np.zeros((b.n_nodes, 6))
    v0[:, :3] = (0.4, -0.2, 0.3)
    out = b.dynamic_solve(dict(hg=b.hg0, v=v0), None, 8, 0.05)
    t = 0.05 * np.arange(8)
    assert np.allclose(out["hg"][:, :, :3, 3], b.x0[None] + t[:, None, None] * v0[None, :, :3], atol=1e-12)


def test_projection_closed_form():
    """aero/utils.py:666-697: force sum and moment (R x0) x f per strip."""
    n = 4
    hg = B.BeamOracle.update_hg(np.broadcast_to(np.eye(4), (n, 4, 4)), 0.3 * RNG.normal(size=(n, 6)))
    x0 = RNG.normal(size=(3, n, 3))
    f = RNG.normal(size=(3, n, 3))
    out = B.project_forcing_to_beam([f], hg[:, :3, :3], [np.arange(n)], [x0])
    for j in range(n):
        assert np.allclose(out[j, :3], f[:, j].sum(0))
        mom = sum(np.cross(hg[j, :3, :3] @ x0[i, j], f[i, j]) for i in range(3))
        assert np.allclose(out[j, 3:], mom)


def test_static_aeroelastic_equilibrium_and_persistence():
    """coupled/coupled.py static FSI converges to a state whose aero load
    balances the internal forces; time stepping from it keeps it
    (tests/coupled/dynamic/test_dynamic_equilibrium.py, shortened; with the
    tolerances 100x tighter than the reference test's, the drift bound is
    tightened 2x on this coarser wing: the drift is the equilibrium's
    convergence error)."""
    strict = B.ConvergenceSettings(30, 1e-9, 1e-11, 1e-9, 1e-11)
    cw = B.cantilever_wing(settings=strict, fsi=strict)
    st = cw.static_solve(np.arange(6))
    hg = st["hg"]
    assert hg[-1, 2, 3] > 1e-4  # the wing bends up
    f = cw._forces(st["aero"], hg, False)
    f_res, _, _, _ = cw.beam.residual(hg, cw.beam.solve_dofs(np.arange(6)), None, f)
    assert np.abs(f_res).max() < 1e-6 * np.abs(f).max()
    init = cw.uvlm.solve_static(hg)
    dyn = cw.dynamic_solve(hg, init, np.arange(6), 6)
    g = dyn["aero"].gamma_b[0]
    assert np.abs(g - g[[0]]).max() < 3e-7
    vp = dyn["varphi"]
    assert np.abs(vp - vp[[0]]).max() < 3e-7


def _cantilever(nn, ei_flap=1.0, ei_lag=4.0, gj=1.0, mu=1.0, jx=1e-2):
    """Uniform cantilever of length 1 along x (clamped at node 0), Euler-Bernoulli
    proportions: axial and shear stiffness 1e6, negligible rotary inertia of
    bending.  With y_vector = y the element's local z is global z, so k_cs[4]
    bends about y (flap, deflection along z) and k_cs[5] about z (lag)."""
    b = B.BeamOracle(nn, B.chain(nn), np.array((0.0, 1.0, 0.0)))
    x = np.zeros((nn, 3))
    x[:, 0] = np.linspace(0.0, 1.0, nn)
    b.set_design_variables(x, np.diag((1e6, 1e6, 1e6, gj, ei_flap, ei_lag)), np.diag((mu, mu, mu, jx, 1e-6, 1e-6)))
    return b


def _modes(b):
    import scipy.linalg

    d = b.make_d(b.hg0)
    z = np.zeros((b.n_nodes, 6))
    els = [b.element(e, d[e], z, z) for e in range(b.n_elem)]
    k = b.assemble_mat([el["k_t"] for el in els])[6:, 6:]
    m = b.assemble_mat([el["m_t"] for el in els])[6:, 6:]
    lam, vec = scipy.linalg.eigh(k, m)
    return np.sqrt(lam), vec


def test_cantilever_natural_frequencies():
    """Mass and stiffness of the element (structure/utils.py, structure/beam.py)
    against the Euler-Bernoulli cantilever: omega_k = (beta_k L)^2 sqrt(EI / (mu L^4))
    with beta_1 L = 1.8751, beta_2 L = 4.6941, for flap (EI = 1) and lag (EI = 4),
    and the clamped-free torsion omega = (pi / 2L) sqrt(GJ / J); the error of the
    constant-strain element falls as h^2."""
    b1, b2 = 1.8751040687, 4.6940911330
    w, _ = _modes(_cantilever(21))
    expect = np.sort([b1 ** 2, 2.0 * b1 ** 2, b2 ** 2, 2.0 * b2 ** 2, 0.5 * np.pi * np.sqrt(1.0 / 1e-2)])
    got = np.sort(w[:5])
    assert np.allclose(got, expect, rtol=5e-3), (got, expect)
    assert abs(w[0] / b1 ** 2 - 1.0) < 1e-3
    e11 = _modes(_cantilever(11))[0][0] / b1 ** 2 - 1.0
    e21 = w[0] / b1 ** 2 - 1.0
    assert 3.0 < e11 / e21 < 5.0


def test_cantilever_free_vibration_period():
    """Generalised-alpha dynamics (structure/time_integration.py) of the first
    flap mode, started by a small velocity in its discrete mode shape: the tip
    moves along z only, with the discrete period 2 pi / omega_1 (to the scheme's
    O((omega dt)^2) elongation) and an amplitude that the spectral radius 0.9
    barely damps at 100 steps per period."""
    b = _cantilever(11)
    w, vec = _modes(b)
    period = 2.0 * np.pi / w[0]
    v0 = np.zeros((b.n_nodes, 6))
    v0[1:] = 1e-4 * vec[:, 0].reshape(-1, 6) / np.abs(vec[:, 0]).max()
    dt = period / 100
    out = b.dynamic_solve(dict(hg=b.hg0, v=v0), np.arange(6), 112, dt)
    tip = out["hg"][:, -1, :3, 3] - out["hg"][0, -1, :3, 3]
    assert np.abs(tip[:, 1]).max() < 1e-6 * np.abs(tip[:, 2]).max()
    s = tip[:, 2]
    t = np.arange(s.size) * dt
    zc = [t[i] - s[i] * dt / (s[i + 1] - s[i]) for i in range(1, s.size - 1) if s[i] * s[i + 1] < 0]
    assert len(zc) == 2
    assert abs(2.0 * (zc[1] - zc[0]) / period - 1.0) < 1e-3
    assert abs(zc[0] / (0.5 * period) - 1.0) < 1e-3
    a1, a2 = np.abs(s[: s.size // 2]).max(), np.abs(s[s.size // 2:]).max()
    assert 0.97 < a2 / a1 <= 1.0
