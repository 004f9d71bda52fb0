"""Pins of the oracle's segment kernels and pair sums (CPU only).

Each check is chosen so that a plausible slip (a squared epsilon, a dropped
|r0|^4, a sign or index error, a transposed filament) fails it:
  * the reference's own kernel tests (tests/algebra/test_kernels.py:9-98);
  * closed forms of a straight filament (perpendicular bisector, infinite line);
  * an independent quadrature of the Biot-Savart line integral;
  * the exact algebraic relation epsilon-kernel = cutoff-form * |c|^2/(|c|^2 + eps|r0|^4);
  * an extended-precision (long double) evaluation;
  * de-duplicated filament sums == brute-force 4-filament ring sums;
  * the square-ring centre closed form 2 sqrt(2) Gamma / (pi a).
"""

import numpy as np
import pytest
from scipy import integrate

from oracle import uvlm_oracle as O


def test_reference_kernel_cases():
    y = np.array([[0.0, 1.0, 0.0], [0.0, -1.0, 0.0]])
    with np.errstate(all="ignore"):
        assert np.all(np.isnan(O.bs_plain(np.zeros(3), y[0], y[1])))
        assert np.all(np.isnan(O.bs_plain(np.array([0.0, 1.0, 0.0]), np.zeros(3), np.array([0.0, -1.0, 0.0]))))
    assert not np.all(np.isnan(O.bs_plain(np.array([0.1, 0, 0]), y[0], y[1])))
    assert np.allclose(O.bs_epsilon(np.zeros(3), y[0], y[1]), 0.0)
    assert np.allclose(O.bs_epsilon(np.array([0.0, 1.0, 0.0]), np.zeros(3), np.array([0.0, -1.0, 0.0])), 0.0)
    x = np.array([0.1, 0.0, 0.0])
    assert np.allclose(O.bs_plain(x, y[0], y[1]), O.bs_epsilon(x, y[0], y[1]), rtol=5e-5)
    # the 5e-5 tolerance is needed: the regularisation is visible at this point
    assert not np.allclose(O.bs_plain(x, y[0], y[1]), O.bs_epsilon(x, y[0], y[1]), rtol=1e-6)
    assert np.allclose(O.bs_cutoff(np.zeros(3), y[0], y[1]), 0.0)
    xc = np.array([1.01 * O.R_CUTOFF, 0, 0])
    assert np.allclose(O.bs_plain(xc, y[0], y[1]), O.bs_cutoff(xc, y[0], y[1]))
    assert np.allclose(O.bs_cutoff(np.array([0.99 * O.R_CUTOFF, 0, 0]), y[0], y[1]), 0.0)


@pytest.mark.parametrize("half_len,d", [(1.0, 0.1), (0.05, 0.3), (3.0, 2.0)])
def test_perpendicular_bisector_closed_form(half_len, d):
    # |v| = Gamma/(4 pi d) * 2L / sqrt(L^2 + d^2), along dl x r (here -z)
    y0, y1 = np.array([0.0, -half_len, 0.0]), np.array([0.0, half_len, 0.0])
    x = np.array([d, 0.0, 0.0])
    v = O.bs_plain(x, y0, y1)
    mag = 1.0 / (4 * np.pi * d) * 2 * half_len / np.sqrt(half_len**2 + d**2)
    assert np.allclose(v, [0.0, 0.0, -mag], rtol=1e-14, atol=0)
    # filament direction flips the sign
    assert np.allclose(O.bs_plain(x, y1, y0), [0.0, 0.0, mag], rtol=1e-14)


def test_infinite_line_limit():
    d = 0.37
    y0, y1 = np.array([0.0, -1e6, 0.0]), np.array([0.0, 1e6, 0.0])
    v = O.bs_plain(np.array([d, 0.0, 0.0]), y0, y1)
    assert abs(-v[2] - 1.0 / (2 * np.pi * d)) / (1.0 / (2 * np.pi * d)) < 1e-9
    # the epsilon kernel's eps |r0|^4 term grows with the filament length: on a
    # 2e6-long filament it dominates |c|^2 and suppresses the influence
    # (the reference's regularisation, utils.py:622), which this pins as well
    v_e = O.bs_epsilon(np.array([d, 0.0, 0.0]), y0, y1)
    c2 = (d * 2e6) ** 2
    assert np.isclose(v_e[2], v[2] * c2 / (c2 + O.EPSILON * (2e6) ** 4), rtol=1e-9)


def _quad_bs(x, y0, y1):
    """Gamma=1: 1/(4 pi) int_0^1 (dl x r)/|r|^3 ds, r = x - (y0 + s r0)."""
    r0 = y1 - y0

    def f(s, k):
        r = x - (y0 + s * r0)
        return np.cross(r0, r)[k] / np.linalg.norm(r) ** 3

    return np.array([integrate.quad(f, 0.0, 1.0, args=(k,), epsabs=0, epsrel=1e-13, limit=200)[0]
                     for k in range(3)]) / (4 * np.pi)


def test_plain_kernel_matches_quadrature():
    rng = np.random.default_rng(0)
    for _ in range(20):
        y0 = rng.normal(size=3)
        y1 = y0 + rng.normal(size=3)
        x = rng.normal(size=3) * 2
        q = _quad_bs(x, y0, y1)
        assert np.allclose(O.bs_plain(x, y0, y1), q, rtol=1e-10, atol=1e-13)


def test_epsilon_relation_to_cutoff_form():
    """bs_epsilon == (cutoff-form value) * |c|^2 / (|c|^2 + eps |r0|^4) exactly:
    pins the unsquared eps and the |r0|^4 of the denominator."""
    rng = np.random.default_rng(1)
    y0 = rng.normal(size=(500, 3))
    y1 = y0 + rng.normal(size=(500, 3)) * 0.05
    x = y0 + rng.normal(size=(500, 3)) * 0.02
    r0, r1, r2 = y1 - y0, x - y0, x - y1
    c = np.cross(r1, r2)
    cc = (c * c).sum(-1)
    r0sq = (r0 * r0).sum(-1)
    factor = cc / (cc + O.EPSILON * r0sq**2)
    assert np.min(factor) < 0.999  # the regularisation is active for some pairs
    v_cut = O.bs_cutoff(x, y0, y1, rcut=0.0)
    assert np.allclose(O.bs_epsilon(x, y0, y1), v_cut * factor[:, None], rtol=1e-12, atol=0)


def test_epsilon_kernel_extended_precision():
    rng = np.random.default_rng(2)
    y0 = rng.normal(size=(200, 3))
    y1 = y0 + rng.normal(size=(200, 3)) * 0.1
    x = rng.normal(size=(200, 3))
    ld = np.longdouble
    v_ld = O.bs_epsilon(x.astype(ld), y0.astype(ld), y1.astype(ld), eps=ld(O.EPSILON))
    assert np.allclose(O.bs_epsilon(x, y0, y1), v_ld.astype(float), rtol=1e-12, atol=1e-15)


def _ring_bruteforce(targets, zeta, gamma, kernel="epsilon"):
    """Sum over rings of 4 filaments each with the ring circulation (no de-duplication)."""
    m, n = gamma.shape
    v = np.zeros_like(targets)
    for i in range(m):
        for j in range(n):
            a, b, c, d = zeta[i, j], zeta[i, j + 1], zeta[i + 1, j + 1], zeta[i + 1, j]
            for p, q in ((a, b), (b, c), (c, d), (d, a)):
                v += gamma[i, j] * O.kernel_fn(kernel)(targets, p, q)
    return v


@pytest.mark.parametrize("kernel", ["epsilon", "plain"])
def test_dedup_equals_ring_bruteforce(kernel):
    rng = np.random.default_rng(4)
    m, n = 3, 5
    zeta = O.make_rectangular_grid(m, n, 1.0, 0.25)
    zeta[..., 1] = np.linspace(0, 2, n + 1)[None, :]
    zeta += rng.normal(size=zeta.shape) * 0.02
    gamma = rng.normal(size=(m, n))
    targets = rng.normal(size=(40, 3)) + np.array([0.5, 1.0, 0.3])
    v = O.compute_v_ind(targets, [zeta], [gamma], kernel)
    # ring orientation: (i,j)->(i,j+1)->(i+1,j+1)->(i+1,j) is the orientation whose
    # de-duplicated chordwise filament (i,j)->(i+1,j) carries G[i,j]-G[i,j-1]
    ref = -_ring_bruteforce(targets, zeta, gamma, kernel)
    assert np.allclose(v, ref, rtol=1e-12, atol=1e-14)


def test_square_ring_centre_closed_form():
    a, G = 0.4, 1.7
    zeta = np.zeros((2, 2, 3))
    zeta[1, :, 0] = a
    zeta[:, 1, 1] = a
    v = O.compute_v_ind(np.array([[a / 2, a / 2, 0.0]]), [zeta], [np.array([[G]])], "plain")[0]
    assert np.allclose(np.abs(v), [0, 0, 2 * np.sqrt(2) * G / (np.pi * a)], rtol=1e-14)
    # epsilon kernel differs by the regularisation only
    v_e = O.compute_v_ind(np.array([[a / 2, a / 2, 0.0]]), [zeta], [np.array([[G]])], "epsilon")[0]
    assert np.allclose(v_e, v, rtol=1e-6)


def test_mirror_equals_reflected_source():
    rng = np.random.default_rng(6)
    zeta = O.make_rectangular_grid(3, 4, 1.0, 0.0)
    zeta[..., 1] = np.linspace(0.2, 2.0, 5)[None, :]
    gamma = rng.normal(size=(3, 4))
    x = rng.normal(size=(10, 3))
    mp, mn = np.zeros(3), np.array([0.0, 1.0, 0.0])
    v_m = O.compute_v_ind(x, [zeta], [gamma], "epsilon", mp, mn)
    zr = O.mirror_grid(zeta, mp, mn)
    v_ref = O.compute_v_ind(x, [zeta], [gamma], "epsilon") - O.compute_v_ind(x, [zr], [gamma], "epsilon")
    assert np.allclose(v_m, v_ref, rtol=1e-14, atol=1e-16)
    # reflecting the grid reverses its ring orientation: the image carries -gamma
    v_img = O.compute_v_ind(x, [zr], [-gamma], "epsilon")
    assert np.allclose(v_m, O.compute_v_ind(x, [zeta], [gamma], "epsilon") + v_img, rtol=1e-14, atol=1e-16)


def test_linearity_in_gamma():
    rng = np.random.default_rng(8)
    zeta = O.make_rectangular_grid(4, 6, 1.0, 0.25)
    zeta[..., 1] = np.linspace(0, 3, 7)[None, :]
    g1, g2 = rng.normal(size=(4, 6)), rng.normal(size=(4, 6))
    x = rng.normal(size=(30, 3))
    v = O.compute_v_ind(x, [zeta], [2.0 * g1 - 3.0 * g2], "epsilon")
    v_ref = 2.0 * O.compute_v_ind(x, [zeta], [g1], "epsilon") - 3.0 * O.compute_v_ind(x, [zeta], [g2], "epsilon")
    assert np.allclose(v, v_ref, rtol=1e-12, atol=1e-15)


def test_c_restatement_matches_numpy_oracle():
    from oracle import c_oracle

    rng = np.random.default_rng(9)
    x = rng.normal(size=(300, 3))
    y0 = rng.normal(size=(400, 3))
    y1 = y0 + 0.1 * rng.normal(size=(400, 3))
    g = rng.normal(size=400)
    a = c_oracle.pair_sum_eps(x, y0, y1, g)
    b = O.pair_sum(x, y0, y1, g)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
