"""GPU parity of AIC assembly (K2) and the circulation solve (K3)."""

import numpy as np
import pytest
import torch

from oracle import uvlm_oracle as O
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _t(x, dev):
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)


def _wing(m, n, m_star, alpha_deg=5.0, span=10.0):
    zeta0 = O.make_rectangular_grid(m, n, 1.0, 0.25)
    hg = O.rigid_wing_hg(n, span, np.deg2rad(alpha_deg))
    zb = O.vect_product(hg, zeta0)
    s = np.concatenate(([0.0], np.cumsum(np.ones(m_star) * (1.0 / m))))
    zw = zb[-1][None] + np.outer(s, [1.0, 0.0, 0.0])[:, None, :]
    return zb, zw


@pytest.mark.parametrize("kernel", ["epsilon", "cutoff"])
def test_aic_static_with_wake_lumping(dev, kernel):
    from flapjax_amd.aic import compute_aic_solve
    from flapjax_amd.utils import biot_savart_cutoff, biot_savart_epsilon

    kobj = {"epsilon": biot_savart_epsilon, "cutoff": biot_savart_cutoff}[kernel]
    zb, zw = _wing(8, 32, 20)
    c, nc = O.compute_surf_c(zb), O.compute_surf_nc(zb)
    ref = O.compute_aic_solve([c], [nc], [zb], [zw], kernel, kernel)
    A = compute_aic_solve([_t(c, dev)], [_t(nc, dev)], [_t(zb, dev)], [_t(zw, dev)], [kobj], [kobj]).cpu().numpy()
    assert A.shape == (256, 256)
    assert rel_err(A, ref) < 1e-12
    # dynamic (bound only), with a mirror plane
    mp, mn = np.zeros(3), np.array([0.0, 1.0, 0.0])
    ref_b = O.compute_aic_solve([c], [nc], [zb], None, kernel, kernel, mp, mn)
    A_b = compute_aic_solve([_t(c, dev)], [_t(nc, dev)], [_t(zb, dev)], None, [kobj], None, mirror_point=mp,
                            mirror_normal=mn).cpu().numpy()
    assert rel_err(A_b, ref_b) < 1e-12


def test_aic_times_gamma_equals_normal_velocity(dev):
    """A @ gamma == v_ind(c) . n (closed relation between K2 and K1)."""
    from flapjax_amd.aic import compute_aic_solve, compute_v_ind

    zb, _ = _wing(6, 20, 0)
    c, nc = O.compute_surf_c(zb), O.compute_surf_nc(zb)
    A = compute_aic_solve([_t(c, dev)], [_t(nc, dev)], [_t(zb, dev)], None, None, None)
    g = np.random.default_rng(1).normal(size=(6, 20))
    v = compute_v_ind(_t(c, dev), [_t(zb, dev)], [_t(g, dev)], None)
    lhs = (A @ _t(g.ravel(), dev)).cpu().numpy()
    rhs = (v * _t(nc, dev)).sum(-1).reshape(-1).cpu().numpy()
    assert rel_err(lhs, rhs) < 1e-12


@pytest.mark.parametrize("n", [1, 7, 33, 64, 100, 257, 1024, 1500, 2100, 4200])
def test_lu_solve_random(dev, n):
    from flapjax_amd import _kernels as K

    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + np.eye(n) * 0.5
    b = rng.normal(size=n)
    fac = K.lu_factor(_t(A, dev))
    LU, piv = fac.lu, fac.piv
    assert int(fac.info.item()) == 0
    x = K.lu_solve(fac, _t(b, dev)).cpu().numpy()
    ref = np.linalg.solve(A, b)
    # backward error of a stable LU, and agreement with LAPACK
    res = np.linalg.norm(A @ x - b, np.inf) / (np.linalg.norm(A, np.inf) * np.linalg.norm(x, np.inf))
    assert res < 50 * n * np.finfo(float).eps
    assert rel_err(x, ref) < 1e-9
    # the transposed solve (adjoint)
    xt = K.lu_solve(fac, _t(b, dev), trans=True).cpu().numpy()
    assert rel_err(xt, np.linalg.solve(A.T, b)) < 1e-9
    # the permutation and factors reproduce P A = L U (up to 2100: the
    # two-level blocking of the larger sizes included)
    if n <= 2100:
        LUh = LU.cpu().numpy()
        perm = piv.cpu().numpy()[n:]
        Lm = np.tril(LUh, -1) + np.eye(n)
        Um = np.triu(LUh)
        assert rel_err(Lm @ Um, A[perm]) < 1e-13 * max(1.0, n / 64)
    if n >= 1024:
        # LAPACK's pivot rows (random matrix: no near ties)
        import scipy.linalg

        assert np.array_equal(piv.cpu().numpy()[:n], scipy.linalg.lu_factor(A)[1])


def test_lu_pivots_match_lapack_order(dev):
    from flapjax_amd import _kernels as K
    import scipy.linalg

    rng = np.random.default_rng(7)
    A = rng.normal(size=(96, 96))
    fac = K.lu_factor(_t(A, dev))
    LU, piv = fac.lu, fac.piv
    lu_ref, piv_ref = scipy.linalg.lu_factor(A)
    assert np.array_equal(piv.cpu().numpy()[:96], piv_ref)
    assert rel_err(LU.cpu().numpy(), lu_ref) < 1e-12


def test_lu_negative_zero_is_not_a_pivot(dev):
    """-0.0 entries are |a| = 0 to the pivot search (a sign bit left in the
    search key would make them the largest), so LAPACK's rows come out."""
    from flapjax_amd import _kernels as K
    import scipy.linalg

    rng = np.random.default_rng(11)
    A = rng.normal(size=(80, 80))
    A[0, 0] = -0.0
    A[5:40, 3] = -0.0
    fac = K.lu_factor(_t(A, dev))
    assert int(fac.info.item()) == 0
    lu_ref, piv_ref = scipy.linalg.lu_factor(A)
    assert np.array_equal(fac.piv.cpu().numpy()[:80], piv_ref)
    assert rel_err(fac.lu.cpu().numpy(), lu_ref) < 1e-12


def test_singular_reports_info(dev):
    from flapjax_amd import _kernels as K

    A = np.ones((40, 40))
    assert int(K.lu_factor(_t(A, dev)).info.item()) == 2


@pytest.mark.parametrize("n", [256, 1024])
def test_refined_solver_on_aic(dev, n):
    """FP32 LU + FP64 refinement reaches the FP64 solution on a real AIC."""
    from flapjax_amd.aic import compute_aic_solve
    from flapjax_amd.linalg import RefinedSolver, solve

    m = 8 if n == 256 else 16
    zb, zw = _wing(m, n // m, 20)
    c, nc = O.compute_surf_c(zb), O.compute_surf_nc(zb)
    A = compute_aic_solve([_t(c, dev)], [_t(nc, dev)], [_t(zb, dev)], [_t(zw, dev)], None, [None])
    b = _t(-(np.array([1.0, 0.0, 0.0]) * nc).sum(-1).ravel(), dev)
    x64 = solve(A, b).cpu().numpy()
    rs = RefinedSolver()
    xr = rs.solve(A, b).cpu().numpy()
    print(f"refined solve n={n}: sweeps={rs.last_iters}, rel err vs FP64 LU={rel_err(xr, x64):.2e}")
    assert rel_err(xr, x64) < 1e-12
    assert rs.last_iters >= 0 and rs.factorisations == 1
    # a nearby matrix (same kept inverse) is solved to FP64 accuracy too
    A2 = A * (1.0 + 1e-7 * torch.cos(torch.arange(n * n, device=dev, dtype=torch.float64)).reshape(n, n))
    x2 = rs.solve(A2, b).cpu().numpy()
    assert rs.factorisations == 1 and rs.last_iters >= 1
    assert rel_err(x2, np.linalg.solve(A2.cpu().numpy(), b.cpu().numpy())) < 1e-12


@pytest.mark.parametrize("n", [40, 300, 1024])
def test_fp32_factor(dev, n):
    """FP32 LU on the matrix cores: P A = L U to FP32 accuracy, and the FP64
    refinement from it reaches the FP64 solution."""
    from flapjax_amd import _kernels as K

    rng = np.random.default_rng(n + 1)
    A = rng.normal(size=(n, n)) + np.eye(n) * 2
    b = rng.normal(size=n)
    fac = K.lu_factor_f32(torch.as_tensor(A, dtype=torch.float32, device=dev))
    LUh = fac.lu.cpu().numpy().astype(float)
    perm = fac.piv.cpu().numpy()[n:]
    Lm = np.tril(LUh, -1) + np.eye(n)
    assert rel_err(Lm @ np.triu(LUh), A[perm]) < 1e-5
    x, it = K.solve_refined(_t(A, dev), fac, _t(b, dev), max_iter=20)
    assert int(it.item()) > 0
    # dsgesv's stop test holds (backward error within sqrt(n) eps), so the
    # forward error is within cond(A) times that
    xh = x.cpu().numpy()
    bw = np.abs(A @ xh - b).max() / (np.abs(A).sum(1).max() * np.abs(xh).max())
    assert bw <= np.sqrt(n) * 2.0 ** -53
    assert rel_err(xh, np.linalg.solve(A, b)) < 4 * np.linalg.cond(A, np.inf) * np.sqrt(n) * 2.0 ** -53


@pytest.mark.parametrize("n", [1, 37, 64, 65, 300, 1024])
def test_lu_inverse(dev, n):
    """Explicit inverse from the factor (FP64 and FP32) against LAPACK's."""
    from flapjax_amd import _kernels as K

    rng = np.random.default_rng(n + 3)
    A = rng.normal(size=(n, n)) + np.eye(n) * 2
    ref = np.linalg.inv(A)
    X = K.lu_inverse(K.lu_factor(_t(A, dev))).cpu().numpy()
    assert rel_err(X, ref) < 1e-11
    assert np.abs(X @ A - np.eye(n)).max() < 1e-11
    X32 = K.lu_inverse(K.lu_factor_f32(torch.as_tensor(A, dtype=torch.float32, device=dev))).cpu().numpy()
    assert X32.dtype == np.float32
    assert rel_err(X32.astype(float), ref) < 1e-3


@pytest.mark.parametrize("n", [64, 500, 1024])
def test_solve_refined_inv(dev, n):
    """Refinement around an explicit inverse: FP64 and FP32 inverses reach the
    FP64 solution of the CURRENT matrix; failures report negative counts."""
    from flapjax_amd import _kernels as K

    rng = np.random.default_rng(n + 5)
    A = rng.normal(size=(n, n)) + np.eye(n) * 4
    b = rng.normal(size=n)
    ref = np.linalg.solve(A, b)
    X = K.lu_inverse(K.lu_factor(_t(A, dev)))
    x, it = K.solve_refined_inv(_t(A, dev), X, _t(b, dev))
    assert int(it.item()) >= 0
    assert rel_err(x.cpu().numpy(), ref) < 1e-12
    # the stop test: backward error within sqrt(n) eps
    xh = x.cpu().numpy()
    bw = np.abs(A @ xh - b).max() / (np.abs(A).sum(1).max() * np.abs(xh).max())
    assert bw <= np.sqrt(n) * 2.0 ** -53 * 4
    # FP32 inverse: more corrections, same answer
    X32 = K.lu_inverse(K.lu_factor_f32(torch.as_tensor(A, dtype=torch.float32, device=dev)))
    x32, it32 = K.solve_refined_inv(_t(A, dev), X32, _t(b, dev), max_iter=10)
    assert int(it32.item()) >= 1
    assert rel_err(x32.cpu().numpy(), ref) < 1e-12
    # a nearby matrix with the kept inverse
    A2 = A + 1e-6 * rng.normal(size=(n, n))
    x2, it2 = K.solve_refined_inv(_t(A2, dev), X, _t(b, dev))
    assert int(it2.item()) >= 1
    assert rel_err(x2.cpu().numpy(), np.linalg.solve(A2, b)) < 1e-12
    # an unrelated inverse does not pass the stop test: negative count
    Xbad = K.lu_inverse(K.lu_factor(_t(rng.normal(size=(n, n)) + np.eye(n) * 4, dev)))
    _, itb = K.solve_refined_inv(_t(A, dev), Xbad, _t(b, dev), max_iter=2)
    assert int(itb.item()) == -2
    # non-finite right-hand side: negative count
    bn = b.copy()
    bn[n // 2] = np.nan
    _, itn = K.solve_refined_inv(_t(A, dev), X, _t(bn, dev), max_iter=2)
    assert int(itn.item()) < 0
    # zero right-hand side: x = 0, converged at once
    x0, it0 = K.solve_refined_inv(_t(A, dev), X, _t(np.zeros(n), dev))
    assert int(it0.item()) == 0 and not x0.any()
