"""Banded LU solve (fj_band_solve) against LAPACK's banded and dense solvers.

scipy.linalg.solve_banded runs LAPACK dgbsv, the algorithm the kernel
restates (dgbtf2 pivoting and elimination, dgbtrs substitutions), so the
two agree to a few units of rounding; the dense LU (fj_lu_factor, partial
pivoting) picks the same pivot rows, so it agrees to the same level."""

import numpy as np
import pytest
import scipy.linalg as sl
import torch

from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _banded(n, kl, ku, rng, weak_diag=False):
    a = np.zeros((n, n))
    for d in range(-kl, ku + 1):
        idx = np.arange(max(0, -d), min(n, n - d))
        a[idx, idx + d] = rng.normal(size=idx.size)
    if not weak_diag:
        a[np.arange(n), np.arange(n)] += 2.0 * (kl + ku + 1)
    return a


def _to_ab(a, kl, ku):
    n = a.shape[0]
    ab = np.zeros((kl + ku + 1, n))
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            ab[ku + i - j, j] = a[i, j]
    return ab


@pytest.mark.parametrize("n,kl,ku,weak", [(1, 0, 0, False), (7, 2, 3, True), (64, 11, 11, True),
                                          (480, 11, 11, True), (480, 11, 11, False), (300, 0, 5, True),
                                          (200, 17, 4, True), (130, 63, 2, True)])
def test_band_solve_matches_lapack(dev, n, kl, ku, weak):
    from flapjax_amd import _kernels as K

    rng = np.random.default_rng(n + 7 * kl + ku)
    a = _banded(n, kl, ku, rng, weak)  # weak diagonal: rows get exchanged
    b = rng.normal(size=n)
    ref = sl.solve_banded((kl, ku), _to_ab(a, kl, ku), b)
    assert K.band_fits(n, kl, ku)
    x = K.band_solve(torch.as_tensor(a, device=dev), torch.as_tensor(b, device=dev), kl, ku).cpu().numpy()
    err = rel_err(x, ref)
    cond = np.linalg.cond(a)
    print(f"n={n} kl={kl} ku={ku}: rel err {err:.2e}, cond {cond:.1e}")
    assert err <= 64 * np.finfo(float).eps * max(1.0, cond)
    if (kl, ku) == (11, 11):
        # the fixed-bandwidth kernel restates dgbtf2 / dgbtrs operation by
        # operation: the same bits as LAPACK's dgbsv
        assert np.array_equal(x, ref)
    dense = K.lu_solve(K.lu_factor(torch.as_tensor(a, device=dev)), torch.as_tensor(b, device=dev)).cpu().numpy()
    assert rel_err(x, dense) <= 64 * np.finfo(float).eps * max(1.0, cond)
    # A and b are left as they were; x may alias b
    bt = torch.as_tensor(b, device=dev)
    K.band_solve(torch.as_tensor(a, device=dev), bt, kl, ku, out=bt)
    assert np.array_equal(bt.cpu().numpy(), x)


def test_band_solve_limits(dev):
    from flapjax_amd import _kernels as K
    from flapjax_amd._lib import FlapjaxLibError

    assert not K.band_fits(4000, 11, 11)
    assert not K.band_fits(100, 64, 0)
    a = torch.eye(8, dtype=torch.float64, device=dev)
    with pytest.raises(FlapjaxLibError):
        K.band_solve(a, torch.ones(8, dtype=torch.float64, device=dev), 64, 0)


def test_band_solve_singular_column(dev):
    """A zero column propagates to a non-finite solution (as LAPACK's
    solve of a singular system does) instead of raising."""
    from flapjax_amd import _kernels as K

    a = np.diag(np.arange(1.0, 9.0))
    a[3, 3] = 0.0
    x = K.band_solve(torch.as_tensor(a, device=dev), torch.ones(8, dtype=torch.float64, device=dev), 1, 1)
    assert not torch.isfinite(x).all()


def test_band_solve_beam_newton_system(dev):
    """A Newton system of the coupled gust wing (80 two-node elements, 480
    solve DOFs, kl = ku = 11; 37 of its columns pivot off the diagonal),
    tests/golden/beam_newton_system_480.npz: the same bits as dgbsv."""
    from pathlib import Path

    from flapjax_amd import _kernels as K

    d = np.load(Path(__file__).parent / "golden" / "beam_newton_system_480.npz")
    ab, b = d["ab"], d["b"]
    n = ab.shape[1]
    a = np.zeros((n, n))
    for j in range(n):
        for i in range(max(0, j - 11), min(n, j + 12)):
            a[i, j] = ab[11 + i - j, j]
    ref = sl.solve_banded((11, 11), ab, b)
    x = K.band_solve(torch.as_tensor(a, device=dev), torch.as_tensor(b, device=dev), 11, 11).cpu().numpy()
    assert np.array_equal(x, ref)


def test_band_solve_fixed_singular_column(dev):
    """kl = ku = 11 with a zero column: the exact pivot search finds no
    pivot and the solution is not finite (as the general kernel's)."""
    from flapjax_amd import _kernels as K

    rng = np.random.default_rng(11)
    a = _banded(64, 11, 11, rng)
    a[:, 20] = 0.0
    x = K.band_solve(torch.as_tensor(a, device=dev), torch.ones(64, dtype=torch.float64, device=dev), 11, 11)
    assert not torch.isfinite(x).all()
