"""The adjoint driver's restarted GMRES on CPU tensors: nonsymmetric
systems with and without a (right) preconditioner, restarts, warm start."""

import numpy as np
import torch

from flapjax_amd.residuals import gmres


def _system(n=60, seed=0):
    rng = np.random.default_rng(seed)
    a = np.eye(n) * 3.0 + rng.normal(size=(n, n)) / np.sqrt(n)
    b = rng.normal(size=n)
    return torch.as_tensor(a), torch.as_tensor(b)


def test_gmres_solves_nonsymmetric_system():
    a, b = _system()
    x, it, rel = gmres(lambda v: a @ v, b, tol=1e-13, restart=80)
    assert rel < 1e-12 and np.allclose(x.numpy(), np.linalg.solve(a.numpy(), b.numpy()), atol=1e-10)


def test_gmres_restarts_and_preconditioner():
    a, b = _system(seed=1)
    x, it, rel = gmres(lambda v: a @ v, b, tol=1e-12, restart=5, maxiter=200)
    assert rel < 1e-11 and it > 5
    inv = torch.linalg.inv(a + 0.01 * torch.eye(a.shape[0], dtype=a.dtype))
    x2, it2, rel2 = gmres(lambda v: a @ v, b, tol=1e-12, restart=5, maxiter=200, precond=lambda v: inv @ v)
    assert rel2 < 1e-11 and it2 < it
    x3, it3, _ = gmres(lambda v: a @ v, b, x0=x2, tol=1e-12)
    assert it3 <= 1
