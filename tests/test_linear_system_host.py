"""LinearSystem (utils/linear.py:680-822) on CPU tensors: the march, the
removal of u_{n+1} and the continuous-to-discrete maps."""

import numpy as np
import scipy.linalg
import torch

from flapjax_amd.linear_uvlm import LinearSystem


def _sys(seed=0, n=5, k=3, p=4):
    rng = np.random.default_rng(seed)
    t = lambda *s: torch.as_tensor(rng.normal(size=s) * 0.3)  # noqa: E731
    return LinearSystem(t(n, n), t(n, k), t(p, n), t(p, k), 0.1), rng


def test_run_matches_recurrence_and_remove_u_np1():
    s, rng = _sys()
    u = torch.as_tensor(rng.normal(size=(12, 3)))
    x0 = torch.as_tensor(rng.normal(size=5))
    x, y = s.run(u, x0)
    a, b, c, d = (v.numpy() for v in (s.a, s.b, s.c, s.d))
    xr = [x0.numpy()]
    for i in range(1, 12):
        xr.append(a @ xr[-1] + b @ u[i].numpy())
    xr = np.array(xr)
    assert np.allclose(x.numpy(), xr, rtol=1e-13, atol=1e-14)
    assert np.allclose(y.numpy(), xr @ c.T + u.numpy() @ d.T, rtol=1e-13, atol=1e-14)
    # x_n' = A x_{n-1}' + (A B) u_{n-1}: with x' = x - B u the same outputs
    s.remove_u_np1()
    assert np.allclose(s.b.numpy(), a @ b) and np.allclose(s.d.numpy(), c @ b + d)


def test_continuous_to_discrete():
    s, _ = _sys(1)
    s.continuous_time = True
    zoh = s.continuous_to_discrete("zoh")
    n = 5
    e = scipy.linalg.expm(0.1 * np.block([[s.a.numpy(), s.b.numpy()], [np.zeros((3, 8))]]))
    assert np.allclose(zoh.a.numpy(), e[:n, :n]) and np.allclose(zoh.b.numpy(), e[:n, n:])
    tus = s.continuous_to_discrete("tustin")
    inv = np.linalg.inv(np.eye(n) - 0.05 * s.a.numpy())
    assert np.allclose(tus.a.numpy(), inv @ (np.eye(n) + 0.05 * s.a.numpy()))
    assert np.allclose(tus.b.numpy(), inv @ s.b.numpy() * 0.1)
