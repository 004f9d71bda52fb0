"""Differentiation form: adjoint kernels vs central finite differences.

* PairSum backward (dL/dx, dL/dy0, dL/dy1, dL/dg) vs FD of the FP64 forward;
* LinearSolve / aic_solve backward (transposed LU solve + adjoint pair sums);
* the checkpointed differentiable UVLM: its forward equals the FP64 oracle's
  run, and dJ/d(twist, u_inf) matches central FD (step 1e-6) on 8 seeded
  random directions.
"""

import numpy as np
import pytest
import torch

from oracle import uvlm_oracle as O
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _t(x, dev):
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)


def test_pair_sum_backward_vs_fd(dev):
    from flapjax_amd import autograd as AG

    rng = np.random.default_rng(0)
    x = _t(rng.normal(size=(40, 3)), dev)
    y0 = _t(rng.normal(size=(30, 3)), dev)
    y1 = y0 + _t(rng.normal(size=(30, 3)) * 0.3, dev)
    g = _t(rng.normal(size=30), dev)
    w = _t(rng.normal(size=(40, 3)), dev)
    leaves = [t.clone().requires_grad_(True) for t in (x, y0, y1, g)]
    L = (AG.pair_sum(*leaves) * w).sum()
    grads = torch.autograd.grad(L, leaves)
    h = 1e-6
    for li in range(4):
        d = _t(rng.normal(size=leaves[li].shape), dev)
        args_p = [t.detach().clone() for t in leaves]
        args_m = [t.detach().clone() for t in leaves]
        args_p[li] += h * d
        args_m[li] -= h * d
        fd = ((AG.pair_sum(*args_p) - AG.pair_sum(*args_m)) * w).sum() / (2 * h)
        an = (grads[li] * d).sum()
        assert abs(float(fd - an)) <= 1e-6 * max(1.0, abs(float(an))), (li, float(fd), float(an))


def _bs_torch(x, y0, y1, g, kind, eps, rcut):
    """The three filament kernels written out in torch FP64 (autograd gives
    the exact derivatives of this expression)."""
    from flapjax_amd import _lib as Lb

    r1 = x[:, None, :] - y0[None]
    r2 = x[:, None, :] - y1[None]
    r0 = (y1 - y0)[None].expand_as(r1)
    c = torch.cross(r1, r2, dim=-1)
    e2 = 0.0 if kind == Lb.KERNEL_PLAIN else eps * eps
    s1 = (r1 * r1).sum(-1, keepdim=True) + e2
    s2 = (r2 * r2).sum(-1, keepdim=True) + e2
    d = r1 / torch.sqrt(s1) - r2 / torch.sqrt(s2)
    num = (r0 * d).sum(-1)
    den = (c * c).sum(-1)
    if kind == Lb.KERNEL_EPSILON:
        den = den + eps * (r0 * r0).sum(-1) ** 2
    q = g[None] * num / den
    if kind == Lb.KERNEL_CUTOFF:
        a = (r0 * r1).sum(-1)
        dist2 = (r1 * r1).sum(-1) - a * a / (r0 * r0).sum(-1)
        q = torch.where(dist2 > rcut * rcut, q, torch.zeros_like(q))
    return (c * q[..., None]).sum(1) / (4.0 * np.pi)


@pytest.mark.parametrize("kname", ["biot_savart_epsilon", "biot_savart_cutoff", "biot_savart"])
@pytest.mark.parametrize("n_tgt,n_seg", [(37, 45), (301, 1037), (600, 9000)])
def test_pair_sum_backward_one_pass_vs_torch(dev, kname, n_tgt, n_seg):
    """The one-pass adjoint kernel (filament tiles x target chunks, a
    transposed butterfly for the target side, fixed-order partial sums)
    against torch autograd of the same FP64 expression: all four gradients,
    ragged tile and chunk edges included, norm-wise within 1e-11."""
    from flapjax_amd import autograd as AG
    from flapjax_amd import utils as U

    kern = getattr(U, kname)
    rng = np.random.default_rng(n_tgt + n_seg)
    x = _t(rng.uniform(-2, 2, size=(n_tgt, 3)), dev)
    y0 = _t(rng.uniform(-2, 2, size=(n_seg, 3)), dev)
    y1 = y0 + _t(rng.normal(size=(n_seg, 3)) * 0.2, dev)
    g = _t(rng.normal(size=n_seg), dev)
    w = _t(rng.normal(size=(n_tgt, 3)), dev)
    a = [t.clone().requires_grad_(True) for t in (x, y0, y1, g)]
    v_dev = AG.pair_sum(*a, kernel=kern)
    got = torch.autograd.grad((v_dev * w).sum(), a)
    b = [t.clone().requires_grad_(True) for t in (x, y0, y1, g)]
    v_ref = _bs_torch(*b, kern.kind, kern.eps, kern.rcut)
    want = torch.autograd.grad((v_ref * w).sum(), b)

    assert rel_err(v_dev.detach().cpu().numpy(), v_ref.detach().cpu().numpy()) < 1e-12
    for i, (p, q) in enumerate(zip(got, want)):
        e = rel_err(p.cpu().numpy(), q.cpu().numpy())
        assert e < 1e-11, (kname, i, e)


def test_linear_solve_backward(dev):
    from flapjax_amd import autograd as AG

    rng = np.random.default_rng(1)
    n = 70
    A = _t(rng.normal(size=(n, n)) + 4 * np.eye(n), dev).requires_grad_(True)
    b = _t(rng.normal(size=n), dev).requires_grad_(True)
    w = _t(rng.normal(size=n), dev)
    L = (AG.linear_solve(A, b) * w).sum()
    gA, gb = torch.autograd.grad(L, (A, b))
    An, bn, wn = A.detach().cpu().numpy(), b.detach().cpu().numpy(), w.cpu().numpy()
    lam = np.linalg.solve(An.T, wn)
    x = np.linalg.solve(An, bn)
    assert np.allclose(gb.cpu().numpy(), lam, rtol=1e-10, atol=1e-12)
    assert np.allclose(gA.cpu().numpy(), -np.outer(lam, x), rtol=1e-10, atol=1e-12)


def _wing(m, n, dev):
    zeta0 = _t(O.make_rectangular_grid(m, n, 1.0, 0.25), dev)
    hg = _t(O.rigid_wing_hg(n, 4.0, np.deg2rad(4.0)), dev)
    return zeta0, hg


def test_aic_solve_geometry_gradient(dev):
    from flapjax_amd import adjoint as A
    from flapjax_amd import autograd as AG

    m, n = 4, 8
    zeta0, hg = _wing(m, n, dev)
    rng = np.random.default_rng(2)
    u = _t([1.0, 0.0, 0.1], dev)
    w = _t(rng.normal(size=m * n), dev)

    def f(twist):
        zb = A.apply_hg(hg, A.twisted_grid(zeta0, twist))
        c, nc = A.colloc(zb), A.normals(zb)
        b = -(u.expand_as(c) * nc).sum(-1).reshape(-1)
        s = torch.arange(6, dtype=F64, device=dev) * 0.25
        zw = zb[-1][None] + s[:, None, None] * (u / torch.linalg.norm(u))
        return (AG.aic_solve(b, c.reshape(-1, 3), nc.reshape(-1, 3), [zb], [zw]) * w).sum()

    tw = _t(rng.normal(size=n + 1) * 0.02, dev).requires_grad_(True)
    (g,) = torch.autograd.grad(f(tw), tw)
    h = 1e-6
    for _ in range(4):
        d = _t(rng.normal(size=n + 1), dev)
        fd = (f(tw.detach() + h * d) - f(tw.detach() - h * d)) / (2 * h)
        an = (g * d).sum()
        assert abs(float(fd - an)) <= 2e-6 * max(1.0, abs(float(an))), (float(fd), float(an))


def _model_setup(dev, m=4, n=8, m_star=8, steps=12, k=4):
    from flapjax_amd import adjoint as A

    dt = 1.0 / m
    zeta0 = _t(O.make_rectangular_grid(m, n, 1.0, 0.25), dev)
    hg_t, hgd_t = O.pitch_plunge_motion(n, 4.0, steps, dt, np.deg2rad(4.0), np.deg2rad(3.0), 0.05, 0.4)
    model = A.DifferentiableUVLM(m, n, m_star, dt, free_wake=True, checkpoint_every=k)
    return model, zeta0, _t(hg_t, dev), _t(hgd_t, dev), hg_t, hgd_t


def test_differentiable_forward_matches_oracle(dev):
    """Zero twist: the differentiable run is the FP64 UVLM of the oracle."""
    model, zeta0, hg_t, hgd_t, hg_np, hgd_np = _model_setup(dev)
    m, n, m_star = model.m, model.n, model.m_star
    J = model.objective(torch.zeros(n + 1, dtype=F64, device=dev), _t([1.0, 0, 0], dev), zeta0, hg_t, hgd_t)
    orc = O.UVLMOracle([(m, n, m_star)], np.arange(n + 1), free_wake=True)
    orc.set_design_variables(model.dt, O.Constant(np.array([1.0, 0, 0]), 1.225), zeta0.cpu().numpy(), hg_np[0])
    d = orc.solve_prescribed_dynamic(orc.solve_static(hg=hg_np[0]), hg_np, hgd_np)
    J_ref = model.dt * sum(d.f_steady[0][i][..., 2].sum() + d.f_unsteady[0][i][..., 2].sum()
                           for i in range(1, hg_np.shape[0]))
    assert abs(float(J) - J_ref) <= 1e-10 * abs(J_ref)


def test_uvlm_adjoint_vs_central_fd(dev):
    """dJ/d(twist, u_inf) by the checkpointed adjoint vs central FD (h = 1e-6)
    on 8 seeded random directions."""
    from flapjax_amd import adjoint as A

    model, zeta0, hg_t, hgd_t, _, _ = _model_setup(dev)
    n = model.n
    rng = np.random.default_rng(3)
    tw0 = _t(rng.normal(size=n + 1) * 0.01, dev)
    u0 = _t([1.0, 0.0, 0.02], dev)
    J, g_tw, g_u = A.gradient(model, tw0, u0, zeta0, hg_t, hgd_t)
    h = 1e-6
    worst = 0.0
    for _ in range(8):
        dtw = _t(rng.normal(size=n + 1), dev)
        du = _t(rng.normal(size=3), dev)
        with torch.no_grad():
            jp = model.objective(tw0 + h * dtw, u0 + h * du, zeta0, hg_t, hgd_t)
            jm = model.objective(tw0 - h * dtw, u0 - h * du, zeta0, hg_t, hgd_t)
        fd = float((jp - jm) / (2 * h))
        an = float((g_tw * dtw).sum() + (g_u * du).sum())
        err = abs(fd - an) / max(abs(an), 1e-12)
        worst = max(worst, err)
    print(f"adjoint vs central FD: worst relative error over 8 directions = {worst:.2e}")
    assert worst < 1e-5
