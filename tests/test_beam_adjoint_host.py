"""Differentiable beam residual (flapjax_amd/beam_adjoint.py) on the CPU.

* The torch residual, static and dynamic, equals the oracle's residual
  (oracle/beam_oracle.py, which restates structure/beam.py:1749-1843) and the
  torch SE(3) maps equal the oracle's on both sides of the small-angle
  switches.
* The implicit-derivative Functions: with a Newton solve written here in
  torch standing in for the device solve, their gradients match central
  finite differences of the solve in every input."""

import numpy as np
import pytest
import torch

from flapjax_amd import beam_adjoint as BA
from flapjax_amd.beam import TimeIntegrator
from oracle import beam_oracle as B

F64 = torch.float64


def _beam(n=6, seed=0):
    rng = np.random.default_rng(seed)
    b = B.BeamOracle(n, B.chain(n), np.array((0.0, 0.0, 1.0)), k_cs_index=np.arange(n - 1) // 2)
    coords = np.zeros((n, 3))
    coords[:, 1] = np.linspace(0.0, 3.0, n)
    k = np.stack([np.diag(rng.uniform(1.0, 3.0, 6) * np.array((1e4, 1e4, 1e4, 1e2, 1e2, 1e2)))
                  for _ in range((n - 2) // 2 + 1)])
    m = np.diag(rng.uniform(1.0, 2.0, 6))
    b.set_design_variables(coords, k, m)
    tb = BA.BeamTables(n, b.conn, b.o0, b.l0, b.d0, b.k_cs, b.m_cs, b.k_cs_index, b.m_cs_index, "cpu")
    return b, tb, rng


@pytest.mark.parametrize("scale", [1e-8, 1e-3, 0.4, 2.5])
def test_maps_match_oracle(scale):
    rng = np.random.default_rng(1)
    for _ in range(4):
        h = rng.normal(size=6) * scale
        ht = torch.as_tensor(h)
        assert np.allclose(BA.exp_se3(ht).numpy(), B.exp_se3(h), rtol=0, atol=1e-15)
        assert np.allclose(BA.t_se3(ht).numpy(), B.t_se3(h), rtol=0, atol=1e-14)
        assert np.allclose(BA.t_inv_se3(ht).numpy(), B.t_inv_se3(h), rtol=0, atol=1e-14)
        g = B.exp_se3(h)
        assert np.allclose(BA.log_se3(torch.as_tensor(g)).numpy(), B.log_se3(g), rtol=0, atol=1e-13)
        for s in (0.25, 0.5):
            assert np.allclose(BA.q_mat(s, ht, torch.eye(6, dtype=F64)).numpy(), B.q_mat(s, h, np.eye(6)), atol=1e-14)


def test_residual_matches_oracle():
    b, tb, rng = _beam(7)
    n = b.n_nodes
    hg = np.stack([b.hg0[i] @ B.exp_se3(rng.normal(size=6) * 0.2) for i in range(n)])
    v, vd, fdd, ffol = (rng.normal(size=(n, 6)) for _ in range(4))
    sd = np.arange(6 * n)
    t = torch.as_tensor
    ref = b.residual(hg, sd, ffol, fdd)[0]
    got = BA.beam_residual(tb, t(hg), None, t(ffol), t(fdd)).numpy()
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    ref = b.residual(hg, sd, ffol, fdd, True, v, vd)[0]
    got = BA.beam_residual(tb, t(hg), None, t(ffol), t(fdd), t(v), t(vd)).numpy()
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    ks = rng.uniform(0.5, 1.5, b.k_cs.shape[0])
    b2, tb2, _ = _beam(7)
    b2.set_design_variables(b.x0, b.k_cs * ks[:, None, None], b.m_cs)
    ref = b2.residual(hg, sd, None, fdd)[0]
    got = BA.beam_residual(tb, t(hg), t(ks), None, t(fdd)).numpy()
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


class _TorchSolver(BA.BeamSolver):
    """BeamSolver with the device Newton replaced by a torch Newton on the
    same residual (exact Jacobian), for checking the backward passes."""

    def __init__(self, tb, hg0, prescribed, ti=None):
        self.tables, self.hg0, self.ti = tb, hg0, ti
        n6 = 6 * tb.n_nodes
        sd = np.array(sorted(set(range(n6)) - set(prescribed)))
        self.sd = self.sd_long = torch.as_tensor(sd)
        self.n_s = len(sd)
        self.iters = []

    @staticmethod
    def _newton(res, x):
        for _ in range(12):
            dx = torch.linalg.solve(torch.func.jacrev(res)(x), res(x))
            x = x - dx
            if float(dx.abs().max()) <= 1e-15 * (1.0 + float(x.abs().max())):
                break
        return x

    def static(self, f_dead, k_scale):
        res = lambda xs: self.reduce(BA.beam_residual(self.tables, self.hg0 @ BA.exp_se3(self.expand(xs, f_dead)),  # noqa: E731
                                                      k_scale, None, f_dead))
        xs = self._newton(res, torch.zeros(self.n_s, dtype=F64))
        return self.hg0 @ BA.exp_se3(self.expand(xs, f_dead))

    def dynamic(self, hg_prev, phi, v, v_dot, f_dead, k_scale):
        res = lambda xs: BA.DynamicBeamSolve.residual(self, xs, hg_prev, phi, v, v_dot, f_dead, k_scale)  # noqa: E731
        xs = self._newton(res, self.reduce(phi))
        return BA.DynamicBeamSolve.full_state(self, xs, phi, v, v_dot)[0]


def _fd_check(fun, args, idx, rng, h=1e-6, tol=1e-6, manifold=False):
    """Directional derivatives of fun against the fourth-order central
    difference (-f(2h) + 8 f(h) - 8 f(-h) + f(-2h)) / 12h along random
    directions; with manifold, argument 0 (node frames) moves as
    hg exp(s d)."""
    leaves = [a.clone().requires_grad_(i in idx) for i, a in enumerate(args)]
    val = fun(*leaves)
    grads = torch.autograd.grad(val, [leaves[i] for i in idx])
    for gi, i in zip(grads, idx):
        on_manifold = manifold and i == 0
        d = torch.as_tensor(rng.normal(size=tuple(args[i].shape[:-2]) + (6,) if on_manifold else args[i].shape))

        def at(sc):
            a = [x.clone() for x in args]
            a[i] = args[i] @ BA.exp_se3(sc * d) if on_manifold else args[i] + sc * d
            return fun(*a)

        hi = h if on_manifold else h * float(args[i].pow(2).mean().sqrt())  # step relative to the input's size
        fd = (-at(2 * hi) + 8 * at(hi) - 8 * at(-hi) + at(-2 * hi)) / (12 * hi)
        if on_manifold:
            dm = torch.func.jvp(lambda e: args[i] @ BA.exp_se3(e * d), (torch.zeros((), dtype=F64),),
                                (torch.ones((), dtype=F64),))[1]
            an = (gi * dm).sum()
        else:
            an = (gi * d).sum()
        assert abs(float(fd - an)) <= tol * max(1.0, abs(float(an))), (i, float(fd), float(an))


def test_coloured_jacobian_equals_dense():
    """The coloured forward-mode Jacobian equals torch.func.jacrev."""
    b, tb, rng = _beam(6)
    n = b.n_nodes
    ti = TimeIntegrator(0.9, 0.01)
    hgp = torch.as_tensor(np.stack([b.hg0[i] @ B.exp_se3(rng.normal(size=6) * 0.1) for i in range(n)]))
    sv = _TorchSolver(tb, hgp, range(6), ti)
    phi0, v0, vd0, f = (torch.as_tensor(rng.normal(size=(n, 6)) * s) for s in (0.05, 0.5, 5.0, 20.0))
    ks = torch.as_tensor(rng.uniform(0.8, 1.2, b.k_cs.shape[0]))

    def res(xs):
        return BA.DynamicBeamSolve.residual(sv, sv.reduce(phi0) + xs, hgp, phi0, v0, vd0, f, ks)

    x = torch.zeros(sv.n_s, dtype=F64)
    jc = sv.jacobian(res, x)
    jd = torch.func.jacrev(res)(x)
    assert torch.allclose(jc, jd, rtol=1e-12, atol=1e-12 * float(jd.abs().max()))
    assert sv._colour[1].shape[1] == 18  # a chain needs three node colours


def test_static_solve_gradient():
    b, tb, rng = _beam(5)
    n = b.n_nodes
    hg0 = torch.as_tensor(b.hg0)
    sv = _TorchSolver(tb, hg0, range(6))
    f = torch.as_tensor(rng.normal(size=(n, 6)) * 20.0)
    ks = torch.as_tensor(rng.uniform(0.8, 1.2, b.k_cs.shape[0]))
    w = torch.as_tensor(rng.normal(size=(n, 4, 4)))

    def fun(f_, k_):
        return (BA.StaticBeamSolve.apply(f_, k_, sv) * w).sum()

    _fd_check(fun, [f, ks], [0, 1], rng)


def test_dynamic_solve_gradient():
    b, tb, rng = _beam(4)
    n = b.n_nodes
    ti = TimeIntegrator(0.9, 0.01)
    hgp = torch.as_tensor(np.stack([b.hg0[i] @ B.exp_se3(rng.normal(size=6) * 0.05) for i in range(n)]))
    hgp[0] = torch.as_tensor(b.hg0[0])
    sv = _TorchSolver(tb, hgp, range(6), ti)
    phi0, v0, vd0 = (torch.as_tensor(rng.normal(size=(n, 6)) * s) for s in (0.01, 0.5, 5.0))
    f = torch.as_tensor(rng.normal(size=(n, 6)) * 20.0)
    ks = torch.as_tensor(rng.uniform(0.8, 1.2, b.k_cs.shape[0]))
    w = [torch.as_tensor(rng.normal(size=(n, 6))) for _ in range(3)]

    def fun(hg_, p_, v_, vd_, f_, k_):
        phi_s = BA.DynamicBeamSolve.apply(hg_, p_, v_, vd_, f_, k_, sv)
        out = BA.DynamicBeamSolve.full_state(sv, phi_s, p_, v_, vd_)
        return sum((o * wi).sum() for o, wi in zip(out, w))

    # the step's accelerations scale rounding in phi by beta' ~ 1/dt^2, so
    # the differences take a wider step (fourth-order stencil)
    _fd_check(fun, [hgp, phi0, v0, vd0, f, ks], [0, 1, 2, 3, 4, 5], rng, h=1e-3, manifold=True)


@pytest.mark.parametrize("relative", [True, False])
def test_flow_expression_matches_oracle(relative):
    """The torch flowfield of the coupled adjoint (coupled_adjoint.flow_at)
    equals the oracle's 1-cos gust (aero/flowfields.py:101-181) and constant
    flow at points inside, outside and on both sides of the gust."""
    from flapjax_amd.coupled_adjoint import flow_at
    from flapjax_amd.flowfields import Constant, OneMinusCosine
    from oracle import uvlm_oracle as U

    rng = np.random.default_rng(3)
    u = np.array((10.0, 0.0, 1.0))
    x = rng.uniform(-8.0, 8.0, size=(200, 3))
    for t in (0.0, 0.3, 0.9):
        g = OneMinusCosine(u, 1.225, relative, 2.5, 0.7, gust_x0=(-4.0, 0.0, 0.0))
        og = U.OneMinusCosine(u, 1.225, relative, gust_length=2.5, gust_amplitude=0.7,
                              gust_x0=np.array((-4.0, 0.0, 0.0)))
        assert np.allclose(flow_at(g, torch.as_tensor(x), t).numpy(), og(x, t), rtol=0, atol=1e-14)
        c = Constant(u, 1.225, relative)
        assert np.allclose(flow_at(c, torch.as_tensor(x), t).numpy(), U.Constant(u, 1.225, relative)(x, t))
