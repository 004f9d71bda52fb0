"""GPU parity of the Biot-Savart kernels and pair sums (K1) against the FP64 oracle.

Tolerances: FP64 path 1e-12 relative (norm-wise) for pair sums; single pairs
1e-13.  The mixed form reports its error and is held to 1e-5.
"""

import numpy as np
import pytest
import torch

from oracle import uvlm_oracle as O
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _t(x, dev):
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)


def test_reference_kernel_cases(dev):
    """tests/algebra/test_kernels.py:9-98 of the reference, on the device kernels."""
    from flapjax_amd.utils import R_CUTOFF, biot_savart, biot_savart_cutoff, biot_savart_epsilon

    y = np.array([[0.0, 1.0, 0.0], [0.0, -1.0, 0.0]])
    # plain: NaN at the midpoint and on the collinear extension, finite off the line
    assert torch.isnan(biot_savart(_t([0, 0, 0], dev), _t(y, dev))).all()
    assert torch.isnan(biot_savart(_t([0, 1, 0], dev), _t([[0, 0, 0], [0, -1, 0]], dev))).all()
    assert not torch.isnan(biot_savart(_t([0.1, 0, 0], dev), _t(y, dev))).all()
    # epsilon: zero at the midpoint and collinear points, ~plain off the line
    assert torch.allclose(biot_savart_epsilon(_t([0, 0, 0], dev), _t(y, dev)), torch.zeros(3, device=dev, dtype=torch.float64))
    assert torch.allclose(biot_savart_epsilon(_t([0, 1, 0], dev), _t([[0, 0, 0], [0, -1, 0]], dev)),
                          torch.zeros(3, device=dev, dtype=torch.float64))
    bs = biot_savart(_t([0.1, 0, 0], dev), _t(y, dev))
    bse = biot_savart_epsilon(_t([0.1, 0, 0], dev), _t(y, dev))
    assert torch.allclose(bs, bse, rtol=5e-5)
    # cutoff
    assert torch.allclose(biot_savart_cutoff(_t([0, 0, 0], dev), _t(y, dev)), torch.zeros(3, device=dev, dtype=torch.float64))
    x = _t([1.01 * R_CUTOFF, 0, 0], dev)
    assert torch.allclose(biot_savart(x, _t(y, dev)), biot_savart_cutoff(x, _t(y, dev)))
    assert torch.allclose(biot_savart_cutoff(_t([0.99 * R_CUTOFF, 0, 0], dev), _t(y, dev)),
                          torch.zeros(3, device=dev, dtype=torch.float64))


@pytest.mark.parametrize("kernel", ["epsilon", "cutoff", "plain"])
def test_single_pairs_match_oracle(dev, kernel):
    from flapjax_amd import _kernels as K
    from flapjax_amd import _lib as L

    rng = np.random.default_rng(11)
    ns, nt = 300, 500
    y0 = rng.normal(size=(ns, 3))
    y1 = y0 + rng.normal(size=(ns, 3)) * 0.3
    g = rng.uniform(-1, 1, ns)
    x = rng.normal(size=(nt, 3)) * 2.0
    kind = {"epsilon": L.KERNEL_EPSILON, "cutoff": L.KERNEL_CUTOFF, "plain": L.KERNEL_PLAIN}[kernel]
    seg = K.segments_from_endpoints(_t(np.stack((y0, y1), 1), dev), _t(g, dev), kind, O.EPSILON)
    v = K.pair_sum(seg, _t(x, dev), kind, O.EPSILON, O.R_CUTOFF).cpu().numpy()
    ref = O.pair_sum(x, y0, y1, g, kernel)
    assert rel_err(v, ref) < 1e-12


def test_pair_sum_random_4096x8192(dev):
    from flapjax_amd import _kernels as K
    from flapjax_amd import _lib as L

    rng = np.random.default_rng(0)
    nt, ns = 4096, 8192
    y0 = rng.uniform(-1, 1, (ns, 3))
    y1 = y0 + rng.normal(size=(ns, 3)) * 0.05
    g = rng.uniform(-0.1, 0.1, ns)
    x = rng.uniform(-1, 1, (nt, 3))
    seg = K.segments_from_endpoints(_t(np.stack((y0, y1), 1), dev), _t(g, dev), L.KERNEL_EPSILON, O.EPSILON)
    v = K.pair_sum(seg, _t(x, dev), L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF).cpu().numpy()
    ref = O.pair_sum(x, y0, y1, g, "epsilon")
    assert rel_err(v, ref) < 1e-12
    # the mixed form: FP32 far field, FP64 near field
    vm = K.pair_sum(seg, _t(x, dev), L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF, L.PREC_MIXED).cpu().numpy()
    err = rel_err(vm, ref)
    print(f"mixed-form relative error (random 4096x8192): {err:.3e}")
    assert err < 1e-5


def _cfg0_lattice():
    m, n, m_star = 8, 32, 20
    zeta0 = O.make_rectangular_grid(m, n, 1.0, 0.25)
    hg = O.rigid_wing_hg(n, 10.0, np.deg2rad(5.0))
    zb = O.vect_product(hg, zeta0)
    s = np.concatenate(([0.0], np.cumsum(np.ones(m_star) * 0.125)))
    zw = zb[-1][None] + np.outer(s, [1.0, 0.0, 0.0])[:, None, :]
    rng = np.random.default_rng(3)
    gb = rng.uniform(-0.1, 0.1, (m, n))
    gw = np.broadcast_to(gb[-1], (m_star, n)).copy()
    return zb, zw, gb, gw


def test_compute_v_ind_cfg0_lattice(dev):
    from flapjax_amd.aic import compute_v_ind
    from flapjax_amd.utils import biot_savart_epsilon, compute_surf_c

    zb, zw, gb, gw = _cfg0_lattice()
    c = O.compute_surf_c(zb)
    ref = O.compute_v_ind(c, [zb, zw], [gb, gw], "epsilon")
    v = compute_v_ind(_t(c, dev), [_t(zb, dev), _t(zw, dev)], [_t(gb, dev), _t(gw, dev)],
                      [biot_savart_epsilon, biot_savart_epsilon]).cpu().numpy()
    assert rel_err(v, ref) < 1e-12
    # the device collocation points equal the oracle's bit for bit
    assert np.array_equal(compute_surf_c(_t(zb, dev)).cpu().numpy(), c)
    # with a mirror plane
    mp, mn = np.zeros(3), np.array([0.0, 1.0, 0.0])
    ref_m = O.compute_v_ind(c, [zb, zw], [gb, gw], "epsilon", mp, mn)
    v_m = compute_v_ind(_t(c, dev), [_t(zb, dev), _t(zw, dev)], [_t(gb, dev), _t(gw, dev)],
                        [biot_savart_epsilon] * 2, mirror_point=mp, mirror_normal=mn).cpu().numpy()
    assert rel_err(v_m, ref_m) < 1e-12


def test_empty_and_ragged(dev):
    from flapjax_amd import _kernels as K
    from flapjax_amd import _lib as L
    from flapjax_amd.aic import compute_v_ind

    # no filaments -> zeros; one target; 65 targets (ragged tile)
    for nt in (1, 65, 130):
        x = np.random.default_rng(nt).normal(size=(nt, 3))
        v = K.pair_sum(None, _t(x, dev), L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF).cpu().numpy()
        assert np.all(v == 0.0)
    # a wake with zero rows contributes nothing
    zb, zw, gb, gw = _cfg0_lattice()
    c = O.compute_surf_c(zb)
    v0 = compute_v_ind(_t(c, dev), [_t(zb, dev), _t(zw[:1], dev)], [_t(gb, dev), _t(gw[:0], dev)], None)
    v1 = compute_v_ind(_t(c, dev), [_t(zb, dev)], [_t(gb, dev)], None)
    assert torch.equal(v0, v1)
    # ragged filament counts around the chunk boundaries
    rng = np.random.default_rng(5)
    for ns in (1, 3, 31, 33, 127, 1000):
        y0 = rng.normal(size=(ns, 3))
        y1 = y0 + 0.1 * rng.normal(size=(ns, 3))
        g = rng.normal(size=ns)
        x = rng.normal(size=(77, 3))
        seg = K.segments_from_endpoints(_t(np.stack((y0, y1), 1), dev), _t(g, dev), L.KERNEL_EPSILON, O.EPSILON)
        v = K.pair_sum(seg, _t(x, dev), L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF).cpu().numpy()
        assert rel_err(v, O.pair_sum(x, y0, y1, g)) < 1e-12


def test_convect_epilogue(dev):
    """FJ_OUT_CONVECT: x + dt (U + v) equals the oracle's propagate step."""
    from flapjax_amd import _kernels as K
    from flapjax_amd import _lib as L
    from flapjax_amd.flowfields import Constant

    zb, zw, gb, gw = _cfg0_lattice()
    y0, y1, g = O.surface_segments(zb, gb)
    x = zw[:-1].reshape(-1, 3)
    ff = Constant(np.array([1.0, 0.0, 0.05]), 1.225, True)
    seg = K.segments_from_grid(_t(zb, dev), _t(gb, dev), L.KERNEL_EPSILON, O.EPSILON)
    out = K.pair_sum(seg, _t(x, dev), L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF, mode=L.OUT_CONVECT, flow=ff,
                     time=0.3, dt=0.0625).cpu().numpy()
    ref = x + 0.0625 * (np.array([1.0, 0.0, 0.05]) + O.pair_sum(x, y0, y1, g))
    assert rel_err(out, ref) < 1e-14


def _lattice_case(dev, kind, rng):
    """Bound-like and wake-like grids (ring rows not a multiple of the strip
    height), a mirrored copy, an empty wake, and targets on and off the
    vertices."""
    from flapjax_amd import _kernels as K

    grids, lat, off = [], [], 0
    shapes = [(7, 13, False), (9, 13, False), (0, 13, False), (7, 13, True), (2, 5, False), (1, 1, False)]
    verts = []
    for m, n, mir in shapes:
        z = O.make_rectangular_grid(m, n, 1.0, 0.25)
        z[..., 1] = np.linspace(0.0, 3.0, n + 1)[None, :]
        z += rng.normal(size=z.shape) * 0.02
        z[..., 0] += 1.5 * (len(grids) % 2)
        verts.append(z.reshape(-1, 3))
        zt = _t(z, dev)
        gt = _t(rng.uniform(-1, 1, (m, n)), dev) if m * n else None
        grids.append((zt, gt, mir))
        lat.append((off, m, n))
        off += m * (n + 1) + (m + 1) * n
    mirror = (np.array([0.0, -0.5, 0.0]), np.array([0.0, 1.0, 0.0]))
    seg = K.segments_from_grids(grids, kind, O.EPSILON, mirror=mirror)
    pts = rng.normal(size=(700, 3)) * 1.5
    if kind != 2:  # the un-regularised kernel is singular on its own vertices
        pts = np.concatenate([pts] + verts)
    return seg, lat, pts


@pytest.mark.parametrize("kernel", ["epsilon", "cutoff", "plain"])
def test_lattice_pair_sum_equals_table_pair_sum(dev, kernel):
    """fj_pair_sum_lattices (per-vertex terms shared between filaments) gives
    the table pair sum of the same filaments, and the oracle's, to 1e-12."""
    from flapjax_amd import _kernels as K
    from flapjax_amd import _lib as L

    kind = {"epsilon": L.KERNEL_EPSILON, "cutoff": L.KERNEL_CUTOFF, "plain": L.KERNEL_PLAIN}[kernel]
    rng = np.random.default_rng(21)
    seg, lat, pts = _lattice_case(dev, kind, rng)
    x = _t(pts, dev)
    v_lat = K.pair_sum_lattices(seg, lat, x, kind, O.EPSILON, O.R_CUTOFF).cpu().numpy()
    v_tab = K.pair_sum(seg, x, kind, O.EPSILON, O.R_CUTOFF).cpu().numpy()
    s = seg.cpu().numpy()
    ref = O.pair_sum(pts, s[:, 0:3], s[:, 3:6], s[:, 10], kernel)
    assert np.isfinite(v_lat).all()
    assert rel_err(v_lat, v_tab) < 1e-12
    assert rel_err(v_lat, ref) < 1e-12


def test_lattice_pair_sum_modes_and_sizes(dev):
    """Epilogue modes (store, add, convect) and launch plans from one target
    to many column blocks."""
    import flapjax_amd as fj
    from flapjax_amd import _kernels as K
    from flapjax_amd import _lib as L

    rng = np.random.default_rng(22)
    seg, lat, pts = _lattice_case(dev, L.KERNEL_EPSILON, rng)
    flow = fj.Constant(np.array([1.0, 0.0, 0.2]), 1.225, True)
    for nt in (1, 63, 65, 4000):
        x = _t(pts[:nt] if nt <= len(pts) else rng.normal(size=(nt, 3)), dev)
        ref = K.pair_sum(seg, x, L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF)
        got = K.pair_sum_lattices(seg, lat, x, L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF)
        assert rel_err(got.cpu().numpy(), ref.cpu().numpy()) < 1e-12
        acc = torch.ones_like(x)
        K.pair_sum_lattices(seg, lat, x, L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF, mode=L.OUT_ADD, out=acc)
        assert rel_err((acc - 1.0).cpu().numpy(), ref.cpu().numpy()) < 1e-12
        cv = K.pair_sum_lattices(seg, lat, x, L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF, mode=L.OUT_CONVECT,
                                 flow=flow, time=0.3, dt=0.05)
        cref = K.pair_sum(seg, x, L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF, mode=L.OUT_CONVECT, flow=flow,
                          time=0.3, dt=0.05)
        assert rel_err(cv.cpu().numpy(), cref.cpu().numpy()) < 1e-14
    # only empty lattices: zero velocity / pure convection
    x = _t(pts[:10], dev)
    z = K.pair_sum_lattices(seg, [(lat[2][0], 0, 13)], x, L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF)
    assert (z == 0).all()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_split_pair_sums(dev, precision):
    """fj_pair_sum_lattices_split / fj_pair_sum_split: two target ranges with
    their own epilogues (store, add, convect) in one launch equal two
    separate pair sums, for splits at 0, inside a 64-target tile and at the
    end, in the lattice (FP64) and filament-table (FP64 and mixed) forms."""
    import flapjax_amd as fj
    from flapjax_amd import _kernels as K
    from flapjax_amd import _lib as L

    prec = L.PREC_FP64 if precision == "fp64" else L.PREC_MIXED
    tol = 1e-13 if precision == "fp64" else 1e-6
    rng = np.random.default_rng(23)
    seg, lat, pts = _lattice_case(dev, L.KERNEL_EPSILON, rng)
    flow = fj.Constant(np.array([1.0, 0.0, 0.2]), 1.225, True)
    x = _t(pts[:300], dev)
    ref_s = K.pair_sum(seg, x, L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF, prec)
    ref_c = K.pair_sum(seg, x, L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF, prec, L.OUT_CONVECT, flow=flow, time=0.3,
                       dt=0.05)
    lattice_forms = [lat, None] if precision == "fp64" else [None]
    for lats in lattice_forms:
        for n_a in (0, 37, 200, 300):
            out_a = torch.ones((n_a, 3), dtype=torch.float64, device=dev)
            out_b = torch.empty((300 - n_a, 3), dtype=torch.float64, device=dev)
            K.pair_sum_split(seg, x, n_a, L.KERNEL_EPSILON, O.EPSILON, O.R_CUTOFF, out_a, L.OUT_ADD, out_b,
                             L.OUT_CONVECT, prec, flow=flow, time=0.3, dt=0.05, lattices=lats)
            if n_a:
                assert rel_err((out_a - 1.0).cpu().numpy(), ref_s[:n_a].cpu().numpy()) < tol, (lats is None, n_a)
            if n_a < 300:
                assert rel_err(out_b.cpu().numpy(), ref_c[n_a:].cpu().numpy()) < tol, (lats is None, n_a)
