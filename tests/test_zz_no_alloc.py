"""fj_pair_sum_lattices_noalloc (the allocation-free entry point, fixed device
scratch) against fj_pair_sum_lattices (caller workspace), both called
through the C-ABI: bit-identical results on two calls in a row, the second
with a smaller input, and a clear error for a call beyond the scratch size.
"""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _wing_grids(dev, m, n, m_star, seed):
    """A bound lattice and a jittered free-wake lattice with random ring
    strengths, and the wake vertices as targets (a rollup)."""
    rng = np.random.default_rng(seed)
    x = np.linspace(0.0, 1.0, m + 1) - 0.25
    y = np.linspace(0.0, 10.0, n + 1)
    zb = np.zeros((m + 1, n + 1, 3))
    zb[..., 0], zb[..., 1] = x[:, None], y[None, :]
    zb[..., 2] = -0.08 * zb[..., 0]
    s = np.arange(m_star + 1) / m
    zw = zb[-1][None] + np.stack([s, 0 * s, 0 * s], -1)[:, None, :]
    zw[1:, :, 2] += 0.01 * rng.normal(size=(m_star, n + 1))
    gb = 0.1 * rng.normal(size=(m, n))
    gw = 0.1 * rng.normal(size=(m_star, n))
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)  # noqa: E731
    return [(t(zb), t(gb), False), (t(zw), t(gw), False)], t(zw[1:].reshape(-1, 3))


def _both(dev, m, n, m_star, seed, mode):
    from flapjax_amd import _kernels as K
    from flapjax_amd import _lib as L

    lib = L.lib()
    grids, tgt = _wing_grids(dev, m, n, m_star, seed)
    seg = K.segments_from_grids(grids, L.KERNEL_EPSILON, 1e-7)
    lats = K.lattices_of(grids)
    refs = K.lattice_refs(lats)
    nt = tgt.shape[0]
    ws = torch.empty(lib.fj_pair_sum_lattices_workspace_bytes(nt, len(lats), refs) // 8 + 1, dtype=torch.float64,
                     device=dev)
    assert ws.numel() * 8 <= L.NOALLOC_SCRATCH_BYTES + 8
    fd = K.flow_desc(__import__("flapjax_amd").Constant(np.array([1.0, 0.0, 0.0]), 1.225, True))
    a = torch.full_like(tgt, 0.5) if mode == L.OUT_ADD else torch.empty_like(tgt)
    b = a.clone()
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    args = (p(seg), len(lats), refs, p(tgt), nt)
    tail = (L.KERNEL_EPSILON, 1e-7, 1e-4, mode, ctypes.byref(fd), 0.3, 0.0625)
    rc = lib.fj_pair_sum_lattices(*args, p(a), *tail, p(ws), L.stream())
    assert rc == 0
    rc = lib.fj_pair_sum_lattices_noalloc(*args, p(b), *tail, L.stream())
    assert rc == 0
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_noalloc_equals_workspace_entry_two_calls(dev, mode):
    big = _both(dev, 16, 64, 120, 1, mode)
    small = _both(dev, 4, 12, 9, 2, mode)
    for a, b in (big, small):
        assert np.isfinite(a).all()
        assert np.array_equal(a, b)


def test_noalloc_rejects_oversized_call(dev):
    from flapjax_amd import _kernels as K
    from flapjax_amd import _lib as L

    lib = L.lib()
    # 2000 x 64 wake lattice with 200k targets: partials alone exceed 32 MiB
    lats = [(0, 2000, 64)]
    refs = K.lattice_refs(lats)
    need = lib.fj_pair_sum_lattices_workspace_bytes(200000, 1, refs)
    assert need > L.NOALLOC_SCRATCH_BYTES
    tgt = torch.zeros((200000, 3), dtype=torch.float64, device=dev)
    out = torch.empty_like(tgt)
    seg = torch.zeros((K.n_segments(2000, 64), L.SEG_STRIDE), dtype=torch.float64, device=dev)
    with pytest.raises(L.FlapjaxLibError, match="scratch"):
        K.pair_sum_lattices_noalloc(seg, lats, tgt, L.KERNEL_EPSILON, 1e-7, 1e-4, L.OUT_STORE, out)
