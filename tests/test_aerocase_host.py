"""Host logic of the AeroCase accessors (aero/data_structures.py:240-615) on
CPU tensors: time-step selection, snapshot/batched errors, views and copies."""

import numpy as np
import pytest
import torch

from flapjax_amd.data_structures import AeroCase, GridDiscretisation


def _case(n_tstep, seed=0):
    g = torch.Generator().manual_seed(seed)
    gd = (GridDiscretisation(2, 3, 4), GridDiscretisation(1, 2, 0))
    c = AeroCase.zeros(gd, n_tstep, "cpu")
    for name in ("gamma_b", "gamma_w", "zeta_b", "zeta_w", "f_steady", "alpha", "gamma_b_dot"):
        for a in getattr(c, name):
            a.copy_(torch.rand(a.shape, generator=g, dtype=torch.float64))
    c.t[:] = np.arange(n_tstep) * 0.1
    return c


def test_snapshot_and_batched_step_selection():
    s, b = _case(1), _case(5)
    assert not s.is_batched and b.is_batched and b.n_surf == 2
    with pytest.raises(ValueError, match="i_ts must be provided"):
        b.gamma_full()
    assert [x.shape for x in s.gamma_full()] == [(2, 3), (1, 2), (4, 3), (0, 2)]
    z = b.zeta_full(3)
    assert torch.equal(z[0], b.zeta_b[0][3]) and torch.equal(z[2], b.zeta_w[0][3])
    st = b.get_states(-1)
    assert torch.equal(st.gamma_b_dot[1], b.gamma_b_dot[1][4])
    with pytest.raises(IndexError):
        b.get_states(5)


def test_getitem_setitem_round_trip():
    b, other = _case(5, 1), _case(5, 2)
    snap = other[2]
    assert not snap.is_batched and snap.n_tstep == 1
    assert snap.gamma_b[0].data_ptr() == other.gamma_b[0][2].data_ptr()  # a view
    b[4] = snap
    for name in ("gamma_b", "zeta_w", "f_steady", "alpha"):
        for x, y in zip(getattr(b, name), getattr(other, name)):
            assert torch.equal(x[4], y[2])
    assert b.t[4] == other.t[2]
    with pytest.raises(TypeError):
        snap[0] = snap
    with pytest.raises(TypeError):
        b[1] = other


def test_surface_views_and_set_at_ts():
    b = _case(3)
    sp = b.get_surf_snapshot(1, 0)
    assert sp.i_ts == 1 and torch.equal(sp.gamma_w, b.gamma_w[0][1]) and sp.surf_b_name == "surf_0_bound"
    with pytest.raises(TypeError):
        b.get_surface(0)
    s = _case(1)
    assert torch.equal(s.get_surface(1).zeta_b, s.zeta_b[1][0])
    vals = [torch.full((2, 3), 7.0, dtype=torch.float64), torch.full((1, 2), -1.0, dtype=torch.float64)]
    b.set_arraylist_at_ts("gamma_b", vals, 2)
    assert torch.all(b.gamma_b[0][2] == 7.0) and torch.all(b.gamma_b[1][2] == -1.0)
    init = AeroCase.initialise(s, 4)
    assert init.n_tstep == 4 and torch.equal(init.gamma_b[0][0], s.gamma_b[0][0])


def test_v_ind_needs_kernels():
    with pytest.raises(ValueError, match="kernels"):
        _case(1).get_v_ind(torch.zeros((2, 3), dtype=torch.float64))
