"""Host logic of the serving tick (no device work): the case records a tick
hands to fj_step_batched follow each case's (previous, current) state
buffers exactly as swapping the buffers every step would, and the final
state read back is the last step's."""

from types import SimpleNamespace

import numpy as np
import pytest
import torch


def _fake_slot(m, n, ms, steps, i0):
    f = lambda *shape: torch.zeros(shape, dtype=torch.float64)  # noqa: E731
    sl = SimpleNamespace(req=SimpleNamespace(m=m, n=n, m_star=ms, steps=steps), dt=1.0 / m, i=i0,
                         dof=torch.arange(n + 1, dtype=torch.int32), zeta0=f(m + 1, n + 1, 3),
                         zbd=f(m + 1, n + 1, 3), A=f(m * n, m * n), Ainv=f(m * n, m * n), gbd=f(m, n),
                         fs=f(m + 1, n + 1, 3), ws=f(64), hg=f(steps, n + 1, 4, 4), hgd=f(steps, n + 1, 4, 4),
                         lift=f(steps))
    zb, zbp, gb, gbp = f(m + 1, n + 1, 3), f(m + 1, n + 1, 3), f(m, n), f(m, n)
    zw, zwn, gw, gwn = f(ms + 1, n + 1, 3), f(ms + 1, n + 1, 3), f(ms, n), f(ms, n)
    sl.zb_pair, sl.gb_pair, sl.zw_pair, sl.gw_pair = (zb, zbp), (gb, gbp), (zw, zwn), (gw, gwn)
    # the per-step swaps the records must reproduce, starting from step i0
    sl.sim = dict(zb=zb, zb_prev=zbp, gb=gb, gb_prev=gbp, zw=zw, zw_next=zwn, gw=gw, gw_next=gwn)
    for _ in range(i0):
        _swap_before(sl.sim)
        _swap_after(sl.sim)
    return sl


def _swap_before(s):
    s["zb_prev"], s["zb"] = s["zb"], s["zb_prev"]
    s["gb_prev"], s["gb"] = s["gb"], s["gb_prev"]


def _swap_after(s):
    s["zw"], s["zw_next"] = s["zw_next"], s["zw"]
    s["gw"], s["gw_next"] = s["gw_next"], s["gw"]


def test_tick_records_follow_the_state_buffer_pairs(monkeypatch):
    from flapjax_amd import _lib as L
    from flapjax_amd.serving import ServingEngine, _Slot

    if not L.LIB_PATH.exists():
        pytest.skip("library not built (run python -m flapjax_amd.build)")
    monkeypatch.setattr(L, "lib", L.load)  # size queries only: host functions
    eng = ServingEngine(device=torch.device("cpu"))
    slots = [_fake_slot(4, 8, 5, 20, 0), _fake_slot(6, 12, 3, 20, 3), _fake_slot(2, 5, 0, 20, 8)]
    plan = eng._tick_plan(slots)
    assert plan["ws"] > 0 and plan["table"] > 0
    for _ in range(4):
        rec = eng._fill_records(plan, slots)
        for k, sl in enumerate(slots):
            s = sl.sim
            _swap_before(s)
            i = sl.i + 1
            r = rec[k]
            assert (r["m"], r["n"], r["m_star"]) == (sl.req.m, sl.req.n, sl.req.m_star)
            assert r["time"] == i * sl.dt
            assert r["hg"] == sl.hg[i].data_ptr() and r["hg_dot"] == sl.hgd[i].data_ptr()
            assert r["lift"] == sl.lift.data_ptr() + 8 * i
            assert r["zeta_b"] == s["zb"].data_ptr() and r["zeta_b_nm1"] == s["zb_prev"].data_ptr()
            assert r["gamma_b"] == s["gb"].data_ptr() and r["gamma_b_nm1"] == s["gb_prev"].data_ptr()
            assert r["zeta_w_nm1"] == s["zw"].data_ptr() and r["zeta_w"] == s["zw_next"].data_ptr()
            assert r["gamma_w_nm1"] == s["gw"].data_ptr() and r["gamma_w"] == s["gw_next"].data_ptr()
            assert r["aic"] == sl.A.data_ptr() and r["aic_inv"] == sl.Ainv.data_ptr()
            _swap_after(s)
            sl.i += 1
            gb, zw = _Slot.final_state(sl)
            assert gb.data_ptr() == s["gb"].data_ptr() and zw.data_ptr() == s["zw"].data_ptr()
    # one stop-test word per case, consecutive
    words = rec["iters"].astype(np.int64)
    assert np.all(np.diff(words) == 4)
