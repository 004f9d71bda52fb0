"""Control surfaces through UVLM's grid_func (aero/uvlm.py:333-339, 537-621,
1119-1406) on the device against the oracle.

An aileron over the aft third of the chord and the first third of the span,
built with flapjax_amd.utils.add_control_surface (torch) on the device and
oracle.uvlm_oracle.add_control_surface (NumPy) in the oracle; the grid
velocity of a moving aileron comes from torch.func.jvp on the device and a
complex step in the oracle (both pinned in tests/test_oracle_control_surfaces.py).
Bars: circulation and forces within 1e-10, wake vertices within 1e-8
(norm-wise, max |a - b| / max |b|).
"""

import numpy as np
import pytest
import torch

from oracle import uvlm_oracle as O
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

M, N, MS, STEPS, SPAN = 6, 12, 8, 8, 4.0
FLAP_M, FLAP_N = slice(4, None), slice(0, 5)


def _motion():
    dt = 1.0 / M
    hg_t, hgd_t = O.pitch_plunge_motion(N, SPAN, STEPS, dt, np.deg2rad(5.0), np.deg2rad(3.0), 0.05, 0.4)
    t = dt * np.arange(STEPS)
    ang = np.deg2rad(1.0) + np.deg2rad(4.0) * np.sin(2.3 * t)
    vel = np.deg2rad(4.0) * 2.3 * np.cos(2.3 * t)
    return dt, hg_t, hgd_t, ang, vel


def _oracle(free_wake):
    dt, hg_t, hgd_t, ang, vel = _motion()
    zeta0 = O.make_rectangular_grid(M, N, 1.0, 0.25)

    def gf(x0, *, aileron):
        return [O.add_control_surface(x0[0], aileron, FLAP_M, FLAP_N)]

    o = O.UVLMOracle([(M, N, MS)], np.arange(N + 1), free_wake=free_wake, grid_func=gf)
    o.set_design_variables(dt, O.Constant(np.array([1.0, 0.0, 0.1]), 1.225, True), zeta0, hg_t[0],
                           reference_cs_angles={"aileron": ang[0]})
    st = o.solve_static(hg=hg_t[0])
    dyn = o.solve_prescribed_dynamic(st, hg_t, hgd_t, cs_ang_t={"aileron": ang}, cs_vel_t={"aileron": vel})
    return st, dyn


def _device(free_wake, grid_func=True, angles=True):
    import flapjax_amd as fj
    from flapjax_amd.utils import add_control_surface
    from flapjax_amd.uvlm import UVLM

    dt, hg_t, hgd_t, ang, vel = _motion()
    if not angles:
        ang, vel = 0.0 * ang, 0.0 * vel

    def gf(x0, *, aileron):
        return [add_control_surface(x0[0], aileron, FLAP_M, FLAP_N)]

    uv = UVLM([(M, N, MS)], np.arange(N + 1), free_wake=free_wake, grid_func=gf if grid_func else None)
    uv.set_design_variables(dt, fj.Constant(np.array([1.0, 0.0, 0.1]), 1.225, True),
                            O.make_rectangular_grid(M, N, 1.0, 0.25), hg_t[0],
                            reference_cs_angles={"aileron": ang[0]} if grid_func else None)
    st = uv.solve_static(hg=hg_t[0])
    cs = dict(cs_ang_t={"aileron": ang}, cs_vel_t={"aileron": vel}) if grid_func else {}
    dyn = uv.solve_prescribed_dynamic(st, hg_t, hgd_t, **cs)
    torch.cuda.synchronize()
    return st, dyn


@pytest.mark.parametrize("free_wake", [False, True])
def test_aileron_static_and_dynamic_vs_oracle(dev, free_wake):
    ost, odyn = _oracle(free_wake)
    st, dyn = _device(free_wake)
    for key in ("zeta_b", "gamma_b", "f_steady"):
        err = rel_err(getattr(st, key)[0][0].cpu().numpy(), getattr(ost, key)[0][0])
        print(f"static {key}: {err:.2e}")
        assert err < 1e-10, (key, err)
    for key, tol in (("zeta_b", 1e-12), ("zeta_b_dot", 1e-12), ("gamma_b", 1e-10), ("f_steady", 1e-10),
                     ("f_unsteady", 1e-10), ("zeta_w", 1e-8)):
        err = rel_err(getattr(dyn, key)[0].cpu().numpy(), getattr(odyn, key)[0])
        print(f"dynamic {key}: {err:.2e}")
        assert err < tol, (key, err)
    # the aileron moves the flow solution (the comparison is not vacuous)
    plain = _device(free_wake, grid_func=False)[1]
    assert rel_err(plain.gamma_b[0].cpu().numpy(), odyn.gamma_b[0]) > 1e-4


def test_undeflected_grid_function_equals_plain_path(dev):
    """A grid function held at zero deflection takes the general step path
    and gives the fused path's results."""
    a = _device(False, grid_func=True, angles=False)[1]
    b = _device(False, grid_func=False)[1]
    for key in ("gamma_b", "f_steady", "f_unsteady", "zeta_w", "zeta_b_dot"):
        err = rel_err(getattr(a, key)[0].cpu().numpy(), getattr(b, key)[0].cpu().numpy())
        assert err < 1e-12, (key, err)


def test_control_velocities_and_pairs(dev):
    from flapjax_amd.utils import cs_ang_to_cs_vel

    ang = np.array([0.1, 0.15, 0.12, 0.3])
    vel = cs_ang_to_cs_vel({"a": ang}, 0.5)["a"].cpu().numpy()
    assert np.array_equal(vel, np.concatenate(([0.0], np.diff(ang) / 0.5)))
    from flapjax_amd.utils import add_control_surface
    from flapjax_amd.uvlm import UVLM

    uv = UVLM([(M, N, MS)], np.arange(N + 1),
              grid_func=lambda x0, *, aileron: [add_control_surface(x0[0], aileron, FLAP_M, FLAP_N)])
    dt, hg_t, hgd_t, ang, vel = _motion()
    import flapjax_amd as fj

    uv.set_design_variables(dt, fj.Constant(np.array([1.0, 0.0, 0.1]), 1.225, True),
                            O.make_rectangular_grid(M, N, 1.0, 0.25), hg_t[0], reference_cs_angles={"aileron": 0.0})
    with pytest.raises(ValueError, match="Missing pair of control angles and velocities"):
        uv.hg_dot_to_zeta_b_dot(hg_t[1], hgd_t[1], {"aileron": 0.1}, {})


def test_coupled_aileron_vs_oracle(dev):
    """Aitken static FSI at the reference aileron angle, then FSI time steps
    with a moving aileron (velocities by backward differences on both sides,
    coupled/coupled.py:322-342): the device beam + UVLM follow the oracle's
    iterates, held to 10x the oracle's one-ulp-trig floor."""
    from oracle import beam_oracle as B
    from tests.test_gpu_beam import _beam_from_oracle, _held

    import flapjax_amd as fj
    from flapjax_amd import UVLM, GridDiscretisation
    from flapjax_amd.beam import ConvergenceSettings
    from flapjax_amd.coupled import CoupledAeroelastic
    from flapjax_amd.utils import add_control_surface

    fm, fn = slice(1, None), slice(0, 2)
    n_t = 8
    ocw = B.cantilever_wing(m=2, n_nodes=5, m_star=6, k_diag=(1e6, 1e6, 1e6, 1e3, 5e3, 1e5), u_inf=(10.0, 0.0, 1.0))
    ou = ocw.uvlm
    ang = 0.05 + 0.1 * np.sin(3.0 * ou.dt * np.arange(n_t))

    def oracle_run():
        ou.grid_func = lambda x0, *, aileron: [O.add_control_surface(x0[0], aileron, fm, fn)]
        ou.set_design_variables(ou.dt, ou.flowfield, ou.zeta_b0, ocw.beam.hg0, reference_cs_angles={"aileron": ang[0]})
        st_ = ocw.static_solve(np.arange(6))
        init = ou.solve_static(st_["hg"])
        return st_, ocw.dynamic_solve(st_["hg"], init, np.arange(6), n_t, cs_ang_t={"aileron": ang})

    ref_st, ref = oracle_run()
    with B.ulp_shifted_trig():
        pst, pdyn = oracle_run()
    ob = ocw.beam
    beam = _beam_from_oracle(ob)
    m, n, ms = ou.grid_disc[0]
    uv = UVLM([GridDiscretisation(m, n, ms)], np.arange(ob.n_nodes), mirror_point=np.zeros(3),
              mirror_normal=np.array((0.0, 1.0, 0.0)),
              grid_func=lambda x0, *, aileron: [add_control_surface(x0[0], aileron, fm, fn)])
    s = ocw.fsi_settings
    cw = CoupledAeroelastic(beam, uv, ConvergenceSettings(s.max_n_iter, s.rel_disp_tol, s.abs_disp_tol,
                                                          s.rel_force_tol, s.abs_force_tol))
    ff = ou.flowfield
    cw.set_design_variables(ob.x0, ob.k_cs, ob.m_cs, None, ou.dt, fj.Constant(ff.u_inf, ff.rho, True), ou.zeta_b0[0],
                            cs_angles_reference={"aileron": ang[0]})
    st = cw.static_solve(np.arange(6))
    assert st.structure.iters["fsi"] == ref_st["iters"]
    _held("static hg", st.structure.hg.cpu().numpy(), ref_st["hg"], pst["hg"])
    dyn = cw.dynamic_solve(st, np.arange(6), n_t, cs_ang_t={"aileron": ang})
    assert dyn.structure.iters["fsi"] == ref["info"]["fsi_iters"]
    _held("hg", dyn.structure.hg.cpu().numpy(), ref["hg"], pdyn["hg"])
    _held("gamma_b", dyn.aero.gamma_b[0].cpu().numpy(), ref["aero"].gamma_b[0], pdyn["aero"].gamma_b[0])
    _held("zeta_b_dot", dyn.aero.zeta_b_dot[0].cpu().numpy(), ref["aero"].zeta_b_dot[0], pdyn["aero"].zeta_b_dot[0])
    _held("f_ext_aero", dyn.structure.f_ext_aero.cpu().numpy(), ref["f_ext_aero"], pdyn["f_ext_aero"])
