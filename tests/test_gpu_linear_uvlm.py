"""LinearUVLM (aero/linear/linear_uvlm.py) on the device.

  * the reference's nonlinear-vs-linear tests (tests/aero/test_linear_planar_wing.py:68-380):
    heave and pitch with a prescribed wake, pitch with a frozen wake, and a
    1-cos gust through the bound/wake upwash inputs, with their tolerances,
    the nonlinear side being this build's own drop-in UVLM;
  * exactness of the linearisation: the Jacobian [A B; C D] against central
    differences of the step, and the linear response to a small motion
    against the nonlinear response minus the reference, divided by the
    amplitude (agreement to first order in the amplitude).
"""

import numpy as np
import pytest
import torch

from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

U_INF = np.array((10.0, 0.0, 1.0))
M, N, M_STAR, C_REF, B_REF = 4, 8, 10, 1.0, 5.0
DT = C_REF / (M * np.linalg.norm(U_INF))
N_TSTEP = int(np.ceil(1.0 / DT))


def _planar_wing(flowfield, ea=0.0):
    """tests/aero/test_linear_planar_wing.py:35-66 (cut-off kernel,
    gamma_dot relaxation 1)."""
    from flapjax_amd.uvlm import UVLM
    from flapjax_amd.utils import biot_savart_cutoff, make_rectangular_grid

    hg = np.zeros((N + 1, 4, 4))
    hg[:, 3, 3] = 1.0
    hg[:, :3, :3] = np.eye(3)
    hg[:, 1, 3] = np.linspace(0.0, B_REF, N + 1)
    uvlm = UVLM([(M, N, M_STAR)], np.arange(N + 1), kernel=biot_savart_cutoff, gamma_dot_relaxation=1.0)
    uvlm.set_design_variables(DT, flowfield, make_rectangular_grid(M, N, C_REF, ea), hg)
    return uvlm, uvlm.solve_static(), hg


def _constant():
    import flapjax_amd as fj

    return fj.Constant(U_INF, 1.225, True)


def _heave(hg0, ampl=0.1):
    t = np.arange(N_TSTEP) * DT
    om = 0.5 * np.pi * 3.0
    hg_t = np.broadcast_to(hg0, (N_TSTEP, N + 1, 4, 4)).copy()
    hg_t[:, :, 2, 3] += (ampl * 0.5 * (1.0 - np.cos(om * t)))[:, None]
    hgd = np.zeros_like(hg_t)
    hgd[:, :, 2, 3] = (ampl * om * 0.5 * np.sin(om * t))[:, None]
    return hg_t, hgd


def _pitch(hg0, ampl=np.deg2rad(4.0)):
    """Rotation about y by alpha(t) with the reference test's skew rate block
    (small-angle hg_dot, tests/aero/test_linear_planar_wing.py:168-170)."""
    t = np.arange(N_TSTEP) * DT
    om = 0.5 * np.pi * 3.0
    a = ampl * 0.5 * (1.0 - np.cos(om * t))
    ad = ampl * om * 0.5 * np.sin(om * t)
    hg_t = np.broadcast_to(hg0, (N_TSTEP, N + 1, 4, 4)).copy()
    c, s = np.cos(a), np.sin(a)
    rot = np.zeros((N_TSTEP, 3, 3))
    rot[:, 0, 0], rot[:, 0, 2], rot[:, 1, 1], rot[:, 2, 0], rot[:, 2, 2] = c, s, 1.0, -s, c
    hg_t[:, :, :3, :3] = rot[:, None]
    hgd = np.zeros_like(hg_t)
    hgd[:, :, 0, 2] = ad[:, None]
    hgd[:, :, 2, 0] = -ad[:, None]
    return hg_t, hgd


def _linear_run(uvlm, static, dyn, wake_type, upwash, flowfield=None):
    from flapjax_amd.linear_uvlm import AeroInputUnflattened

    lin = uvlm.linearise(reference=static, wake_type=wake_type, bound_upwash=upwash, wake_upwash=upwash,
                         unsteady_force=True)
    dz = [z - r[0][None] for z, r in zip(dyn.zeta_b, lin.reference.zeta_b)]
    u = AeroInputUnflattened(zeta_b=dz, zeta_b_dot=[z.clone() for z in dyn.zeta_b_dot], nu_b=None, nu_w=None)
    return lin, lin.run(u, flowfield=flowfield)


def _close(a, b, atol):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    err = float(np.max(np.abs(a - b)))
    assert err <= atol, (err, atol)
    return err


def test_linear_heaving_wing(dev):
    uvlm, static, hg0 = _planar_wing(_constant())
    hg_t, hgd = _heave(hg0)
    dyn = uvlm.solve_prescribed_dynamic(static, hg_t, hgd)
    _, res = _linear_run(uvlm, static, dyn, "prescribed", False)
    _close(dyn.gamma_b[0], res.x_t_tot.gamma_b[0], 5e-3)
    _close(dyn.gamma_w[0], res.x_t_tot.gamma_w[0], 5e-3)
    _close(dyn.zeta_w[0], res.x_t_tot.zeta_w[0], 1e-3)
    _close(dyn.f_steady[0], res.y_t_tot.f_steady[0], 1e-1)
    _close(dyn.f_unsteady[0], res.y_t_tot.f_unsteady[0], 2e-2)


def test_linear_pitching_wing(dev):
    uvlm, static, hg0 = _planar_wing(_constant())
    hg_t, hgd = _pitch(hg0)
    dyn = uvlm.solve_prescribed_dynamic(static, hg_t, hgd)
    _, res = _linear_run(uvlm, static, dyn, "prescribed", False)
    _close(dyn.gamma_b[0], res.x_t_tot.gamma_b[0], 6e-2)
    _close(dyn.gamma_w[0], res.x_t_tot.gamma_w[0], 6e-2)
    _close(dyn.zeta_w[0], res.x_t_tot.zeta_w[0], 1e-3)
    _close(dyn.f_steady[0], res.y_t_tot.f_steady[0], 3e-1)
    _close(dyn.f_unsteady[0], res.y_t_tot.f_unsteady[0], 1e-1)


def test_linear_pitching_wing_frozen_wake(dev):
    uvlm, static, hg0 = _planar_wing(_constant(), ea=1.0)
    hg_t, hgd = _pitch(hg0)
    dyn = uvlm.solve_prescribed_dynamic(static, hg_t, hgd)
    lin, res = _linear_run(uvlm, static, dyn, "frozen", False)
    assert res.x_t_tot.zeta_w is None and lin.sys.n_states == 2 * M * N + M_STAR * N
    _close(dyn.gamma_b[0], res.x_t_tot.gamma_b[0], 4e-2)
    _close(dyn.gamma_w[0], res.x_t_tot.gamma_w[0], 4e-2)
    _close(dyn.f_steady[0], res.y_t_tot.f_steady[0], 3e-1)
    _close(dyn.f_unsteady[0], res.y_t_tot.f_unsteady[0], 1e-1)


def test_linear_cosine_gust(dev):
    import flapjax_amd as fj

    ff = fj.OneMinusCosine(U_INF, 1.225, True, gust_length=2.0, gust_amplitude=0.4, gust_x0=(-5.0, 0.0, 0.0))
    uvlm, static, hg0 = _planar_wing(ff)
    hg_t = np.broadcast_to(hg0, (N_TSTEP, N + 1, 4, 4)).copy()
    dyn = uvlm.solve_prescribed_dynamic(static, hg_t, np.zeros_like(hg_t))
    _, res = _linear_run(uvlm, static, dyn, "prescribed", True, flowfield=ff)
    assert float(dyn.gamma_b[0].abs().max() - static.gamma_b[0].abs().max()) > 0.05  # the gust does act
    _close(dyn.gamma_b[0], res.x_t_tot.gamma_b[0], 3e-1)
    _close(dyn.gamma_w[0], res.x_t_tot.gamma_w[0], 3e-1)
    _close(dyn.zeta_w[0], res.x_t_tot.zeta_w[0], 8e-2)
    _close(dyn.f_steady[0], res.y_t_tot.f_steady[0], 6e-1)
    _close(dyn.f_unsteady[0], res.y_t_tot.f_unsteady[0], 8e-1)


@pytest.mark.parametrize("wake_type", ["prescribed", "free"])
def test_jacobian_equals_central_differences(dev, wake_type):
    """[A B; C D] times random directions against central differences of the
    step at the reference (h = 1e-6): exact linearisation of every block."""
    import flapjax_amd as fj

    ff = fj.OneMinusCosine(U_INF, 1.225, True, gust_length=2.0, gust_amplitude=0.4, gust_x0=(-0.5, 0.0, 0.0))
    uvlm, static, _ = _planar_wing(ff)
    lin = uvlm.linearise(reference=static, wake_type=wake_type, bound_upwash=True, wake_upwash=True)
    s = lin.sys
    x0 = lin._pack(lin._ref_values(lin.states), lin.states)
    u0 = lin._pack(lin._ref_values(lin.inputs), lin.inputs)
    rng = np.random.default_rng(3)
    for _ in range(3):
        dx = torch.as_tensor(rng.normal(size=x0.shape), device=dev)
        du = torch.as_tensor(rng.normal(size=u0.shape), device=dev)
        h = 1e-6

        def f(sgn):
            with torch.no_grad():
                xn, yn = lin.step(lin._unpack(u0 + sgn * h * du, lin.inputs), lin._unpack(x0 + sgn * h * dx, lin.states))
            return torch.cat((lin._pack(xn, lin.states), lin._pack(yn, lin.outputs)))

        fd = (f(1.0) - f(-1.0)) / (2 * h)
        an = torch.cat((s.a @ dx + s.b @ du, s.c @ dx + s.d @ du))
        assert rel_err(an.cpu().numpy(), fd.cpu().numpy()) < 1e-6


def test_small_motion_first_order(dev):
    """The linear response to a heave of amplitude 1e-5 equals the nonlinear
    response minus the reference to first order: relative agreement O(1e-5)."""
    uvlm, static, hg0 = _planar_wing(_constant())
    hg_t, hgd = _heave(hg0, ampl=1e-5)
    dyn = uvlm.solve_prescribed_dynamic(static, hg_t, hgd)
    _, res = _linear_run(uvlm, static, dyn, "prescribed", False)
    d_nl = (dyn.gamma_b[0] - static.gamma_b[0][0][None]).cpu().numpy()
    d_lin = res.x_t.gamma_b[0].cpu().numpy()
    assert np.max(np.abs(d_lin)) > 0
    assert rel_err(d_lin, d_nl) < 1e-4
    f_nl = (dyn.f_steady[0] - static.f_steady[0][0][None]).cpu().numpy()
    assert rel_err(res.y_t.f_steady[0].cpu().numpy(), f_nl) < 1e-4
