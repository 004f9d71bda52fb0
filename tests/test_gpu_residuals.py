"""The discrete-adjoint residual layer (aero/uvlm.py:1529-2047) and the
GMRES adjoint driver with the frozen-wake preconditioner
(coupled/gradients/coupled.py:848-990, 1234-1249) on the device.

  * every residual vanishes (to the solve's rounding) on the drop-in UVLM's
    own history: the layer states the same equations the step solves;
  * the discrete adjoint of a prescribed-motion run, solved per step by
    GMRES on reverse passes through the residuals, gives the gradient of the
    time-integrated lift that reverse mode through the differentiable rerun
    gives (itself checked against central differences in test_gpu_adjoint);
    the frozen-wake preconditioner cuts the GMRES iterations.
"""

import numpy as np
import pytest
import torch

from oracle import uvlm_oracle as O
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _case(free_wake, m=4, n=8, m_star=8, steps=7, span=4.0):
    import flapjax_amd as fj
    from flapjax_amd.uvlm import UVLM

    dt = 1.0 / m
    zeta0 = O.make_rectangular_grid(m, n, 1.0, 0.25)
    hg_t, hgd_t = O.pitch_plunge_motion(n, span, steps, dt, np.deg2rad(4.0), np.deg2rad(3.0), 0.05, 1.2)
    uv = UVLM([(m, n, m_star)], np.arange(n + 1), free_wake=free_wake)
    uv.set_design_variables(dt, fj.Constant(np.array([1.0, 0.0, 0.0]), 1.225, True), zeta0, hg_t[0])
    return uv, zeta0, hg_t, hgd_t


@pytest.mark.parametrize("free_wake", [False, True])
def test_residuals_vanish_on_the_drop_in_history(dev, free_wake):
    from flapjax_amd.residuals import UVLMResiduals

    uv, _, hg_t, hgd_t = _case(free_wake)
    d = uv.solve_prescribed_dynamic(uv.solve_static(hg=hg_t[0]), hg_t, hgd_t)
    res = UVLMResiduals(uv)
    hg = torch.as_tensor(hg_t, device=dev)
    hgd = torch.as_tensor(hgd_t, device=dev)
    nb, nw = 4 * 8, 8 * 8
    for i in range(2, hg_t.shape[0]):
        qn, qp = d.get_states(i), d.get_states(i - 1)
        # the structure's force state: the case's own forces, local frame
        f = [d.f_steady[0][i] + d.f_unsteady[0][i]]
        f_beam = d.project_forcing_to_beam(i, hg[i][:, :3, :3], [uv.zeta_b0[0]], True)
        rt = hg[i][:, :3, :3].transpose(-1, -2)
        f_loc = torch.cat(((rt @ f_beam[:, :3, None])[..., 0], (rt @ f_beam[:, 3:, None])[..., 0]), -1)
        r = res.timestep_residual(float(d.t[i]), hg[i - 1], hg[i], hgd[i], qn, qp, f_loc)
        g = qn.gamma_b[0].abs().max()
        assert float(r[:nb].abs().max()) < 1e-10 * float(g)
        assert float(r[nb:nb + nw].abs().max()) == 0.0  # the wake circulation shift is exact
        assert float(r[nb + nw:2 * nb + nw].abs().max()) < 1e-9 * float(qn.gamma_b_dot[0].abs().max() + g)
        assert float(r[2 * nb + nw:-f_loc.numel()].abs().max()) < 1e-13 * float(qn.zeta_w[0].abs().max())
        assert float(r[-f_loc.numel():].abs().max()) < 1e-10 * float(f_loc.abs().max())
        del f


@pytest.mark.parametrize("free_wake", [False, True])
def test_discrete_adjoint_gradient(dev, free_wake):
    from flapjax_amd.adjoint import DifferentiableUVLM, gradient
    from flapjax_amd.residuals import PrescribedAdjoint

    uv, zeta0, hg_t, hgd_t = _case(free_wake)
    z0 = torch.as_tensor(zeta0, device=dev)
    hg = torch.as_tensor(hg_t, device=dev)
    hgd = torch.as_tensor(hgd_t, device=dev)
    tw = torch.as_tensor(np.random.default_rng(0).normal(size=9) * 0.02, device=dev)
    u = torch.tensor([1.0, 0.0, 0.05], dtype=torch.float64, device=dev)
    model = DifferentiableUVLM(4, 8, 8, uv.dt, 1.225, free_wake, uv.gamma_dot_relaxation, uv.kernel)
    J_ref, g_tw_ref, g_u_ref = gradient(model, tw, u, z0, hg, hgd)
    adj = PrescribedAdjoint(uv, gmres_tol=1e-13)
    J, g_tw, g_u = adj.gradient(tw, u, z0, hg, hgd)
    assert abs(J - float(J_ref)) <= 1e-12 * abs(float(J_ref))
    e_tw = rel_err(g_tw.cpu().numpy(), g_tw_ref.cpu().numpy())
    e_u = rel_err(g_u.cpu().numpy(), g_u_ref.cpu().numpy())
    iters = [it for it, _ in adj.gmres_iters]
    print(f"free={free_wake}: dJ/dtwist {e_tw:.2e}, dJ/dU {e_u:.2e}, GMRES steps per time step {iters}")
    assert e_tw < 1e-8 and e_u < 1e-8
    assert max(r for _, r in adj.gmres_iters) < 1e-12
    plain = PrescribedAdjoint(uv, gmres_tol=1e-13, precond=False)
    plain.gradient(tw, u, z0, hg, hgd)
    assert sum(iters) < sum(it for it, _ in plain.gmres_iters)
