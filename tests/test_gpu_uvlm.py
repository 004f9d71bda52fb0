"""GPU parity of full UVLM solves against the FP64 oracle.

Bar (from the build target): circulation, induced velocity and vertex forces
within 1e-10 relative (norm-wise), wake vertex positions after N steps within
1e-8 relative.
"""

import numpy as np
import pytest
import torch

from oracle import uvlm_oracle as O
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10
TOL_WAKE = 1e-8


def _setup(m, n, m_star, span, alpha_deg, free_wake=False, kernel="epsilon", mirror=False, precision="fp64",
           u_inf=(1.0, 0.0, 0.0), gust=None):
    import flapjax_amd as fj
    from flapjax_amd.uvlm import UVLM
    from flapjax_amd.utils import biot_savart_cutoff, biot_savart_epsilon

    zeta0 = O.make_rectangular_grid(m, n, 1.0, 0.25)
    hg = O.rigid_wing_hg(n, span, np.deg2rad(alpha_deg))
    dt = 1.0 / (m * np.linalg.norm(u_inf))
    mp = np.zeros(3) if mirror else None
    mn = np.array([0.0, 1.0, 0.0]) if mirror else None
    orc = O.UVLMOracle([(m, n, m_star)], np.arange(n + 1), kernel=kernel, free_wake=free_wake, mirror_point=mp,
                       mirror_normal=mn)
    if gust is None:
        ff_o = O.Constant(np.array(u_inf), 1.225, True)
        ff_g = fj.Constant(np.array(u_inf), 1.225, True)
    else:
        ff_o = O.OneMinusCosine(np.array(u_inf), 1.225, True, gust_length=gust[0], gust_amplitude=gust[1])
        ff_g = fj.OneMinusCosine(np.array(u_inf), 1.225, True, gust_length=gust[0], gust_amplitude=gust[1])
    orc.set_design_variables(dt, ff_o, zeta0, hg)
    kobj = {"epsilon": biot_savart_epsilon, "cutoff": biot_savart_cutoff}[kernel]
    gpu = UVLM([(m, n, m_star)], np.arange(n + 1), kernel=kobj, free_wake=free_wake, mirror_point=mp,
               mirror_normal=mn, precision=precision)
    gpu.set_design_variables(dt, ff_g, zeta0, hg)
    return orc, gpu, dt


def _cmp(case_g, case_o, keys, tol, i=None):
    errs = {}
    for key in keys:
        a = getattr(case_g, key)[0].cpu().numpy()
        b = getattr(case_o, key)[0]
        if i is not None:
            a, b = a[i], b[i]
        errs[key] = rel_err(a, b)
    bad = {k: v for k, v in errs.items() if v > tol}
    assert not bad, f"parity failures: {bad} (all: {errs})"
    return errs


@pytest.mark.parametrize("horseshoe", [False, True])
def test_static_cfg0(dev, horseshoe):
    """8 x 32 rigid wing, 5 deg, flat wake of 20 rows (static solve)."""
    orc, gpu, _ = _setup(8, 32, 20, 10.0, 5.0)
    co = orc.solve_static(horseshoe=horseshoe)
    cg = gpu.solve_static(horseshoe=horseshoe)
    errs = _cmp(cg, co, ("gamma_b", "gamma_w", "f_steady", "alpha", "c", "nc", "zeta_b", "zeta_w"), TOL)
    print("static cfg0 errors:", errs)


def test_static_mirror_and_cutoff(dev):
    orc, gpu, _ = _setup(5, 10, 20, 5.0, 3.0, kernel="cutoff", mirror=True, u_inf=(10.0, 0.0, 1.0))
    _cmp(gpu.solve_static(), orc.solve_static(), ("gamma_b", "gamma_w", "f_steady", "alpha"), TOL)


@pytest.mark.parametrize("free_wake", [False, True])
def test_dynamic_pitch_plunge(dev, free_wake):
    """Small pitch-plunge run: 8 x 16, 12 wake rows, 10 steps."""
    m, n, m_star, steps = 8, 16, 12, 10
    orc, gpu, dt = _setup(m, n, m_star, 4.0, 5.0, free_wake=free_wake)
    hg_t, hg_dot_t = O.pitch_plunge_motion(n, 4.0, steps, dt, np.deg2rad(5.0), np.deg2rad(3.0), 0.05, 0.4 * 4)
    so = orc.solve_static(hg=hg_t[0])
    sg = gpu.solve_static(hg=hg_t[0])
    do = orc.solve_prescribed_dynamic(so, hg_t, hg_dot_t)
    dg = gpu.solve_prescribed_dynamic(sg, hg_t, hg_dot_t)
    errs = _cmp(dg, do, ("gamma_b", "gamma_w", "gamma_b_dot", "f_steady", "f_unsteady", "alpha"), TOL)
    werr = _cmp(dg, do, ("zeta_w",), TOL_WAKE)
    print("dynamic errors:", errs, werr)


def test_dynamic_gust_mirror(dev):
    """1-cos gust on a mirrored half wing with free wake."""
    m, n, m_star, steps = 4, 10, 8, 8
    orc, gpu, dt = _setup(m, n, m_star, 3.0, 2.0, free_wake=True, mirror=True, gust=(1.0, 0.05))
    hg_t = np.stack([O.rigid_wing_hg(n, 3.0, np.deg2rad(2.0))] * steps)
    hg_dot_t = np.zeros_like(hg_t)
    do = orc.solve_prescribed_dynamic(orc.solve_static(), hg_t, hg_dot_t)
    dg = gpu.solve_prescribed_dynamic(gpu.solve_static(), hg_t, hg_dot_t)
    _cmp(dg, do, ("gamma_b", "gamma_w", "f_steady", "f_unsteady"), TOL)
    _cmp(dg, do, ("zeta_w",), TOL_WAKE)


def test_mixed_precision_error_is_reported(dev):
    """The mixed form (FP32 bulk + FP64 refinement) against the FP64 oracle."""
    m, n, m_star, steps = 8, 16, 12, 10
    orc, gpu, dt = _setup(m, n, m_star, 4.0, 5.0, free_wake=True, precision="mixed")
    hg_t, hg_dot_t = O.pitch_plunge_motion(n, 4.0, steps, dt, np.deg2rad(5.0), np.deg2rad(3.0), 0.05, 1.6)
    do = orc.solve_prescribed_dynamic(orc.solve_static(hg=hg_t[0]), hg_t, hg_dot_t)
    dg = gpu.solve_prescribed_dynamic(gpu.solve_static(hg=hg_t[0]), hg_t, hg_dot_t)
    errs = {k: rel_err(getattr(dg, k)[0].cpu().numpy(), getattr(do, k)[0])
            for k in ("gamma_b", "f_steady", "zeta_w")}
    print("mixed-form errors vs FP64 oracle:", errs)
    assert errs["gamma_b"] < 1e-6 and errs["f_steady"] < 1e-5 and errs["zeta_w"] < 1e-6


@pytest.mark.parametrize("free_wake", [False, True])
def test_variable_wake_discretisation(dev, free_wake):
    """delta_w path (aero/utils.py:452-491) vs the oracle, heaving wing as in
    the reference's tests/aero/test_variable_wake_disc.py."""
    import flapjax_amd as fj
    from flapjax_amd.uvlm import UVLM

    m, n, m_star, steps = 4, 8, 10, 8
    dt = 1.0 / (m * 10.0)
    zeta0 = O.make_rectangular_grid(m, n, 1.0, 0.0)
    hg = O.rigid_wing_hg(n, 5.0, 0.0)
    t = np.arange(steps) * dt
    w = 2 * np.pi * 3.0
    hg_t = np.stack([hg] * steps)
    hg_t[:, :, 2, 3] += (0.1 * 0.5 * (1 - np.cos(w * t)))[:, None]
    hgd_t = np.zeros_like(hg_t)
    hgd_t[:, :, 2, 3] = (0.1 * w * 0.5 * np.sin(w * t))[:, None]
    delta_w = np.linspace(1.0, 1.6, m_star) * dt * 10.0
    orc = O.UVLMOracle([(m, n, m_star)], np.arange(n + 1), free_wake=free_wake, gamma_dot_relaxation=1.0)
    orc.set_design_variables(dt, O.Constant(np.array([10.0, 0, 0]), 1.225), zeta0, hg, delta_w=[delta_w])
    gpu = UVLM([(m, n, m_star)], np.arange(n + 1), free_wake=free_wake, gamma_dot_relaxation=1.0,
               variable_wake_disc=True)
    gpu.set_design_variables(dt, fj.Constant(np.array([10.0, 0, 0]), 1.225, True), zeta0, hg, delta_w=[delta_w])
    do = orc.solve_prescribed_dynamic(orc.solve_static(), hg_t, hgd_t)
    dg = gpu.solve_prescribed_dynamic(gpu.solve_static(), hg_t, hgd_t)
    _cmp(dg, do, ("gamma_b", "gamma_w", "f_steady"), TOL)
    _cmp(dg, do, ("zeta_w",), TOL_WAKE)


def test_non_finite_motion_propagates_like_the_reference():
    """A NaN in the prescribed motion at step k gives NaN results from step k
    on (with a warning), not an exception -- the reference's behaviour."""
    m, n, m_star, steps = 4, 8, 6, 6
    _, gpu, dt = _setup(m, n, m_star, 4.0, 5.0, free_wake=True)
    hg_t, hg_dot_t = O.pitch_plunge_motion(n, 4.0, steps, dt, np.deg2rad(5.0), np.deg2rad(3.0), 0.05, 1.6)
    hg_t = hg_t.copy()
    hg_t[4, 2, 0, 3] = np.nan
    with pytest.warns(UserWarning, match="Non-finite"):
        d = gpu.solve_prescribed_dynamic(gpu.solve_static(hg=hg_t[0]), hg_t, hg_dot_t)
    g = d.gamma_b[0].cpu().numpy()
    assert np.isfinite(g[:4]).all()
    assert not np.isfinite(g[4]).all()


@pytest.mark.parametrize("mirror,two_surf,free", [(False, False, True), (True, False, True), (False, True, True),
                                                   (False, False, False)])
def test_fused_step_equals_general_path(mirror, two_surf, free):
    """The fused dynamic step (one geometry launch per surface, shared
    filament tables, fused forces) gives the results of the general
    base_solve_from_grid path, step by step."""
    import flapjax_amd as fj
    from flapjax_amd.uvlm import UVLM

    m, n, m_star, steps = 4, 8, 6, 7
    span = 4.0
    zeta0 = O.make_rectangular_grid(m, n, 1.0, 0.25)
    hg_t, hg_dot_t = O.pitch_plunge_motion(n, span, steps, 0.25, np.deg2rad(5.0), np.deg2rad(3.0), 0.05, 1.6)
    shapes = [(m, n, m_star)] * (2 if two_surf else 1)
    dofs = [np.arange(n + 1)] * len(shapes)
    z0s = [zeta0, zeta0 + np.array([0.0, 0.0, 0.5])] if two_surf else zeta0
    mp = np.zeros(3) if mirror else None
    mn = np.array([0.0, 1.0, 0.0]) if mirror else None
    runs = []
    for fused in (True, False):
        uv = UVLM(shapes, dofs, free_wake=free, mirror_point=mp, mirror_normal=mn)
        uv.set_design_variables(0.25, fj.Constant(np.array([1.0, 0.0, 0.0]), 1.225, True), z0s, hg_t[0])
        if not fused:
            uv._fast_step_ok = lambda *a: False
        runs.append(uv.solve_prescribed_dynamic(uv.solve_static(hg=hg_t[0]), hg_t, hg_dot_t))
    a, b = runs
    for key in ("gamma_b", "gamma_w", "zeta_w", "f_steady", "f_unsteady", "gamma_b_dot", "alpha", "c", "nc",
                "zeta_b_dot"):
        for s in range(len(shapes)):
            x, y = getattr(a, key)[s].cpu().numpy(), getattr(b, key)[s].cpu().numpy()
            assert rel_err(x, y) <= 1e-15, (key, s, rel_err(x, y))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_forces_and_next_rollup_in_one_launch(precision):
    """Free wake, one surface: the forces' pair sum of step n and the rollup
    of step n + 1 run as one launch (fj_pair_sum_lattices_split in FP64,
    fj_pair_sum_split in the mixed form); the results equal separate
    launches -- in FP64 to the last bit's order (the same partial plan), in
    the mixed form within its FP32 far-field error (the fused launch groups
    the FP32 partial sums differently) -- also when a step is solved twice
    (the second solve of step n finds no mark for n and rolls up itself) and
    when the next step is given another pose buffer as hg_nm1 (no fusion)."""
    import torch

    import flapjax_amd as fj
    from flapjax_amd.uvlm import UVLM
    from flapjax_amd.utils import to_device

    m, n, m_star, steps = 6, 12, 8, 10
    zeta0 = O.make_rectangular_grid(m, n, 1.0, 0.25)
    hg_np, hgd_np = O.pitch_plunge_motion(n, 4.0, steps, 0.2, np.deg2rad(5.0), np.deg2rad(3.0), 0.05, 1.2)
    hg, hgd = to_device(hg_np), to_device(hgd_np)
    mp, mn = np.zeros(3), np.array([0.0, 1.0, 0.0])

    def run(fuse, resolve=(), copy_at=()):
        uv = UVLM([(m, n, m_star)], np.arange(n + 1), free_wake=True, mirror_point=mp, mirror_normal=mn,
                  precision=precision)
        uv.set_design_variables(0.2, fj.Constant(np.array([1.0, 0.0, 0.1]), 1.225, True), zeta0, hg_np[0])
        uv.fuse_rollup = fuse
        case = uv.solve_static(hg=hg[0]).to_dynamic(0, steps)
        for i in range(1, steps):
            h_nm1 = hg[i - 1].clone() if i in copy_at else hg[i - 1]
            uv.case_solve(case, i, hg[i], h_nm1, hgd[i], False, False)
            if i in resolve:
                uv.case_solve(case, i, hg[i], h_nm1, hgd[i], False, False)
        torch.cuda.synchronize()
        return case

    ref = run(False)
    for kw in ({}, {"resolve": (3, 4)}, {"copy_at": (5,)}):
        got = run(True, **kw)
        for key in ("gamma_b", "gamma_w", "zeta_w", "f_steady", "f_unsteady", "gamma_b_dot", "alpha"):
            e = rel_err(getattr(got, key)[0].cpu().numpy(), getattr(ref, key)[0].cpu().numpy())
            tol = 1e-15 if precision == "fp64" else (1e-4 if key in ("f_steady", "f_unsteady", "gamma_b_dot")
                                                     else 1e-5)
            assert e <= tol, (kw, key, e)
