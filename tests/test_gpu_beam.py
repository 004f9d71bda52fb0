"""Device beam and aeroelastic coupling against the oracle
(oracle/beam_oracle.py).

Tolerances.  The reference's SE(3) tangent formulas cancel above their
small-angle switch: a one-ulp difference in sin or cos (between any two
libraries, the reference's XLA included) moves T(h) by about
eps / |omega|^2, and the Newton iterations carry that into the solution, so
two correct FP64 evaluations of the reference's formulas agree only to a
problem-dependent level.  Each solve below is therefore run twice in the
oracle, once with sin and cos moved by one ulp (``ulp_shifted_trig``); the
difference is that level ("floor"), and the device result is held to
10 x floor (and never looser than needed: node maps on well-conditioned
inputs and the large-load static solve, whose floor is ~6e-12, are held to
1e-11 / 1e-10 with identical Newton iteration counts).
"""

import numpy as np
import pytest
import torch

from oracle import beam_oracle as B
from oracle import uvlm_oracle as U
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _held(name, dev_val, ref_val, pert_val, factor=10.0, least=1e-12):
    """Device within factor x the oracle's one-ulp-trig sensitivity."""
    err = rel_err(dev_val, ref_val)
    floor = rel_err(pert_val, ref_val)
    print(f"{name}: device {err:.3e}, floor {floor:.3e}")
    assert err <= max(least, factor * floor), (name, err, floor)


def _beam_from_oracle(ob, settings=None, **kw):
    from flapjax_amd.beam import DEFAULT_STRUCT_CONVERGENCE_SETTINGS, BeamStructure, ConvergenceSettings

    s = ob.settings
    cs = ConvergenceSettings(s.max_n_iter, s.rel_disp_tol, s.abs_disp_tol, s.rel_force_tol, s.abs_force_tol)
    b = BeamStructure(ob.n_nodes, ob.conn, ob.y_vector, relaxation_factor=ob.relaxation_factor,
                      spectral_radius=ob.spectral_radius,
                      struct_convergence_settings=cs if settings is None else settings, **kw)
    b.set_design_variables(ob.x0, ob.k_cs, ob.m_cs)
    assert DEFAULT_STRUCT_CONVERGENCE_SETTINGS == ConvergenceSettings(*B.DEFAULT_STRUCT.__dict__.values())
    return b


def test_node_maps(dev):
    from flapjax_amd import _lib as L
    from flapjax_amd.beam import BeamStructure

    rng = np.random.default_rng(3)
    n = 64
    h = rng.normal(size=(n, 6))
    h[:, 3:] *= np.where(np.arange(n) % 4 == 0, 0.0, 0.5)[:, None]  # a quarter in the series branch
    h[1::4, 3:] = rng.normal(size=(len(h[1::4]), 3)) * 3e-6            # and below |w|^2 = 1e-10
    hg0 = np.stack([B.exp_se3(x) for x in rng.normal(size=(n, 6)) * 0.7])
    ref_up = np.stack([a @ B.exp_se3(x) for a, x in zip(hg0, h)])
    ref_t = np.stack([B.t_se3(x) for x in h])
    ref_rel = np.stack([B.hg_to_d(a, b) for a, b in zip(hg0, ref_up)])
    beam = BeamStructure.__new__(BeamStructure)
    hg_d = torch.tensor(hg0, device=dev)
    out = torch.empty_like(hg_d)
    tt = torch.empty((n, 6, 6), dtype=torch.float64, device=dev)
    hh = torch.empty((n, 6), dtype=torch.float64, device=dev)
    beam._nodes(L.BEAM_NODE_UPDATE, hg_d, x=torch.tensor(h, device=dev), hg_ref=hg_d, hg_out=out, t_out=tt, h_out=hh)
    big = np.linalg.norm(h[:, 3:], axis=1) > 0.1
    assert rel_err(out.cpu().numpy(), ref_up) < 1e-14
    assert rel_err(tt.cpu().numpy()[big], ref_t[big]) < 1e-13
    small = np.linalg.norm(h[:, 3:], axis=1) ** 2 <= B.SMALL_ANG_THRESH
    assert rel_err(tt.cpu().numpy()[small], ref_t[small]) < 1e-14
    assert rel_err(hh.cpu().numpy(), ref_rel) < 1e-11
    v = rng.normal(size=(n, 6))
    rate = beam.make_hg_dot(hg_d, torch.tensor(v, device=dev)).cpu().numpy()
    assert rel_err(rate, np.stack([a @ B.tilde4(x) for a, x in zip(hg0, v)])) < 1e-15
    # force and moment into and out of the node frames (make_f_dead_ext,
    # structure/beam.py:1387-1395), with a time axis broadcast over hg
    f = rng.normal(size=(3, n, 6))
    r = hg0[:, :3, :3]
    loc = BeamStructure._to_local(torch.tensor(f, device=dev), hg_d).cpu().numpy()
    exp = np.concatenate((np.einsum("nji,tnj->tni", r, f[..., :3]), np.einsum("nji,tnj->tni", r, f[..., 3:])), -1)
    assert loc.shape == f.shape and rel_err(loc, exp) < 1e-15
    back = BeamStructure._to_global(torch.tensor(loc, device=dev), hg_d).cpu().numpy()
    assert rel_err(back, f) < 1e-14


def test_static_geradin_large_load(dev):
    """tests/structure/static_primal/test_geradin_beam.py, 600 kN dead tip
    load, 3 load steps: the same Newton iterates as the oracle."""
    ob = B.geradin_beam(20)
    f = np.zeros((ob.n_nodes, 6))
    f[-1, 2] = -600000.0
    ref = ob.static_solve(np.arange(6), f_ext_dead=f, load_steps=3)
    b = _beam_from_oracle(ob)
    out = b.static_solve(torch.arange(6), f_ext_dead=torch.tensor(f, device=dev), load_steps=3)
    assert out.iters == ref["iters"]
    assert rel_err(out.hg.cpu().numpy(), ref["hg"]) < 1e-10
    assert rel_err(out.varphi.cpu().numpy(), ref["varphi"]) < 1e-9
    assert np.isclose(out.hg[-1, 2, 3].item(), -2.157409, rtol=1e-2)


def test_static_geradin_small_load(dev):
    """The 600 N case: element rotations of ~5e-5 sit just above the
    reference's small-angle switch, where its formulas are noise-dominated
    (its own Newton loop stalls at the iteration limit in two load steps);
    held to the oracle's rounding floor and to the reference's golden value."""
    ob = B.geradin_beam(20)
    f = np.zeros((ob.n_nodes, 6))
    f[-1, 2] = -600.0
    ref = ob.static_solve(np.arange(6), f_ext_dead=f, load_steps=3)
    with B.ulp_shifted_trig():
        pert = ob.static_solve(np.arange(6), f_ext_dead=f, load_steps=3)
    b = _beam_from_oracle(ob)
    out = b.static_solve(torch.arange(6), f_ext_dead=torch.tensor(f, device=dev), load_steps=3)
    _held("hg", out.hg.cpu().numpy(), ref["hg"], pert["hg"])
    assert np.isclose(out.hg[-1, 2, 3].item(), -2.6819e-3, rtol=1e-2)


def test_static_follower_bent_beam(dev):
    """Follower loads on a curved, non-uniform three-dimensional beam."""
    n = 7
    ob = B.BeamOracle(n, B.chain(n), np.array((0.0, 0.0, 1.0)), k_cs_index=np.arange(n - 1) % 2)
    coords = np.zeros((n, 3))
    coords[:, 1] = np.linspace(0.0, 3.0, n)
    coords[:, 0] = 0.2 * np.sin(np.arange(n))
    k = np.stack((np.diag((4e4, 3e4, 3e4, 60.0, 90.0, 150.0)), np.diag((5e4, 2e4, 4e4, 80.0, 70.0, 110.0))))
    ob.set_design_variables(coords, k, None)
    f = np.zeros((n, 6))
    f[-1] = (3.0, -2.0, 40.0, 5.0, -4.0, 2.0)
    f[3, 2] = -15.0
    ref = ob.static_solve(np.arange(6), f_ext_follower=f, load_steps=2)
    b = _beam_from_oracle(ob, k_cs_index=np.arange(n - 1) % 2)
    out = b.static_solve(np.arange(6), f_ext_follower=f, load_steps=2)
    assert out.iters == ref["iters"]
    assert rel_err(out.hg.cpu().numpy(), ref["hg"]) < 1e-10


def test_dynamic_rigid_rotation(dev):
    """tests/structure/dynamic_primal/test_const_initial_rotational_velocity.py
    (about the beam axis): the device trajectory equals the oracle's and the
    closed form angle = omega t."""
    omega, dt, n_t = 10.0, 0.001, 40
    ob = B.BeamOracle(3, np.array([[0, 1], [1, 2]]), np.array((0.0, 1.0, 0.0)), spectral_radius=1.0)
    coords = np.zeros((3, 3))
    coords[:, 0] = (-4.56, 0.0, 4.56)
    ob.set_design_variables(coords, np.diag(np.full(6, 1e4)), np.eye(6))
    v0 = np.zeros((3, 6))
    v0[:, 3] = omega
    ref = ob.dynamic_solve(dict(hg=ob.hg0, v=v0), None, n_t, dt)
    b = _beam_from_oracle(ob)
    init = b.reference_configuration().to_dynamic()
    init.v = torch.tensor(v0, device=dev)
    out = b.dynamic_solve(init, None, n_t, dt)
    assert rel_err(out.hg.cpu().numpy(), ref["hg"]) < 1e-12
    assert rel_err(out.v.cpu().numpy(), ref["v"]) < 1e-12
    th = omega * dt * np.arange(n_t)
    ang = np.array([B.log_so3(h[1, :3, :3])[0] for h in out.hg.cpu().numpy()])
    assert np.allclose(ang, np.arctan2(np.sin(th), np.cos(th)), atol=1e-10)


def test_dynamic_cantilever_loads(dev):
    """A cantilever released under time-varying follower and dead tip loads:
    same trajectory (frames, velocities, accelerations) as the oracle."""
    n = 6
    ob = B.BeamOracle(n, B.chain(n), np.array((0.0, 0.0, 1.0)), spectral_radius=0.8)
    coords = np.zeros((n, 3))
    coords[:, 1] = np.linspace(0.0, 4.0, n)
    m = np.diag((0.9, 0.9, 0.9, 0.05, 0.03, 0.07))
    ob.set_design_variables(coords, np.diag((1e5, 1e5, 1e5, 300.0, 400.0, 800.0)), m)
    n_t, dt = 25, 0.01
    t = dt * np.arange(n_t)
    ff = np.zeros((n_t, n, 6))
    ff[:, -1, 2] = 20.0 * np.sin(7.0 * t)
    fd = np.zeros((n_t, n, 6))
    fd[:, -1, 0] = 8.0 * t
    fd[:, 2, 4] = 3.0
    ref = ob.dynamic_solve(dict(hg=ob.hg0), np.arange(6), n_t, dt, f_ext_follower=ff, f_ext_dead=fd)
    with B.ulp_shifted_trig():
        pert = ob.dynamic_solve(dict(hg=ob.hg0), np.arange(6), n_t, dt, f_ext_follower=ff, f_ext_dead=fd)
    b = _beam_from_oracle(ob)
    out = b.dynamic_solve(None, np.arange(6), n_t, dt, f_ext_follower=ff, f_ext_dead=fd)
    for name in ("hg", "v", "v_dot", "a"):
        _held(name, getattr(out, name).cpu().numpy(), ref[name], pert[name])
    assert out.hg[-1, -1, 2, 3].abs().item() > 1e-3  # the tip moved


def _coupled(ocw, dev):
    from flapjax_amd import UVLM, Constant, GridDiscretisation
    from flapjax_amd.beam import ConvergenceSettings
    from flapjax_amd.coupled import CoupledAeroelastic

    ob, ou = ocw.beam, ocw.uvlm
    beam = _beam_from_oracle(ob)
    m, n, ms = ou.grid_disc[0]
    uv = UVLM([GridDiscretisation(m, n, ms)], np.arange(ob.n_nodes), mirror_point=np.zeros(3),
              mirror_normal=np.array((0.0, 1.0, 0.0)))
    s = ocw.fsi_settings
    cw = CoupledAeroelastic(beam, uv, ConvergenceSettings(s.max_n_iter, s.rel_disp_tol, s.abs_disp_tol,
                                                          s.rel_force_tol, s.abs_force_tol))
    ff = ou.flowfield
    flow = Constant(ff.u_inf, ff.rho, True)
    cw.set_design_variables(ob.x0, ob.k_cs, ob.m_cs, None, ou.dt, flow, ou.zeta_b0[0])
    return cw


def test_project_forces(dev):
    from flapjax_amd.coupled import project_forcing_to_beam

    rng = np.random.default_rng(5)
    n = 6
    hg = np.stack([B.exp_se3(x) for x in rng.normal(size=(n, 6)) * 0.5])
    x0 = [rng.normal(size=(4, 5, 3)), rng.normal(size=(3, 4, 3))]
    f = [rng.normal(size=(4, 5, 3)), rng.normal(size=(3, 4, 3))]
    g = [rng.normal(size=(4, 5, 3)), rng.normal(size=(3, 4, 3))]
    dm = [np.arange(5), np.array([5, 4, 3, 2])]
    ref = B.project_forcing_to_beam([a + b for a, b in zip(f, g)], hg[:, :3, :3], dm, x0)
    t = lambda a: torch.tensor(a, device=dev)  # noqa: E731
    out = project_forcing_to_beam([t(a) for a in f], t(hg)[:, :3, :3], dm, [t(a) for a in x0],
                                  f_extra=[t(a) for a in g])
    assert rel_err(out.cpu().numpy(), ref) < 1e-14


def test_coupled_static_and_dynamic(dev):
    """Aitken-relaxed static FSI, then FSI time steps from it with a 1-cos
    gust: the device beam + UVLM follow the oracle's iterates."""
    gust = U.OneMinusCosine(np.array((10.0, 0.0, 1.0)), 1.225, True, gust_length=1.0, gust_amplitude=0.5)
    ocw = B.cantilever_wing(m=2, n_nodes=5, m_star=6, k_diag=(1e6, 1e6, 1e6, 1e3, 5e3, 1e5), u_inf=(10.0, 0.0, 1.0),
                            flow=gust)
    n_t = 8

    def oracle_run():
        st_ = ocw.static_solve(np.arange(6))
        init = ocw.uvlm.solve_static(st_["hg"])
        return st_, ocw.dynamic_solve(st_["hg"], init, np.arange(6), n_t)

    ref_st, ref = oracle_run()
    with B.ulp_shifted_trig():
        pst, pdyn = oracle_run()
    from flapjax_amd import OneMinusCosine

    cw = _coupled(ocw, dev)
    cw.aero.flowfield = OneMinusCosine((10.0, 0.0, 1.0), 1.225, True, 1.0, 0.5)
    st = cw.static_solve(np.arange(6))
    assert st.structure.iters["fsi"] == ref_st["iters"]
    _held("static hg", st.structure.hg.cpu().numpy(), ref_st["hg"], pst["hg"])
    _held("static gamma", st.aero.gamma_b[0][0].cpu().numpy(), ref_st["aero"]["gamma_b"][0],
          pst["aero"]["gamma_b"][0])
    dyn = cw.dynamic_solve(st, np.arange(6), n_t)
    assert dyn.structure.iters["fsi"] == ref["info"]["fsi_iters"]
    _held("hg", dyn.structure.hg.cpu().numpy(), ref["hg"], pdyn["hg"])
    _held("gamma_b", dyn.aero.gamma_b[0].cpu().numpy(), ref["aero"].gamma_b[0], pdyn["aero"].gamma_b[0])
    _held("zeta_w", dyn.aero.zeta_w[0].cpu().numpy(), ref["aero"].zeta_w[0], pdyn["aero"].zeta_w[0])
    _held("f_ext_aero", dyn.structure.f_ext_aero.cpu().numpy(), ref["f_ext_aero"], pdyn["f_ext_aero"])


def test_device_newton_loop_equals_host_loop(dev):
    """The dynamic Newton loop with its stop test on the device
    (fj_beam_newton_dynamic: batches of guarded iterations, one status read
    per batch) runs the iterations the host-driven loop runs, to the bit:
    same trajectory and the same iteration count at every step, including a
    loop that stops at the iteration limit."""
    from flapjax_amd.beam import ConvergenceSettings

    n = 6
    ob = B.BeamOracle(n, B.chain(n), np.array((0.0, 0.0, 1.0)), spectral_radius=0.8)
    coords = np.zeros((n, 3))
    coords[:, 1] = np.linspace(0.0, 4.0, n)
    ob.set_design_variables(coords, np.diag((1e5, 1e5, 1e5, 300.0, 400.0, 800.0)),
                            np.diag((0.9, 0.9, 0.9, 0.05, 0.03, 0.07)))
    n_t, dt = 20, 0.01
    t = dt * np.arange(n_t)
    ff = np.zeros((n_t, n, 6))
    ff[:, -1, 2] = 40.0 * np.sin(7.0 * t)
    outs = []
    for conv in (None, ConvergenceSettings(2, None, None, 1e-14, None)):
        for device in (True, False):
            b = _beam_from_oracle(ob)
            if conv is not None:
                b.struct_convergence_settings = conv
            b.device_newton = device
            outs.append(b.dynamic_solve(None, np.arange(6), n_t, dt, f_ext_follower=ff))
    for a, c in ((outs[0], outs[1]), (outs[2], outs[3])):
        assert a.iters["struct"] == c.iters["struct"]
        for name in ("hg", "v", "v_dot"):
            assert torch.equal(getattr(a, name), getattr(c, name)), name
    assert max(outs[0].iters["struct"]) >= 2 and set(outs[2].iters["struct"]) == {2}


def test_ga_nodes_equals_separate_launches(dev):
    """fj_beam_ga_nodes (generalised-alpha state + node frames in one
    launch) gives the bits of fj_beam_ga followed by fj_beam_nodes."""
    from flapjax_amd import _lib as L
    from flapjax_amd.beam import BeamStructure

    rng = np.random.default_rng(9)
    nn = 81
    am, af = 0.3, 0.4
    beam = BeamStructure.__new__(BeamStructure)
    t = lambda *shape: torch.tensor(rng.normal(size=shape), device=dev)  # noqa: E731
    phi_a, v_a, vd_a, v_p, vd_p, a_p = (0.1 * t(nn, 6) for _ in range(6))
    hg_p = torch.tensor(np.stack([B.exp_se3(x) for x in rng.normal(size=(nn, 6))]), device=dev)
    lib = L.lib()
    e = lambda: torch.empty((nn, 6), dtype=torch.float64, device=dev)  # noqa: E731
    # alpha point: phi_n, v_n, copy of phi_a, frames and rates
    o = [e(), e(), e()]
    L.check(lib.fj_beam_ga(L.GA_ALPHA, 6 * nn, am, af, 0.0, 0.0, 0.0, L.ptr(phi_a), L.ptr(v_a), L.ptr(v_p), None,
                           None, None, *[L.ptr(x) for x in o], None, L.stream()), "fj_beam_ga")
    hg_ref = torch.empty((nn, 4, 4), dtype=torch.float64, device=dev)
    beam._nodes(L.BEAM_NODE_UPDATE, hg_p, x=o[0], hg_out=hg_ref)
    rate_ref = beam.make_hg_dot(hg_ref, o[1])
    f = [e(), e(), e()]
    hg_f, rate_f = torch.empty_like(hg_ref), torch.empty_like(hg_ref)
    L.check(lib.fj_beam_ga_nodes(L.GA_ALPHA, nn, am, af, L.ptr(phi_a), L.ptr(v_a), None, L.ptr(v_p), None, None,
                                 L.ptr(hg_p), L.ptr(f[0]), L.ptr(f[1]), None, None, L.ptr(f[2]), L.ptr(hg_f),
                                 L.ptr(rate_f), L.stream()), "fj_beam_ga_nodes")
    for x, y in zip(o + [hg_ref, rate_ref], f + [hg_f, rate_f]):
        assert torch.equal(x, y)
    # step completion: phi, v, v_dot, a and the frames
    o = [e(), e(), e(), e()]
    L.check(lib.fj_beam_ga(L.GA_FINISH, 6 * nn, am, af, 0.0, 0.0, 0.0, L.ptr(phi_a), L.ptr(v_a), L.ptr(vd_a),
                           L.ptr(v_p), L.ptr(vd_p), L.ptr(a_p), *[L.ptr(x) for x in o], L.stream()), "fj_beam_ga")
    beam._nodes(L.BEAM_NODE_UPDATE, hg_p, x=o[0], hg_out=hg_ref)
    f = [e(), e(), e(), e()]
    L.check(lib.fj_beam_ga_nodes(L.GA_FINISH, nn, am, af, L.ptr(phi_a), L.ptr(v_a), L.ptr(vd_a), L.ptr(v_p),
                                 L.ptr(vd_p), L.ptr(a_p), L.ptr(hg_p), *[L.ptr(x) for x in f], None, L.ptr(hg_f),
                                 None, L.stream()), "fj_beam_ga_nodes")
    for x, y in zip(o + [hg_ref], f + [hg_f]):
        assert torch.equal(x, y)


def test_fused_step_end_equals_separate_launches(dev):
    """fj_beam_step_end (step completion, frames, force projection, local
    copy and next predictor in one launch) gives the bits of the six
    separate launches over coupled time steps with a gust."""
    from flapjax_amd import OneMinusCosine
    from flapjax_amd.beam import BeamStructure

    gust = U.OneMinusCosine(np.array((10.0, 0.0, 1.0)), 1.225, True, gust_length=1.0, gust_amplitude=0.5)
    ocw = B.cantilever_wing(m=2, n_nodes=5, m_star=6, k_diag=(1e6, 1e6, 1e6, 1e3, 5e3, 1e5), u_inf=(10.0, 0.0, 1.0),
                            flow=gust)
    runs = []
    for fused in (True, False):
        cw = _coupled(ocw, dev)
        cw.aero.flowfield = OneMinusCosine((10.0, 0.0, 1.0), 1.225, True, 1.0, 0.5)
        cw.structure.fused_step_end = fused
        assert BeamStructure.fused_step_end  # the default
        st = cw.static_solve(np.arange(6))
        runs.append(cw.dynamic_solve(st, np.arange(6), 8))
    a, b = runs
    for name in ("hg", "v", "v_dot", "a", "f_ext_aero"):
        assert torch.equal(getattr(a.structure, name), getattr(b.structure, name)), name
    for name in ("gamma_b", "zeta_w", "f_steady", "f_unsteady"):
        assert torch.equal(getattr(a.aero, name)[0], getattr(b.aero, name)[0]), name
    assert a.structure.iters == b.structure.iters
