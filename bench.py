"""Benchmark of the UVLM hot path on one MI355X.

Workload: unsteady rigid-wing UVLM with free wake -- 16 x 64 bound panels
(chord 1, span 10), 200 wake rows, pitch-plunge motion (mean alpha 5 deg,
pitch 3 deg, plunge 0.05 chord, reduced frequency 0.2), dt = chord / (16 U),
full wake-on-wake rollup every step.  A "step" is one UVLM time step on the
device (rollup + wake convection, AIC, circulation solve, forces).

Metric: Biot-Savart pair interactions per second over the whole step
(4.52e8 de-duplicated (target, filament) pairs per step), for the plain FP64
drop-in (``value``) and the mixed form (FP32 far field + FP64 near field,
FP32 LU + FP64 refinement).  Also reported: per-step p50/p99 latency, cases/s
for a whole 200-step case, the roofline of the dominant kernel (the rollup
pair-sum launch, 13,000 x 27,992 pairs, 57 FLOP per pair, FP64 VALU bound),
and the CPU baseline (the oracle's C restatement of the same pair sum, OpenMP
over all host threads, on a bounded sample).

    python bench.py [--gpus 1] [--steps 40] [--warmup 5]
"""

from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

M, N, M_STAR, SPAN = 16, 64, 200, 10.0
N_TSTEP = 200
ALPHA0, PITCH, PLUNGE, K_RED = 5.0, 3.0, 0.05, 0.2
F_PAIR = 57  # algorithmic FLOP per pair interaction
PEAK_FP64 = 78.6  # TFLOP/s, MI355X FP64 vector (datasheet)
# HBM bytes per FP64 rollup launch (records, partial and reduce kernels of
# the lattice pair sum), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs
# of `bench.py --profile-rollup`, each scaled by the same counter on a
# 256 MiB device copy (FETCH_SIZE reads 0.5x the copied bytes on MI355X,
# WRITE_SIZE 1.0x); re-measured whenever the pair-sum launch shape changes.
ROLLUP_TRAFFIC_BYTES = 47_463_000


def quiesce():
    """Collect, then freeze every object alive so far (gc.freeze) before a
    timed region: a full collection otherwise rescans the interpreter's ~1e6
    long-lived objects (torch, numpy, scipy) and stalls a step every few
    hundred by ~40 ms.  A serving process would freeze after start-up the
    same way; the library itself allocates no more per step than before."""
    gc.collect()
    gc.freeze()


def n_seg(m, n):
    return m * (n + 1) + (m + 1) * n


def pairs_per_step(m=M, n=N, m_star=M_STAR):
    s_b, s_w = n_seg(m, n), n_seg(m_star, n)
    rollup = m_star * (n + 1) * (s_b + s_w)
    forces = s_b * (s_b + s_w)
    upwash = m * n * s_w
    aic = m * n * s_b
    return rollup + forces + upwash + aic, rollup


def pitch_plunge(n, span, n_tstep, dt):
    """hg (n_tstep, n+1, 4, 4) and hg_dot = d(hg)/dt (exact rotation rate
    theta_dot dR/dtheta plus the plunge rate) for the pitch-plunge motion."""
    omega = 2.0 * K_RED  # k = omega c / (2 U), c = U = 1
    t = np.arange(n_tstep) * dt
    th = np.deg2rad(ALPHA0) + np.deg2rad(PITCH) * np.sin(omega * t)
    thd = np.deg2rad(PITCH) * omega * np.cos(omega * t)
    h = PLUNGE * np.sin(omega * t)
    hd = PLUNGE * omega * np.cos(omega * t)
    hg = np.zeros((n_tstep, n + 1, 4, 4))
    c, s = np.cos(th), np.sin(th)
    hg[:, :, 0, 0] = c[:, None]
    hg[:, :, 0, 2] = s[:, None]
    hg[:, :, 1, 1] = 1.0
    hg[:, :, 2, 0] = -s[:, None]
    hg[:, :, 2, 2] = c[:, None]
    hg[:, :, 1, 3] = np.linspace(0.0, span, n + 1)[None, :]
    hg[:, :, 2, 3] = h[:, None]
    hg[:, :, 3, 3] = 1.0
    hgd = np.zeros_like(hg)
    hgd[:, :, 0, 0] = (-s * thd)[:, None]
    hgd[:, :, 0, 2] = (c * thd)[:, None]
    hgd[:, :, 2, 0] = (-c * thd)[:, None]
    hgd[:, :, 2, 2] = (-s * thd)[:, None]
    hgd[:, :, 2, 3] = hd[:, None]
    return hg, hgd


def make_case(precision, aic_reuse=True):
    import flapjax_amd as fj
    from flapjax_amd.uvlm import UVLM
    from flapjax_amd.utils import make_rectangular_grid

    dt = 1.0 / M
    hg, hgd = pitch_plunge(N, SPAN, N_TSTEP, dt)
    uv = UVLM([(M, N, M_STAR)], np.arange(N + 1), free_wake=True, precision=precision, aic_reuse=aic_reuse)
    uv.set_design_variables(dt, fj.Constant(np.array([1.0, 0.0, 0.0]), 1.225, True),
                            make_rectangular_grid(M, N, 1.0, 0.25), hg[0])
    return uv, hg, hgd


def static_bench(reps=20):
    """The steady case (8 x 32 rings, span 10, 5 deg pitch, 20 wake rows,
    dt = c/(8U)): one static solve (geometry, AIC with the lumped wake,
    factorisation and refined solve, forces) from a fresh solver per case,
    event-timed over ``reps`` cases; pairs per case = AIC + forces."""
    import torch

    import flapjax_amd as fj
    from flapjax_amd.uvlm import UVLM
    from flapjax_amd.utils import make_rectangular_grid, to_device

    m, n, ms = 8, 32, 20
    hg = to_device(pitch_plunge(n, SPAN, 1, 1.0 / m)[0][0])

    def case():
        uv = UVLM([(m, n, ms)], np.arange(n + 1), free_wake=False)
        uv.set_design_variables(1.0 / m, fj.Constant(np.array([1.0, 0.0, 0.0]), 1.225, True),
                                make_rectangular_grid(m, n, 1.0, 0.25), hg)
        return uv.solve_static(hg=hg)

    case()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s.record()
    for _ in range(reps):
        case()
    e.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    s_b, s_w = n_seg(m, n), n_seg(ms, n)
    pairs = m * n * (s_b + s_w) + s_b * (s_b + s_w)
    return {"workload": "steady rigid wing, 8x32 rings, 20 wake rows, 5 deg", "ms_per_case": round(1e3 * wall, 4),
            "device_ms_per_case": round(s.elapsed_time(e) / reps, 4), "cases_per_s": round(1.0 / wall, 2),
            "pairs_per_case": pairs}


def time_form(precision, steps, warmup, aic_reuse=True, whole_case=True):
    import torch

    uv, hg_np, hgd_np = make_case(precision, aic_reuse)
    from flapjax_amd.utils import to_device

    hg, hgd = to_device(hg_np), to_device(hgd_np)
    static = uv.solve_static(hg=hg[0])
    case = static.to_dynamic(0, N_TSTEP)
    i = 1

    def step():
        nonlocal case, i
        if i >= N_TSTEP:
            case = static.to_dynamic(0, N_TSTEP)
            i = 1
        uv.case_solve(case, i, hg[i], hg[i - 1], hgd[i], False, False)
        i += 1

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    quiesce()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    t0 = time.perf_counter()
    ev[0].record()
    for k in range(steps):
        step()
        ev[k + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([ev[k].elapsed_time(ev[k + 1]) for k in range(steps)])
    if not whole_case:
        return {"wall_s": wall, "ms_per_step": 1e3 * wall / steps, "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "case_s": float("nan")}
    # whole case (static + 199 dynamic steps) end to end
    torch.cuda.synchronize()
    quiesce()
    tc = time.perf_counter()
    s2 = uv.solve_static(hg=hg[0])
    c2 = s2.to_dynamic(0, N_TSTEP)
    for j in range(1, N_TSTEP):
        uv.case_solve(c2, j, hg[j], hg[j - 1], hgd[j], False, False)
    torch.cuda.synchronize()
    case_s = time.perf_counter() - tc
    return {"wall_s": wall, "ms_per_step": 1e3 * wall / steps, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "case_s": case_s, "case": c2}


FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "cfg1_free_wake_16x64.npz")


def errors_vs_oracle(case):
    """Per-result errors (max |a - b| / max |b|) of a whole 200-step case
    against the oracle's FP64 run of the same workload (the committed
    fixture tests/golden/cfg1_free_wake_16x64.npz), beside the oracle's own
    rounding floor (its run with every pair sum summed in reverse order):
    gamma_b over steps 1-50 (max) and at step 199, and gamma_w, zeta_w,
    f_steady, f_unsteady, alpha and the induced velocity at the collocation
    points at steps 50 and 199."""
    if not os.path.exists(FIXTURE):
        return None
    ref = np.load(FIXTURE)

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    kept = [int(k) for k in ref["steps_kept"]]
    g = case.gamma_b[0][kept].cpu().numpy()
    eg = [rel(g[i], ref["gamma_b"][i]) for i in range(len(kept))]
    early = [i for i, k in enumerate(kept) if k <= 50]
    out = {"gamma_b_steps_1_50": float("%.3g" % max(eg[i] for i in early)),
           "gamma_b_199": float("%.3g" % eg[-1])}
    floor = {"gamma_b_steps_1_50": float("%.3g" % max(ref["floor_gamma_b"][i] for i in early)),
             "gamma_b_199": float("%.3g" % ref["floor_gamma_b"][-1])}
    for st in (int(x) for x in ref["state_steps"]):
        for key in ("gamma_w", "zeta_w", "f_steady", "f_unsteady", "alpha"):
            out[f"{key}_{st}"] = float("%.3g" % rel(getattr(case, key)[0][st].cpu().numpy(), ref[f"{key}_{st}"]))
            floor[f"{key}_{st}"] = float("%.3g" % ref[f"floor_{key}_{st}"])
        v = case.get_v_ind(case.c[0][st], st).cpu().numpy()
        out[f"v_ind_c_{st}"] = float("%.3g" % rel(v, ref[f"v_ind_c_{st}"]))
        floor[f"v_ind_c_{st}"] = float("%.3g" % ref[f"floor_v_ind_c_{st}"])
    return {"err": out, "oracle_floor": floor}


# Coupled gust response: a cantilever beam along y (span 10) of 40 stiffness
# stations, each a pair of two-node SE(3) elements (81 nodes, one per
# spanwise station of the 16 x 80 lattice), mirrored about y = 0, freestream
# U = 10 at 5 deg, 300 wake rows, dt = c/(16U), 1-cos vertical gust of
# 0.05 U with L = 5 chords (the profile spans 2L = 10 chords) reaching the
# wing after 1 chord of travel, 300 steps, from the static aeroelastic
# equilibrium.  Cross-section mass 10 per unit length and unit inertias 10,
# as in the reference's adjoint wing model: at 0.75 per unit length the
# fluid's added mass (pi rho c^2 / 4 = 0.96) exceeds the structure's and
# the unrelaxed partitioned FSI iteration of the dynamic solve diverges.
CPL_M, CPL_N, CPL_M_STAR, CPL_STEPS, CPL_STATIONS = 16, 80, 300, 300, 40
CPL_EI_FLAP, CPL_EI_LAG, CPL_GJ, CPL_EA, CPL_GA = 2e4, 4e6, 1e4, 1e8, 1e8
CPL_MASS, CPL_INERTIA, CPL_U = 10.0, (1.0, 1.0, 1.0), 10.0


def coupled_design(seed=3, stations=CPL_STATIONS):
    """Cross-section tables of the coupled wing: k_cs (stations, 6, 6) with
    each stiffness scaled by 1 + U(-0.1, 0.1) (seeded), m_cs (6, 6).  Local
    axes of an element along +y with y_vector = z: x spanwise, y = global z,
    z = global x, so component 4 is in-plane (lag) and 5 out-of-plane (flap)
    bending."""
    rng = np.random.default_rng(seed)
    base = np.array((CPL_EA, CPL_GA, CPL_GA, CPL_GJ, CPL_EI_LAG, CPL_EI_FLAP))
    k = np.stack([np.diag(base * (1.0 + 0.1 * rng.uniform(-1.0, 1.0, 6))) for _ in range(stations)])
    m = np.diag((CPL_MASS, CPL_MASS, CPL_MASS) + tuple(CPL_MASS * np.array(CPL_INERTIA)))
    return k, m


def make_coupled(precision="fp64", m=CPL_M, n=CPL_N, m_star=CPL_M_STAR, stations=CPL_STATIONS, span=10.0,
                 free_wake=False, aic_reuse=None):
    """The coupled gust model on the device (n elements, n + 1 nodes, two
    elements per stiffness station)."""
    import flapjax_amd as fj
    from flapjax_amd.beam import BeamStructure
    from flapjax_amd.coupled import CoupledAeroelastic
    from flapjax_amd.uvlm import UVLM
    from flapjax_amd.utils import make_rectangular_grid

    nn = n + 1
    conn = np.stack((np.arange(n), np.arange(1, nn)), axis=1)
    beam = BeamStructure(nn, conn, np.array((0.0, 0.0, 1.0)), k_cs_index=np.arange(n) * stations // n)
    coords = np.zeros((nn, 3))
    coords[:, 1] = np.linspace(0.0, span, nn)
    uv = UVLM([(m, n, m_star)], np.arange(nn), mirror_point=np.zeros(3), mirror_normal=np.array((0.0, 1.0, 0.0)),
              free_wake=free_wake, precision=precision, aic_reuse=aic_reuse)
    cw = CoupledAeroelastic(beam, uv)
    a = np.deg2rad(5.0)
    u_inf = CPL_U * np.array((np.cos(a), 0.0, np.sin(a)))
    gust = fj.OneMinusCosine(u_inf, 1.225, True, 5.0, 0.05 * CPL_U, gust_x0=(-11.0, 0.0, 0.0))
    k, mm = coupled_design(stations=stations)
    cw.set_design_variables(coords, k, mm, None, 1.0 / (m * CPL_U), gust, make_rectangular_grid(m, n, 1.0, 0.25))
    return cw


def coupled_bench(precision, steps=CPL_STEPS):
    """Static aeroelastic equilibrium, then ``steps`` coupled time steps
    (FSI iterations of beam Newton solves and UVLM steps) end to end."""
    import torch

    # warm-up on an instance of its own (first-use costs of the beam and
    # coupling kernels and of the caching allocator stay out of the timing)
    warm = make_coupled(precision)
    warm.dynamic_solve(warm.static_solve(np.arange(6)), np.arange(6), 3)
    del warm
    cw = make_coupled(precision)
    torch.cuda.synchronize()
    quiesce()
    t0 = time.perf_counter()
    st = cw.static_solve(np.arange(6))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dyn = cw.dynamic_solve(st, np.arange(6), steps)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fsi = np.asarray(dyn.structure.iters["fsi"])
    tip = dyn.structure.hg[:, -1, 2, 3].cpu().numpy()
    lift = dyn.aero.f_steady[0][..., 2].sum(dim=(1, 2)).cpu().numpy()
    s_b, s_w = n_seg(CPL_M, CPL_N), n_seg(CPL_M_STAR, CPL_N)
    # per FSI iteration (mirrored lattices: 2 x the filaments): AIC, wake
    # upwash at the collocation points, forces at the filament midpoints
    pairs_it = 2 * (CPL_M * CPL_N * s_b + CPL_M * CPL_N * s_w + s_b * (s_b + s_w))
    dyn_s = t2 - t1
    return {"static_s": round(t1 - t0, 3), "static_fsi_iters": int(st.structure.iters["fsi"]),
            "dynamic_s": round(dyn_s, 3), "ms_per_step": round(1e3 * dyn_s / (steps - 1), 3),
            "fsi_iters_mean": round(float(fsi.mean()), 3), "fsi_iters_max": int(fsi.max()),
            "newton_iters_per_step": round(sum(dyn.structure.iters["struct"]) / (steps - 1), 3),
            "pairs_per_fsi_iter": int(pairs_it),
            "aero_gpairs_per_s": round(pairs_it * fsi.sum() / dyn_s / 1e9, 3),
            "cases_per_s": round(1.0 / (t2 - t0), 4), "tip_z_min_max": [round(float(tip.min()), 6),
                                                                        round(float(tip.max()), 6)],
            "lift_max": round(float(lift.max()), 6)}


# checkpoint 0: the whole 100-step history stays in HBM (no recomputed
# forward; peak reported in the line)
ADJ_STEPS, ADJ_TIP_WEIGHT, ADJ_CHECKPOINT = 100, 1.0, 0
ADJ_CHECK_STEPS, ADJ_CHECK_NEWTON, ADJ_CHECK_FSI = 10, 8, 10


def _adjoint_case(steps, fixed_iterations):
    """The coupled gust model, its differentiable rerun and the production
    objective as a function of (stiffness scalings, twist angles)."""
    import torch

    from flapjax_amd.adjoint import twisted_grid
    from flapjax_amd.beam import ConvergenceSettings
    from flapjax_amd.coupled_adjoint import DifferentiableAeroelastic

    # the check factors every circulation system afresh (as jnp.linalg.solve
    # does): the kept-inverse refinement stops at dsgesv's backward-error
    # bound, which leaves ~cond(A) sqrt(N) eps (~1e-8 here) of noise in the
    # circulation, too much for differences
    cw = make_coupled("fp64", aic_reuse=not fixed_iterations)
    if fixed_iterations:
        # zero tolerances: every Newton and FSI loop runs its full count, far
        # past convergence, so the objective is smooth for the differences
        cw.structure.struct_convergence_settings = ConvergenceSettings(ADJ_CHECK_NEWTON, 0.0, 0.0, 0.0, 0.0)
        cw.fsi_convergence_settings = ConvergenceSettings(ADJ_CHECK_FSI, 0.0, 0.0, 0.0, 0.0)
    zeta0 = cw.aero.zeta_b0[0]
    model = DifferentiableAeroelastic(cw, np.arange(6), checkpoint_every=ADJ_CHECKPOINT)

    def production(k_, t_):
        z0 = twisted_grid(zeta0, torch.cat((torch.zeros(1, dtype=torch.float64, device="cuda"), t_))).contiguous()
        cw.aero.set_design_variables(cw.aero.dt, cw.aero.flowfield, z0, cw.structure.hg0)
        cw.structure.set_stiffness_scaling(k_)
        st = cw.static_solve(np.arange(6))
        dyn = cw.dynamic_solve(st, np.arange(6), steps)
        f = dyn.aero.f_steady[0] + dyn.aero.f_unsteady[0]
        zt = dyn.structure.hg[:, -1, 2, 3]
        return float(cw.aero.dt * (f[1:, ..., 2].sum(dim=(1, 2)) + ADJ_TIP_WEIGHT * (zt[1:] - zt[0]) ** 2).sum())

    return model, production


def adjoint_bench(steps=ADJ_STEPS, n_dir=2, seed=5, check=True):
    """Gradient of J = dt sum_n [lift_n + w (z_tip,n - z_tip,0)^2] over the
    coupled gust case shortened to ``steps`` steps, with respect to the 40
    stiffness scalings and 80 spanwise twist angles: one reverse pass over
    the whole history (kept in HBM) at the solvers' default tolerances, timed
    end to end (after a first, untimed-for-the-ratio call) beside the
    production forward solve.  The check: the same case
    over its first 10 steps with every Newton and FSI loop run a fixed number
    of iterations, gradient against fourth-order central differences of the
    production solve along ``n_dir`` seeded random directions."""
    import torch

    ks = torch.ones(CPL_STATIONS, dtype=torch.float64, device="cuda")
    tw = torch.zeros(CPL_N, dtype=torch.float64, device="cuda")
    model, production = _adjoint_case(steps, False)
    # the first gradient also captures the beam pullback graphs and grows the
    # caching allocator to the history's footprint: timed on its own, then
    # the gradient is timed warm, as the production forward is
    torch.cuda.synchronize()
    quiesce()
    t0 = time.perf_counter()
    model.gradient(ks, tw, steps, ADJ_TIP_WEIGHT)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    quiesce()
    t0 = time.perf_counter()
    J, g_k, g_t = model.gradient(ks, tw, steps, ADJ_TIP_WEIGHT)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    quiesce()
    t1 = time.perf_counter()
    J_prod = production(ks, tw)
    t_fwd = time.perf_counter() - t1
    out = {"workload": f"coupled gust case, {steps} steps; d/d(40 stiffness scalings, 80 twist angles) of "
                       f"dt sum (lift + {ADJ_TIP_WEIGHT} (z_tip - z_tip0)^2), default solver tolerances",
           "gradient_s": round(t_grad, 3), "gradient_first_call_s": round(t_first, 3),
           "production_forward_s": round(t_fwd, 3),
           "gradient_over_forward": round(t_grad / t_fwd, 2), "J": float(J),
           "J_production_rel_diff": float("%.3g" % (abs(float(J) - J_prod) / abs(J_prod))),
           "grad_norm_k": float(g_k.norm()), "grad_norm_twist": float(g_t.norm()),
           "checkpoint_every": ADJ_CHECKPOINT, "fsi_iters_mean": round(float(np.mean(model.fsi_iters)), 3)}
    if not check:
        out["peak_mem_gb"] = round(torch.cuda.max_memory_allocated() / 1e9, 3)
        return out
    model, production = _adjoint_case(ADJ_CHECK_STEPS, True)
    t2 = time.perf_counter()
    J, g_k, g_t = model.gradient(ks, tw, ADJ_CHECK_STEPS, ADJ_TIP_WEIGHT)
    torch.cuda.synchronize()
    t_chk = time.perf_counter() - t2
    # noise floor of the production objective: two designs 1e-13 apart
    j_a, j_b = production(ks, tw), production(ks * (1.0 + 1e-13), tw)
    noise = abs(j_a - j_b) / abs(j_a)
    rng = np.random.default_rng(seed)
    errs = []
    for _ in range(n_dir):
        dk = torch.as_tensor(rng.normal(size=CPL_STATIONS), device="cuda")
        dt_ = torch.as_tensor(rng.normal(size=CPL_N), device="cuda")
        hk, ht = 1e-2, 1e-3
        at = lambda s: production(ks + s * hk * dk, tw + s * ht * dt_)  # noqa: E731
        fd = (-at(2) + 8 * at(1) - 8 * at(-1) + at(-2)) / 12.0
        an = float((g_k * dk).sum()) * hk + float((g_t * dt_).sum()) * ht
        errs.append(abs(fd - an) / max(abs(fd), 1e-300))
    out["check"] = {"steps": ADJ_CHECK_STEPS, "newton_iters": ADJ_CHECK_NEWTON, "fsi_iters": ADJ_CHECK_FSI,
                    "gradient_s": round(t_chk, 3), "J_production_rel_diff":
                        float("%.3g" % (abs(float(J) - production(ks, tw)) / abs(float(J)))),
                    "fd_rel_err": [float("%.3g" % e) for e in errs], "J_noise_rel": float("%.3g" % noise),
                    "note": "differences resolve the gradient only to ~J_noise / h: the unsteady force "
                            "amplifies circulation rounding by relax/dt, which keeps the production "
                            "objective noisy at this size"}
    out["peak_mem_gb"] = round(torch.cuda.max_memory_allocated() / 1e9, 3)
    return out


def rollup_roofline(reps=20):
    """Time the dominant launch alone (rollup pair sum, 13,000 x 27,992) with
    HIP events on the stream it runs on."""
    import torch

    from flapjax_amd import _kernels as K
    from flapjax_amd import _lib as L
    from flapjax_amd.utils import to_device

    rng = np.random.default_rng(0)
    uv, hg_np, _ = make_case("fp64")
    zb = uv.zeta_b_ref[0]
    zw = uv.zeta_w_ref[0].clone()
    zw += to_device(rng.normal(size=zw.shape) * 0.01 * np.array([0.0, 0.0, 1.0]))
    gb = to_device(rng.uniform(-0.1, 0.1, (M, N)))
    gw = to_device(rng.uniform(-0.1, 0.1, (M_STAR, N)))
    seg = torch.cat([K.segments_from_grid(zb, gb, L.KERNEL_EPSILON, 1e-7),
                     K.segments_from_grid(zw, gw, L.KERNEL_EPSILON, 1e-7)])
    tgt = zw[:-1].reshape(-1, 3).contiguous()
    out = torch.empty_like(tgt)
    # the table holds the bound and the wake lattice, as in the step
    lats = [(0, M, N), (M * (N + 1) + (M + 1) * N, M_STAR, N)]
    res = {}
    for name, prec in (("fp64", L.PREC_FP64), ("mixed", L.PREC_MIXED)):
        ws = torch.empty(K.pair_workspace_bytes(tgt.shape[0], seg.shape[0], prec, lats) // 8 + 1,
                         dtype=torch.float64, device=tgt.device)
        for _ in range(3):
            K.pair_sum(seg, tgt, L.KERNEL_EPSILON, 1e-7, 1e-4, prec, out=out, workspace=ws, lattices=lats)
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            K.pair_sum(seg, tgt, L.KERNEL_EPSILON, 1e-7, 1e-4, prec, out=out, workspace=ws, lattices=lats)
        e.record()
        torch.cuda.synchronize()
        ms = s.elapsed_time(e) / reps
        pairs = tgt.shape[0] * seg.shape[0]
        res[name] = {"ms": ms, "pairs": pairs, "gpairs_s": pairs / ms / 1e6,
                     "tflops": pairs * F_PAIR / (ms * 1e-3) / 1e12}
    return res, (tgt.cpu().numpy(), seg.cpu().numpy())


def pair_sweep(exps=range(12, 18), reps=3):
    """Square pair-sum sweep T = S = 2^k (seeded random filaments of length
    ~0.05 in a unit box), FP64 and mixed: Gpairs/s per size."""
    import torch

    from flapjax_amd import _kernels as K
    from flapjax_amd import _lib as L

    rng = np.random.default_rng(0)
    res = {}
    for k in exps:
        n = 1 << k
        y0 = rng.uniform(0.0, 1.0, (n, 3))
        ends = torch.as_tensor(np.stack((y0, y0 + 0.05 * rng.normal(size=(n, 3))), 1), device="cuda")
        seg = K.segments_from_endpoints(ends, torch.as_tensor(rng.uniform(-1, 1, n), device="cuda"),
                                        L.KERNEL_EPSILON, 1e-7)
        tgt = torch.as_tensor(rng.uniform(0.0, 1.0, (n, 3)), device="cuda")
        out = torch.empty_like(tgt)
        row = {}
        for name, prec in (("fp64", L.PREC_FP64), ("mixed", L.PREC_MIXED)):
            K.pair_sum(seg, tgt, L.KERNEL_EPSILON, 1e-7, 1e-4, prec, out=out)
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            for _ in range(reps):
                K.pair_sum(seg, tgt, L.KERNEL_EPSILON, 1e-7, 1e-4, prec, out=out)
            e.record()
            torch.cuda.synchronize()
            row[name] = round(n * n * reps / (s.elapsed_time(e) * 1e-3) / 1e9, 2)
        res[f"2^{k}"] = row
    return res


def convection_bench(m_star=4000, n=1000, reps=20):
    """Prescribed-wake convection step (circulation shift + TE row +
    x + dt U for every wake vertex) on a large synthetic wake: algorithmic
    bytes 48 m*(n+1) + 24(n+1) + 16 m* n per step against HBM."""
    import torch

    import flapjax_amd as fj
    from flapjax_amd import _kernels as K

    dev = "cuda"
    rng = np.random.default_rng(3)
    zb = torch.as_tensor(rng.normal(size=(2, n + 1, 3)), device=dev)
    gb = torch.as_tensor(rng.normal(size=(1, n)), device=dev)
    zw = [torch.as_tensor(rng.normal(size=(m_star + 1, n + 1, 3)), device=dev) for _ in range(2)]
    gw = [torch.as_tensor(rng.normal(size=(m_star, n)), device=dev) for _ in range(2)]
    flow = fj.Constant(np.array([1.0, 0.0, 0.0]), 1.225, True)

    def step(a, b):
        K.wake_shift(gb, gw[a], zb, m_star, gw[b], zw[b])
        K.wake_convect(zw[a], None, flow, 0.0, 0.01, zw[b])

    step(0, 1)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for r in range(reps):
        step(r % 2, 1 - r % 2)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / reps
    nbytes = 48 * m_star * (n + 1) + 24 * (n + 1) + 16 * m_star * n
    return {"wake": f"{m_star} rows x {n} panels", "us_per_step": round(1e3 * ms, 2),
            "bytes_per_step": nbytes, "gb_s": round(nbytes / (ms * 1e-3) / 1e9, 1), "peak_gb_s": 8000}


def profile_rollup(reps=10, copy_mib=256):
    """Counter-collection workload: the rollup pair-sum launch alone, ``reps``
    times per form, then ``reps`` plain device copies of ``copy_mib`` MiB (the
    scale for the HBM byte counters)."""
    import torch

    roof, _ = rollup_roofline(reps)
    src = torch.ones(copy_mib << 17, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"rollup_ms": roof["fp64"]["ms"], "rollup_mixed_ms": roof["mixed"]["ms"],
                      "pairs": roof["fp64"]["pairs"], "copy_bytes": src.numel() * 8, "reps": reps}))


def profile_kernels(reps=5, copy_mib=256):
    """Counter-collection workload for the kernels the roofline figures are
    quoted on, each a fixed number of times: the rollup pair-sum launch (FP64
    and mixed), the FP64 LU factorisation at n = 1024 and 4096, the FP32 LU
    and the explicit inverse at n = 1024, the prescribed-wake convection step
    on a 4000 x 1000 wake, and plain device copies of ``copy_mib`` MiB (the
    scale for the HBM byte counters).  Prints the event-timed mean of each."""
    import torch

    from flapjax_amd import _kernels as K

    out = {}
    roof, _ = rollup_roofline(reps)
    out["rollup_ms"] = roof["fp64"]["ms"]
    out["rollup_mixed_ms"] = roof["mixed"]["ms"]
    rng = np.random.default_rng(4)

    def timed(fn, n_rep):
        fn()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(n_rep):
            fn()
        e.record()
        torch.cuda.synchronize()
        return s.elapsed_time(e) / n_rep

    for n, n_rep in ((1024, reps), (4096, 2)):
        A = torch.as_tensor(rng.normal(size=(n, n)) + n ** 0.5 * np.eye(n), device="cuda")
        ms = timed(lambda: K.lu_factor(A), n_rep)
        out[f"lu_f64_{n}_ms"] = ms
        out[f"lu_f64_{n}_tflops"] = 2.0 / 3.0 * n ** 3 / (ms * 1e-3) / 1e12
        if n == 1024:
            A32 = A.float()
            out["lu_f32_1024_ms"] = timed(lambda: K.lu_factor_f32(A32), n_rep)
            fac = K.lu_factor(A)
            out["inverse_f64_1024_ms"] = timed(lambda: K.lu_inverse(fac), n_rep)
    out["convection"] = convection_bench(reps=4 * reps)
    src = torch.ones(copy_mib << 17, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2 * reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    out["copy_bytes"] = src.numel() * 8
    print(json.dumps(out))


def profile_adjoint(reps=10):
    """Counter-collection workload for the adjoint pair sum at the size of
    the coupled gust case's force pullback: the 2,656 filament midpoints of
    the 16 x 80 bound lattice against the bound and 300-row wake lattices and
    their mirror images about y = 0 (102,072 filaments), seeded circulations
    and upstream weights; ``reps`` event-timed reverse passes (all four
    gradients) after one warm-up."""
    import torch

    from flapjax_amd import autograd as AG
    from flapjax_amd.utils import make_rectangular_grid

    dev = torch.device("cuda")
    rng = np.random.default_rng(6)
    zb = torch.as_tensor(make_rectangular_grid(CPL_M, CPL_N, 1.0, 0.25), device=dev)
    zb = zb.clone()
    zb[..., 1] = torch.linspace(0.0, 10.0, CPL_N + 1, dtype=torch.float64, device=dev)[None, :]
    te = zb[-1]
    s = torch.arange(1, CPL_M_STAR + 1, dtype=torch.float64, device=dev) / CPL_M
    zw = torch.cat((te[None], te[None] + s[:, None, None] * torch.tensor([1.0, 0.0, 0.0], device=dev,
                                                                         dtype=torch.float64)), 0)
    gb = torch.as_tensor(rng.normal(size=(CPL_M, CPL_N)), device=dev)
    gw = torch.as_tensor(rng.normal(size=(CPL_M_STAR, CPL_N)), device=dev)
    mirror = (np.zeros(3), np.array([0.0, 1.0, 0.0]))
    ys0, ys1, gs = [], [], []
    for z, g in ((zb, gb), (zw, gw)):
        for zz in (z, AG.mirror_grid(z, mirror)):
            a, b, c = AG.grid_filaments(zz, g)
            ys0.append(a)
            ys1.append(b)
            gs.append(c)
    y0, y1, g = torch.cat(ys0), torch.cat(ys1), torch.cat(gs)
    mid = torch.cat((0.5 * (zb[:-1] + zb[1:]).reshape(-1, 3), 0.5 * (zb[:, :-1] + zb[:, 1:]).reshape(-1, 3)))
    w = torch.as_tensor(rng.normal(size=mid.shape), device=dev)
    leaves = [t.detach().clone().requires_grad_(True) for t in (mid, y0, y1, g)]

    def once():
        v = AG.pair_sum(*leaves)
        return torch.autograd.grad((v * w).sum(), leaves)

    once()
    torch.cuda.synchronize()
    st, en = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    st.record()
    for _ in range(reps):
        once()
    en.record()
    torch.cuda.synchronize()
    pairs = mid.shape[0] * y0.shape[0]
    ms = st.elapsed_time(en) / reps
    print(json.dumps({"targets": mid.shape[0], "filaments": y0.shape[0], "pairs": pairs, "reps": reps,
                      "forward_plus_reverse_ms": round(ms, 4)}))


def serve(precision, n_cases, rate, m_star_cap, steps_range):
    """Served mixed-size stream (continuous batching), one GPU."""
    import torch

    from flapjax_amd.serving import ServingEngine, poisson_stream

    reqs = poisson_stream(n_cases, rate, seed=2, m_star=m_star_cap, steps_range=steps_range)
    # warm-up on an engine of its own, so that the measured engine starts
    # with an empty inverse cache
    ServingEngine(precision=precision, max_slots=64).run(
        poisson_stream(4, 1e6, seed=5, m_star=8, steps_range=(5, 8)))
    torch.cuda.synchronize()
    eng = ServingEngine(precision=precision, max_slots=64)
    quiesce()
    _, stats = eng.run(reqs)
    return stats


def progress(msg):
    """A progress line on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _timed(fn, min_seconds, max_reps=50):
    t0 = time.perf_counter()
    reps = 0
    while True:
        fn()
        reps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds or reps >= max_reps:
            return reps, el


def cpu_baseline(sample, min_seconds=10.0):
    """The oracle's C restatement of the FP64 pair sum on all host threads
    (``value``), plus the same at one thread and the oracle's chunked NumPy
    pair sum on a smaller sample of the same targets."""
    from oracle import c_oracle
    from oracle import uvlm_oracle as O

    tgt, seg = sample
    y0, y1, g = seg[:, 0:3], seg[:, 3:6], seg[:, 10]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    c_oracle.set_threads(threads)
    reps, el = _timed(lambda: c_oracle.pair_sum_eps(tgt, y0, y1, g), min_seconds)
    pairs = reps * tgt.shape[0] * y0.shape[0]
    out = {"value": pairs / el / 1e9, "unit": "Gpairs/s", "cores": threads, "kind": "port",
           "sample": f"{reps} x rollup pair sum of one step ({tgt.shape[0]} targets x {y0.shape[0]} filaments), "
                     f"FP64 C (gcc -O3 -fopenmp), {el:.1f} s"}
    # one thread, 1/13 of the targets
    t1 = tgt[::13]
    c_oracle.set_threads(1)
    reps, el = _timed(lambda: c_oracle.pair_sum_eps(t1, y0, y1, g), min_seconds / 4)
    c_oracle.set_threads(threads)
    out["single_core"] = {"value": reps * t1.shape[0] * y0.shape[0] / el / 1e9, "cores": 1,
                          "sample": f"{reps} x {t1.shape[0]} targets x {y0.shape[0]} filaments, {el:.1f} s"}
    # the oracle's NumPy form (chunked targets x filaments, as the reference's lax.map rows)
    tn = tgt[::130]
    reps, el = _timed(lambda: O.pair_sum(tn, y0, y1, g, "epsilon"), min_seconds / 4)
    out["numpy"] = {"value": reps * tn.shape[0] * y0.shape[0] / el / 1e9,
                    "sample": f"{reps} x {tn.shape[0]} targets x {y0.shape[0]} filaments, {el:.1f} s"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-cases", type=int, default=256)
    ap.add_argument("--serve-steps", type=int, nargs=2, default=(50, 400))
    ap.add_argument("--serve-full-rate", type=float, default=2.0,
                    help="arrival rate of the stream with wake rows = steps (below its ~2.5 cases/s capacity)")
    ap.add_argument("--serve-full-mixed", action="store_true",
                    help="also serve the wake rows = steps stream in the mixed form (another ~2 minutes)")
    ap.add_argument("--serve-capped-rate", type=float, default=30.0,
                    help="arrival rate of the capped-wake stream below its capacity (~44 cases/s)")
    ap.add_argument("--serve-wake-cap", type=int, default=32,
                    help="wake rows per case of the capped stream = min(steps, cap)")
    ap.add_argument("--profile-rollup", action="store_true",
                    help="only the rollup launch and a copy of known size (for counter runs)")
    ap.add_argument("--coupled-only", action="store_true",
                    help="only the coupled gust response (beam + UVLM FSI), both precision forms")
    ap.add_argument("--coupled-steps", type=int, default=CPL_STEPS)
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the coupled adjoint gradient (shortened gust case)")
    ap.add_argument("--adjoint-steps", type=int, default=ADJ_STEPS)
    ap.add_argument("--adjoint-no-check", action="store_true",
                    help="with --adjoint-only: the gradient and the production forward only (for profiling)")
    ap.add_argument("--profile-adjoint", action="store_true",
                    help="only the adjoint pair sum at the coupled force-pullback size (for counter runs)")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="rollup, LU, inverse, convection and copy kernels a fixed number of times (for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the UVLM path here runs on a single MI355X "
              "(multi-GPU sharding is a design note only); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.profile_rollup:
        profile_rollup()
        return
    if args.profile_kernels:
        profile_kernels()
        return
    if args.profile_adjoint:
        profile_adjoint()
        return
    if args.coupled_only:
        print(json.dumps({"coupled_gust": {p: coupled_bench(p, args.coupled_steps) for p in ("fp64", "mixed")}}))
        return
    if args.adjoint_only:
        print(json.dumps({"coupled_adjoint": adjoint_bench(args.adjoint_steps, check=not args.adjoint_no_check)}))
        return
    pairs, rollup_pairs = pairs_per_step()
    progress("cfg1 step, fp64")
    f64 = time_form("fp64", args.steps, args.warmup)
    progress("cfg1 step, mixed")
    mix = time_form("mixed", args.steps, args.warmup)
    par_f64 = errors_vs_oracle(f64.pop("case"))
    par_mix = errors_vs_oracle(mix.pop("case"))
    lit = time_form("fp64", max(3, args.steps // 4), 1, aic_reuse=False, whole_case=False)
    roof, sample = rollup_roofline()
    srv = {}
    if args.serve_cases > 0:
        def row(st):
            return {"cases_per_s": round(st["cases_per_s"], 3), "p50_s": round(st["p50_s"], 4),
                    "p99_s": round(st["p99_s"], 4), "gpairs_per_s": round(st["gpairs_per_s"], 3),
                    "makespan_s": round(st["makespan_s"], 3)}

        base = (f"{args.serve_cases} Poisson arrivals (seed 2), M in {{4,8,16,32}}, N in {{16,32,64,128}}, "
                f"{args.serve_steps[0]}-{args.serve_steps[1]} steps, pitch 0-5 deg, k 0.05-0.5, free wake")
        full = {"stream": base + f", wake rows = steps, arrivals at {args.serve_full_rate}/s (below capacity: "
                                 "cases/s is arrival-bound, p50/p99 are service latencies)"}
        for prec in ("fp64", "mixed") if args.serve_full_mixed else ("fp64",):
            progress(f"serving: wake rows = steps, {prec}")
            full[prec] = row(serve(prec, args.serve_cases, args.serve_full_rate, None, tuple(args.serve_steps)))
        cap = {"stream": base + f", wake rows = min(steps, {args.serve_wake_cap}) (a capped stream)"}
        for prec in ("fp64", "mixed"):
            progress(f"serving: capped wake, saturated, {prec}")
            cap[prec + "_saturated"] = row(serve(prec, args.serve_cases, 1e4, args.serve_wake_cap,
                                                 tuple(args.serve_steps)))
        progress("serving: capped wake, below capacity, fp64")
        cap[f"fp64_at_{args.serve_capped_rate:g}_per_s"] = row(
            serve("fp64", args.serve_cases, args.serve_capped_rate, args.serve_wake_cap, tuple(args.serve_steps)))
        cap["mixed_speedup_vs_fp64"] = round(cap["mixed_saturated"]["cases_per_s"]
                                             / cap["fp64_saturated"]["cases_per_s"], 3)
        srv = {"wake_rows_eq_steps": full, "capped_wake": cap}
    progress("steady case")
    steady = static_bench()
    progress("coupled gust")
    coupled = {p: coupled_bench(p, args.coupled_steps) for p in ("fp64", "mixed")} if args.coupled_steps > 0 else None
    progress("coupled adjoint")
    adjoint = adjoint_bench(args.adjoint_steps) if args.adjoint_steps > 0 else None
    progress("pair sweep, convection, cpu baseline")
    sweep = pair_sweep()
    conv = convection_bench()
    cpu = None if args.no_cpu else cpu_baseline(sample)
    value = pairs * args.steps / f64["wall_s"] / 1e9
    value_mixed = pairs * args.steps / mix["wall_s"] / 1e9
    line = {
        "metric": "Biot-Savart Gpairs/s (UVLM free-wake step, plain FP64 drop-in)",
        "value": round(value, 3),
        "unit": "Gpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f64["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded pitch-plunge rigid wing)",
        "config": {"workload": "unsteady rigid-wing UVLM, free wake: 16x64 bound panels, 200 wake rows, "
                               "pitch-plunge k=0.2, dt=c/(16U)", "pairs_per_step": pairs,
                   "parallelism": "single GPU"},
        "latency_ms": {"p50": round(f64["p50_ms"], 4), "p99": round(f64["p99_ms"], 4)},
        "cases_per_s": round(1.0 / f64["case_s"], 3),
        "fp64_refactor_each_step": {"ms_per_step": round(lit["ms_per_step"], 4),
                                    "value": round(pairs / lit["ms_per_step"] / 1e6, 3)},
        "parity": {"fp64": par_f64, "note": "errors of a whole 200-step case against the oracle's FP64 run "
                   "(max |a - b| / max |b|); oracle_floor: the oracle against itself with the pair sums summed "
                   "in reverse order -- the free-wake rollup amplifies rounding past step ~50"},
        "mixed": {"value": round(value_mixed, 3), "unit": "Gpairs/s", "dtype": "f32+f64",
                  "ms_per_step": round(mix["ms_per_step"], 4),
                  "latency_ms": {"p50": round(mix["p50_ms"], 4), "p99": round(mix["p99_ms"], 4)},
                  "cases_per_s": round(1.0 / mix["case_s"], 3), "speedup_vs_fp64": round(value_mixed / value, 3),
                  "err": par_mix["err"] if par_mix else None},
        "roofline": {"bound": "valu", "achieved": round(roof["fp64"]["tflops"], 3), "peak": PEAK_FP64,
                     "unit": "TFLOP/s", "frac": round(roof["fp64"]["tflops"] / PEAK_FP64, 4), "traffic": ROLLUP_TRAFFIC_BYTES,
                     "kernel": "k_lattice_records + k_lattice_partial + k_pair_reduce (rollup pair sum)",
                     "launch_ms": round(roof["fp64"]["ms"], 4), "pairs_per_launch": roof["fp64"]["pairs"],
                     "mixed_launch_ms": round(roof["mixed"]["ms"], 4)},
        "serving": srv,
        "steady_case": steady,
        "coupled_gust": coupled,
        "coupled_adjoint": adjoint,
        "pair_sweep_gpairs_s": sweep,
        "convection": conv,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
