"""Writes the golden fixtures of this directory from the FP64 oracle.

The reference (JAX, Python >= 3.12) cannot run in this environment, so the
fixtures are outputs of oracle/uvlm_oracle.py, whose own pins are the tests in
tests/test_oracle_*.py.  They freeze the oracle's results for the two
configurations the GPU parity tests use most, so a change of the oracle is
caught on the CPU and the GPU tests can compare without re-running it.

    python tests/golden/make_golden.py
"""

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from oracle import uvlm_oracle as O  # noqa: E402


def cfg0_static():
    """8 x 32 rigid wing, chord 1, span 10, 5 deg, 20 wake rows, U = (1,0,0)."""
    m, n, m_star = 8, 32, 20
    zeta0 = O.make_rectangular_grid(m, n, 1.0, 0.25)
    hg = O.rigid_wing_hg(n, 10.0, np.deg2rad(5.0))
    o = O.UVLMOracle([(m, n, m_star)], np.arange(n + 1))
    o.set_design_variables(1.0 / m, O.Constant(np.array([1.0, 0.0, 0.0]), 1.225), zeta0, hg)
    c = o.solve_static()
    return {"zeta0": zeta0, "hg": hg, "gamma_b": c.gamma_b[0][0], "gamma_w": c.gamma_w[0][0],
            "f_steady": c.f_steady[0][0], "alpha": c.alpha[0][0], "zeta_w": c.zeta_w[0][0]}


def small_free_wake():
    """8 x 16, 12 wake rows, free wake, pitch-plunge, 10 steps."""
    m, n, m_star, steps = 8, 16, 12, 10
    dt = 1.0 / m
    zeta0 = O.make_rectangular_grid(m, n, 1.0, 0.25)
    hg_t, hgd_t = O.pitch_plunge_motion(n, 4.0, steps, dt, np.deg2rad(5.0), np.deg2rad(3.0), 0.05, 0.4)
    o = O.UVLMOracle([(m, n, m_star)], np.arange(n + 1), free_wake=True)
    o.set_design_variables(dt, O.Constant(np.array([1.0, 0.0, 0.0]), 1.225), zeta0, hg_t[0])
    d = o.solve_prescribed_dynamic(o.solve_static(hg=hg_t[0]), hg_t, hgd_t)
    return {"zeta0": zeta0, "hg_t": hg_t, "hgd_t": hgd_t, "gamma_b": d.gamma_b[0], "gamma_w": d.gamma_w[0],
            "f_steady": d.f_steady[0], "f_unsteady": d.f_unsteady[0], "zeta_w": d.zeta_w[0][-1]}


CFG1_STEPS_KEPT = (1, 10, 25, 40, 50, 60, 80, 100, 150, 199)
CFG1_STATE_STEPS = (50, 199)


def _cfg1_run(reverse):
    """One oracle run of the full unsteady free-wake workload; ``reverse``
    sums every pair sum over the filaments in reverse order (the same
    arithmetic with a different rounding)."""
    from oracle import c_oracle

    def ps(t, y0, y1, g, kernel="epsilon", chunk=256):
        assert kernel == "epsilon"
        if reverse:
            return c_oracle.pair_sum_eps(t, y0[::-1], y1[::-1], g[::-1])
        return c_oracle.pair_sum_eps(t, y0, y1, g)

    m, n, m_star, steps, span = 16, 64, 200, 200, 10.0
    dt = 1.0 / m
    zeta0 = O.make_rectangular_grid(m, n, 1.0, 0.25)
    hg_t, hgd_t = O.pitch_plunge_motion(n, span, steps, dt, np.deg2rad(5.0), np.deg2rad(3.0), 0.05, 0.4)
    O.set_pair_sum(ps)
    try:
        o = O.UVLMOracle([(m, n, m_star)], np.arange(n + 1), free_wake=True)
        o.set_design_variables(dt, O.Constant(np.array([1.0, 0.0, 0.0]), 1.225), zeta0, hg_t[0])
        d = o.solve_prescribed_dynamic(o.solve_static(hg=hg_t[0]), hg_t, hgd_t)
        res = {"gamma_b": d.gamma_b[0][list(CFG1_STEPS_KEPT)]}
        for s in CFG1_STATE_STEPS:
            v_c = O.compute_v_ind(d.c[0][s], [d.zeta_b[0][s], d.zeta_w[0][s]], [d.gamma_b[0][s], d.gamma_w[0][s]],
                                  "epsilon")
            for key in ("gamma_w", "zeta_w", "f_steady", "f_unsteady", "alpha"):
                res[f"{key}_{s}"] = getattr(d, key)[0][s]
            res[f"v_ind_c_{s}"] = v_c
    finally:
        O.set_pair_sum(None)
    return res


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def cfg1_free_wake():
    """The full unsteady free-wake workload: 16 x 64 rings, chord 1, span 10,
    200 wake rows from the static solve, free wake, dt = 1/16, 200 steps of
    pitch 5 + 3 sin(0.4 t) deg and plunge 0.05 sin(0.4 t) (k = 0.2) with the
    exact hg_dot.  The epsilon-kernel pair sums run in the oracle's C
    restatement (same formula, OpenMP), the rest in NumPy.

    The run is made twice, the second time with every pair sum summed over
    the filaments in reverse order.  The two differ only by rounding, and
    their difference ("floor_*": per kept step for gamma_b, per state step
    for the rest) is what two correct FP64 evaluations can agree to: the
    free-wake rollup amplifies rounding (wake vertices approach each other
    in the tip roll-up, and the epsilon kernel's core is 1e-7), from ~1e-16
    at step 10 to ~1e-10 at step 50 and ~1e-2 at step 199 on the wake
    vertices.  Kept: gamma_b at CFG1_STEPS_KEPT, and gamma_w, zeta_w,
    f_steady, f_unsteady, alpha and the induced velocity at the collocation
    points (bound + wake, compute_v_ind) at CFG1_STATE_STEPS.  About 20
    minutes on 8 cores."""
    ref = _cfg1_run(False)
    alt = _cfg1_run(True)
    out = dict(ref)
    out["steps_kept"] = np.array(CFG1_STEPS_KEPT)
    out["state_steps"] = np.array(CFG1_STATE_STEPS)
    out["floor_gamma_b"] = np.array([_rel(alt["gamma_b"][k], ref["gamma_b"][k]) for k in range(len(CFG1_STEPS_KEPT))])
    for key in ref:
        if key != "gamma_b":
            out[f"floor_{key}"] = np.array(_rel(alt[key], ref[key]))
    return out


CFG2_STEPS = 21  # the initial state and 20 coupled time steps
CFG2_STEPS_KEPT = (5, 10, 20)
CFG2_NEWTON, CFG2_FSI = 5, 3  # fixed iteration counts (every tolerance 0)


def _cfg2_model():
    """The coupled gust workload in the oracle: 81-node cantilever along y
    (80 two-node elements, two per stiffness station, 40 seeded stations),
    16 x 80 lattice mirrored about y = 0 with 300 prescribed wake rows,
    U = 10 at 5 deg, 1-cos gust of 0.05 U over 2 x 5 chords starting 11
    chords upstream, dt = c / (16 U) -- the model bench.py times -- with
    CFG2_NEWTON Newton and CFG2_FSI FSI iterations per solve."""
    sys.path.insert(0, str(HERE.parent.parent))
    import bench
    from oracle import beam_oracle as B

    m, n, ms, stations = bench.CPL_M, bench.CPL_N, bench.CPL_M_STAR, bench.CPL_STATIONS
    nn = n + 1
    beam = B.BeamOracle(nn, B.chain(nn), np.array((0.0, 0.0, 1.0)), k_cs_index=np.arange(n) * stations // n,
                        settings=B.ConvergenceSettings(CFG2_NEWTON, 0.0, 0.0, 0.0, 0.0))
    coords = np.zeros((nn, 3))
    coords[:, 1] = np.linspace(0.0, 10.0, nn)
    k, mm = bench.coupled_design(stations=stations)
    beam.set_design_variables(coords, k, mm)
    uv = O.UVLMOracle([(m, n, ms)], np.arange(nn), mirror_point=np.zeros(3),
                      mirror_normal=np.array((0.0, 1.0, 0.0)))
    a = np.deg2rad(5.0)
    u_inf = bench.CPL_U * np.array((np.cos(a), 0.0, np.sin(a)))
    gust = O.OneMinusCosine(u_inf, 1.225, True, gust_length=5.0, gust_amplitude=0.05 * bench.CPL_U,
                            gust_x0=np.array((-11.0, 0.0, 0.0)))
    uv.set_design_variables(1.0 / (m * bench.CPL_U), gust, O.make_rectangular_grid(m, n, 1.0, 0.25), beam.hg0)
    return B.CoupledOracle(beam, uv, fsi_settings=B.ConvergenceSettings(CFG2_FSI, 0.0, 0.0, 0.0, 0.0))


def _cfg2_run(alt):
    """One oracle run; ``alt`` moves the beam's sin and cos by one ulp and
    sums every pair sum over the filaments in reverse order."""
    from oracle import beam_oracle as B
    from oracle import c_oracle

    def ps(t, y0, y1, g, kernel="epsilon", chunk=256):
        assert kernel == "epsilon"
        if alt:
            return c_oracle.pair_sum_eps(t, y0[::-1], y1[::-1], g[::-1])
        return c_oracle.pair_sum_eps(t, y0, y1, g)

    O.set_pair_sum(ps)
    try:
        if alt:
            with B.ulp_shifted_trig():
                cw = _cfg2_model()
                st = cw.static_solve(np.arange(6))
                init = cw.uvlm.solve_static(st["hg"])
                d = cw.dynamic_solve(st["hg"], init, np.arange(6), CFG2_STEPS)
        else:
            cw = _cfg2_model()
            st = cw.static_solve(np.arange(6))
            init = cw.uvlm.solve_static(st["hg"])
            d = cw.dynamic_solve(st["hg"], init, np.arange(6), CFG2_STEPS)
    finally:
        O.set_pair_sum(None)
    keep = list(CFG2_STEPS_KEPT)
    return {"static_hg": st["hg"], "static_gamma_b": st["aero"]["gamma_b"][0],
            "hg": d["hg"][keep], "gamma_b": d["aero"].gamma_b[0][keep], "f_ext_aero": d["f_ext_aero"][keep],
            "zeta_w": d["aero"].zeta_w[0][CFG2_STEPS - 1][:CFG2_STEPS + 1]}


def cfg2_coupled():
    """The coupled gust workload at its full size for 20 steps (static
    Aitken FSI, then generalised-alpha FSI steps), with fixed iteration
    counts so that both sides take the same iterations.  Run twice, the
    second time with the beam's sin and cos moved by one ulp and the pair
    sums summed in reverse order: the difference ("floor_*") is how closely
    two correct FP64 evaluations of the reference's formulas agree on this
    case.  Kept: the static equilibrium (frames, circulation), and at
    CFG2_STEPS_KEPT the node frames, circulation and the beam's aero loads,
    and the wake rows that have been shed by step 20.  About 15 minutes on 8
    cores."""
    ref = _cfg2_run(False)
    alt = _cfg2_run(True)
    out = dict(ref)
    out["steps_kept"] = np.array(CFG2_STEPS_KEPT)
    out["iterations"] = np.array((CFG2_NEWTON, CFG2_FSI))
    for key in ("hg", "gamma_b", "f_ext_aero"):
        out[f"floor_{key}"] = np.array([_rel(alt[key][k], ref[key][k]) for k in range(len(CFG2_STEPS_KEPT))])
    for key in ("static_hg", "static_gamma_b", "zeta_w"):
        out[f"floor_{key}"] = np.array(_rel(alt[key], ref[key]))
    return out

if __name__ == "__main__":
    which = sys.argv[1:] or ["small", "cfg1"]
    if "small" in which:
        np.savez_compressed(HERE / "cfg0_static.npz", **cfg0_static())
        np.savez_compressed(HERE / "free_wake_8x16.npz", **small_free_wake())
    if "cfg1" in which:
        np.savez_compressed(HERE / "cfg1_free_wake_16x64.npz", **cfg1_free_wake())
    if "cfg2" in which:
        np.savez_compressed(HERE / "cfg2_coupled_16x80.npz", **cfg2_coupled())
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)
