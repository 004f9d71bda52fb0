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


if __name__ == "__main__":
    np.savez_compressed(HERE / "cfg0_static.npz", **cfg0_static())
    np.savez_compressed(HERE / "free_wake_8x16.npz", **small_free_wake())
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)
