"""Writes tests/golden/beam_newton_system_480.npz: the last Newton system of
the coupled gust wing (bench.make_coupled: 16x80 mirrored lattice, 80
two-node beam elements, static aeroelastic solve and 8 coupled time steps)
in LAPACK band storage (ab, kl = ku = 11) with its right-hand side.

Needs a GPU (the system is assembled by the device beam):
    python tests/golden/make_beam_system.py
"""

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))


def main():
    import bench

    cw = bench.make_coupled("fp64")
    st = cw.static_solve(np.arange(6))
    cw.dynamic_solve(st, np.arange(6), 8)
    a_dev, r_dev = cw.structure._sys_buffers(480, cw.structure.hg0.device)[:2]
    a, b = a_dev.cpu().numpy(), r_dev.cpu().numpy()
    n, kl, ku = a.shape[0], 11, 11
    ab = np.zeros((kl + ku + 1, n))
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            ab[ku + i - j, j] = a[i, j]
    rows, cols = np.nonzero(a)
    assert np.all(np.abs(rows - cols) <= 11), "entries outside the band"
    np.savez_compressed(Path(__file__).parent / "beam_newton_system_480.npz", ab=ab, b=b)


if __name__ == "__main__":
    main()
