"""The oracle reproduces the committed golden fixtures (tests/golden/)."""

from pathlib import Path

import numpy as np

from tests.golden import make_golden as G

HERE = Path(__file__).resolve().parent / "golden"


def _check(name, fresh):
    ref = np.load(HERE / name)
    for k in ref.files:
        assert np.allclose(fresh[k], ref[k], rtol=1e-13, atol=1e-15), k


def test_cfg0_static_fixture():
    _check("cfg0_static.npz", G.cfg0_static())


def test_free_wake_fixture():
    _check("free_wake_8x16.npz", G.small_free_wake())
