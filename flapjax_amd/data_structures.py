"""Grid discretisation and solution container (aero/data_structures.py:29-615).

``AeroCase`` keeps the reference's field names; every per-surface history is
a device tensor with a leading time axis, so a whole run stays resident in
HBM (cfg 16x64 with 200 wake rows and 200 steps: ~70 MB).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch


@dataclass
class GridDiscretisation:
    """aero/data_structures.py:29-40: chordwise m, spanwise n, wake rows m_star."""

    m: int
    n: int
    m_star: int


FIELDS = ("zeta_b", "zeta_b_dot", "zeta_w", "gamma_b", "gamma_b_dot", "gamma_w", "f_steady", "f_unsteady",
          "alpha", "c", "nc")


def field_shapes(gd: GridDiscretisation) -> dict:
    m, n, ms = gd.m, gd.n, gd.m_star
    return {
        "zeta_b": (m + 1, n + 1, 3),
        "zeta_b_dot": (m + 1, n + 1, 3),
        "zeta_w": (ms + 1, n + 1, 3),
        "gamma_b": (m, n),
        "gamma_b_dot": (m, n),
        "gamma_w": (ms, n),
        "f_steady": (m + 1, n + 1, 3),
        "f_unsteady": (m + 1, n + 1, 3),
        "alpha": (n,),
        "c": (m, n, 3),
        "nc": (m, n, 3),
    }


@dataclass
class AeroCase:
    """Solution history (aero/data_structures.py:43-615): per field a list over
    surfaces of (n_tstep, ...) device tensors; ``t`` is a host array."""

    grid_disc: tuple
    n_tstep: int
    t: np.ndarray
    zeta_b: list = field(default_factory=list)
    zeta_b_dot: list = field(default_factory=list)
    zeta_w: list = field(default_factory=list)
    gamma_b: list = field(default_factory=list)
    gamma_b_dot: list = field(default_factory=list)
    gamma_w: list = field(default_factory=list)
    f_steady: list = field(default_factory=list)
    f_unsteady: list = field(default_factory=list)
    alpha: list = field(default_factory=list)
    c: list = field(default_factory=list)
    nc: list = field(default_factory=list)
    surf_b_names: list = field(default_factory=list)
    surf_w_names: list = field(default_factory=list)

    def __post_init__(self):
        # aero/uvlm.py:330-331 naming
        if not self.surf_b_names:
            self.surf_b_names = [f"surf_{i}_bound" for i in range(len(self.grid_disc))]
        if not self.surf_w_names:
            self.surf_w_names = [f"surf_{i}_wake" for i in range(len(self.grid_disc))]

    @classmethod
    def zeros(cls, grid_disc, n_tstep: int, device) -> "AeroCase":
        case = cls(grid_disc=tuple(grid_disc), n_tstep=n_tstep, t=np.zeros(n_tstep))
        for gd in grid_disc:
            for name, shp in field_shapes(gd).items():
                getattr(case, name).append(torch.zeros((n_tstep, *shp), dtype=torch.float64, device=device))
        return case

    def snapshot(self, i_ts: int) -> dict:
        return {name: [a[i_ts] for a in getattr(self, name)] for name in FIELDS}

    def to_dynamic(self, i_ts: int, n_tstep: int) -> "AeroCase":
        """aero/data_structures.py:570-610: expand snapshot i_ts into a history."""
        dev = self.gamma_b[0].device
        out = AeroCase.zeros(self.grid_disc, n_tstep, dev)
        for name in FIELDS:
            for dst, src in zip(getattr(out, name), getattr(self, name)):
                dst[i_ts].copy_(src[i_ts if self.n_tstep > 1 else 0])
        out.t[i_ts] = self.t[i_ts if self.n_tstep > 1 else 0]
        return out

    def plot(self, directory, plot_bound: bool = True, plot_wake: bool = True, index=None):
        """VTK output (aero/data_structures.py:401-461): per-surface ``.vts``
        files for a single-step case, per-step files and one ``.pvd`` per
        surface and kind for a history."""
        from .vtk_output import plot_case

        return plot_case(self, directory, plot_bound, plot_wake, index)

    def to_numpy(self) -> dict:
        res = {"t": self.t.copy()}
        for name in FIELDS:
            res[name] = [a.detach().cpu().numpy() for a in getattr(self, name)]
        return res
