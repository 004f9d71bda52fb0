"""Background flow fields (aero/flowfields.py:13-200).

Same constructors and attributes as the reference.  Evaluation happens in the
HIP kernels that need U(x, t) (wake convection, right-hand side, forces); a
field crosses the C-ABI as an ``fj_flowfield`` descriptor.  Only the two
reference kinds are supported on the device path; another subclass raises
``NotImplementedError`` when a kernel needs it.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from . import _kernels as K


def _vec3(v, name):
    a = np.asarray(v.detach().cpu() if isinstance(v, torch.Tensor) else v, dtype=np.float64).reshape(-1)
    if a.shape != (3,):
        raise ValueError(f"{name} must have shape (3,), got {a.shape}")
    return a


class FlowField:
    """aero/flowfields.py:13-84."""

    def __init__(self, u_inf, rho, relative_motion: bool):
        self.u_inf = _vec3(u_inf, "u_inf")
        self.rho = float(rho)
        self.u_inf_mag = float(np.linalg.norm(self.u_inf))
        self.u_inf_dir = self.u_inf / self.u_inf_mag
        self.q_inf = 0.5 * self.rho * self.u_inf_mag**2
        self.relative_motion = bool(relative_motion)

    def descriptor(self) -> L.FjFlowfield:
        raise NotImplementedError(f"{type(self).__name__} has no device implementation")

    def vmap_call(self, x: torch.Tensor, t) -> torch.Tensor:
        """U at every point of x (..., 3) at time t (aero/flowfields.py:49-60)."""
        return K.flow_eval(self, x, float(t))

    def surf_vmap_call(self, xs, t):
        return [self.vmap_call(x, t) for x in xs]

    def __call__(self, x, t):
        return self.vmap_call(x, t)

    def to_design_variables(self):
        return {"u_inf": self.u_inf, "rho": self.rho}


class Constant(FlowField):
    """aero/flowfields.py:88-97."""

    def descriptor(self):
        d = L.FjFlowfield()
        d.kind = L.FLOW_CONSTANT
        d.relative_motion = int(self.relative_motion)
        for k in range(3):
            d.u_inf[k] = self.u_inf[k]
        d.gust_length = 1.0
        return d


class OneMinusCosine(FlowField):
    """aero/flowfields.py:101-200."""

    def __init__(self, u_inf, rho, relative_motion, gust_length, gust_amplitude, gust_travel_direction=None,
                 gust_amplitude_direction=None, gust_x0=None):
        super().__init__(u_inf, rho, relative_motion)
        self.gust_amplitude = float(gust_amplitude)
        self.gust_length = float(gust_length)
        td = self.u_inf_dir if gust_travel_direction is None else _vec3(gust_travel_direction, "gust_travel_direction")
        self.gust_travel_direction = td / np.linalg.norm(td)
        ad = np.array((0.0, 0.0, 1.0)) if gust_amplitude_direction is None else _vec3(
            gust_amplitude_direction, "gust_amplitude")
        self.gust_amplitude_direction = ad / np.linalg.norm(ad)
        self.gust_x0 = np.zeros(3) if gust_x0 is None else _vec3(gust_x0, "gust_x0")

    def descriptor(self):
        d = L.FjFlowfield()
        d.kind = L.FLOW_ONE_MINUS_COSINE
        d.relative_motion = int(self.relative_motion)
        for k in range(3):
            d.u_inf[k] = self.u_inf[k]
            d.gust_travel_dir[k] = self.gust_travel_direction[k]
            d.gust_amplitude_dir[k] = self.gust_amplitude_direction[k]
            d.gust_x0[k] = self.gust_x0[k]
        d.gust_length = self.gust_length
        d.gust_amplitude = self.gust_amplitude
        return d

    def to_design_variables(self):
        return {"u_inf": self.u_inf, "rho": self.rho, "gust_amplitude": self.gust_amplitude,
                "gust_length": self.gust_length}
