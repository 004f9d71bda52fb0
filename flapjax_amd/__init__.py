"""flapjax_amd: MI355X-native (gfx950) UVLM aerodynamic hot path.

Drop-in counterparts of the reference's aero hot path (ben-l-p/flapjax v1.2.1,
src/flapjax/aero/): induced velocity (``aic.compute_v_ind``), AIC assembly
(``aic.compute_aic_solve``), the circulation solve (``linalg``), wake
convection (``utils.propagate_wake``) and the UVLM time step (``uvlm.UVLM``).
All compute runs in hand-written HIP kernels behind the C-ABI declared in
``include/flapjax_amd.h`` (``libflapjax_amd.so``); there is no CPU fallback.
"""

from . import _lib
from .data_structures import AeroCase, GridDiscretisation
from .flowfields import Constant, FlowField, OneMinusCosine

__all__ = ["AeroCase", "GridDiscretisation", "Constant", "FlowField", "OneMinusCosine", "library_path",
           "UVLM"]

library_path = _lib.LIB_PATH


def __getattr__(name):
    if name == "UVLM":
        from .uvlm import UVLM

        return UVLM
    raise AttributeError(name)
