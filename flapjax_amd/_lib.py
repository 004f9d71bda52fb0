"""ctypes binding of libflapjax_amd.so (the C-ABI in include/flapjax_amd.h).

The library is built in-tree by ``python -m flapjax_amd.build``.  There is no
fallback: if the library is missing or no GPU is visible, every compute call
raises.  Tensors cross the boundary as raw device pointers; the stream is
torch's current HIP stream.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_int32, c_size_t, c_void_p
from pathlib import Path

import numpy as np
import torch

LIB_PATH = Path(__file__).resolve().parent / "libflapjax_amd.so"

FJ_OK = 0
FJ_ERR_ARG = -1
FJ_ERR_LAUNCH = -2
FJ_ERR_ALLOC = -3
FJ_ERR_SINGULAR = -4
FJ_ERR_NOCONV = -5
FJ_ERR_SCRATCH = -6
NOALLOC_SCRATCH_BYTES = 32 << 20

KERNEL_EPSILON, KERNEL_CUTOFF, KERNEL_PLAIN = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
OUT_STORE, OUT_ADD, OUT_CONVECT = 0, 1, 2
FLOW_CONSTANT, FLOW_ONE_MINUS_COSINE = 0, 1
SEG_STRIDE = 16

_ERRS = {
    FJ_ERR_ARG: "invalid argument",
    FJ_ERR_LAUNCH: "kernel launch failed",
    FJ_ERR_ALLOC: "workspace allocation failed",
    FJ_ERR_SINGULAR: "singular matrix",
    FJ_ERR_NOCONV: "iterative refinement did not converge",
    FJ_ERR_SCRATCH: f"the call needs more than the fixed {NOALLOC_SCRATCH_BYTES >> 20} MiB device scratch",
}


class FlapjaxLibError(RuntimeError):
    pass


class FjFlowfield(ctypes.Structure):
    _fields_ = [
        ("kind", c_int),
        ("relative_motion", c_int),
        ("u_inf", c_double * 3),
        ("gust_length", c_double),
        ("gust_amplitude", c_double),
        ("gust_travel_dir", c_double * 3),
        ("gust_amplitude_dir", c_double * 3),
        ("gust_x0", c_double * 3),
    ]


_SIGS = {
    "fj_version": ([], ctypes.c_char_p),
    "fj_pair_sum_workspace_bytes": ([c_int, c_int, c_int], c_size_t),
    "fj_segments_from_grid": ([c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_double, c_void_p, c_void_p], c_int),
    "fj_pair_sum": ([c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_double, c_double, c_int, c_int,
                     POINTER(FjFlowfield), c_double, c_double, c_void_p, c_void_p], c_int),
    "fj_pair_sum_lattices_workspace_bytes": ([c_int, c_int, c_void_p], c_size_t),
    "fj_pair_sum_lattices": ([c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_int, c_double, c_double,
                              c_int, POINTER(FjFlowfield), c_double, c_double, c_void_p, c_void_p], c_int),
    "fj_pair_sum_split": ([c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_int,
                           c_double, c_double, c_int, POINTER(FjFlowfield), c_double, c_double, c_void_p, c_void_p],
                          c_int),
    "fj_pair_sum_lattices_split": ([c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_void_p,
                                    c_int, c_int, c_double, c_double, POINTER(FjFlowfield), c_double, c_double,
                                    c_void_p, c_void_p], c_int),
    "fj_pair_sum_lattices_noalloc": ([c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_int, c_double,
                                      c_double, c_int, POINTER(FjFlowfield), c_double, c_double, c_void_p], c_int),
    "fj_aic_assemble": ([c_void_p, c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p,
                         c_int, c_double, c_double, c_void_p, c_int, c_int, c_int, c_void_p], c_int),
    "fj_lu_aux_bytes": ([c_int], c_size_t),
    "fj_lu_factor": ([c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "fj_lu_solve": ([c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p], c_int),
    "fj_lu_factor_f32": ([c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "fj_band_solve_capacity": ([], c_size_t),
    "fj_band_solve": ([c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "fj_refine_workspace_bytes": ([c_int], c_size_t),
    "fj_solve_refined": ([c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int,
                          c_double, c_void_p, c_void_p, c_void_p], c_int),
    "fj_solve_refined_f64": ([c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                              c_int, c_double, c_void_p, c_void_p, c_void_p], c_int),
    "fj_lu_inverse": ([c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p], c_int),
    "fj_lu_inverse_f32": ([c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p], c_int),
    "fj_solve_refined_inv": ([c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_double,
                              c_void_p, c_void_p, c_void_p], c_int),
    "fj_solve_refined_inv_f32": ([c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_double,
                                  c_void_p, c_void_p, c_void_p], c_int),
    "fj_segments_from_grids": ([c_int, c_void_p, c_void_p, c_void_p, c_int, c_double, c_void_p, c_void_p], c_int),
    "fj_step_geometry": ([c_void_p] * 5 + [c_int, c_int] + [c_void_p] * 8, c_int),
    "fj_step_forces": ([c_void_p] * 8 + [c_int, c_int, POINTER(FjFlowfield), c_double, c_double, c_double, c_double,
                        c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "fj_grid_from_hg": ([c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p], c_int),
    "fj_colloc_normals": ([c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "fj_filament_midpoints": ([c_void_p, c_int, c_int, c_void_p, c_void_p], c_int),
    "fj_rhs": ([c_void_p, c_void_p, c_void_p, c_void_p, c_int, POINTER(FjFlowfield), c_double, c_void_p, c_void_p], c_int),
    "fj_wake_shift": ([c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p], c_int),
    "fj_wake_convect": ([c_void_p, c_void_p, c_int, c_int, POINTER(FjFlowfield), c_double, c_double, c_void_p,
                         c_void_p], c_int),
    "fj_wake_initialise": ([c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p], c_int),
    "fj_wake_rediscretise_workspace_bytes": ([c_int, c_int], c_size_t),
    "fj_wake_rediscretise": ([c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
                             c_int),
    "fj_static_wake_gamma": ([c_void_p, c_int, c_int, c_int, c_void_p, c_void_p], c_int),
    "fj_segments_from_endpoints": ([c_void_p, c_void_p, c_int, c_int, c_double, c_void_p, c_void_p], c_int),
    "fj_steady_forces": ([c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, POINTER(FjFlowfield),
                          c_double, c_double, c_void_p, c_void_p], c_int),
    "fj_gamma_dot_unsteady": ([c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_double, c_double, c_double,
                               c_void_p, c_void_p, c_void_p], c_int),
    "fj_strip_alpha": ([c_void_p, c_void_p, c_int, c_int, POINTER(FjFlowfield), c_double, c_double, c_void_p,
                        c_void_p], c_int),
    "fj_pair_sum_backward": ([c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_double, c_double, c_void_p,
                              c_int, c_void_p, c_void_p, c_void_p], c_int),
    "fj_pair_sum_backward_workspace_bytes": ([c_int, c_int], c_size_t),
    "fj_batched_task_bytes": ([c_int], c_size_t),
    "fj_batched_workspace_bytes": ([c_int, c_void_p, c_void_p], c_size_t),
    "fj_pair_sum_batched": ([c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_double, c_double, c_void_p, c_void_p,
                             c_void_p, c_void_p], c_int),
    "fj_step_batched_workspace_bytes": ([c_int, c_void_p], c_size_t),
    "fj_step_batched_table_bytes": ([c_int], c_size_t),
    "fj_step_batched": ([c_int, c_void_p, POINTER(FjFlowfield), c_double, c_double, c_int, c_double, c_double,
                         c_int, c_int, c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "fj_step_batched_restage": ([c_int, c_void_p, POINTER(FjFlowfield), c_double, c_double, c_int, c_double,
                                 c_double, c_int, c_int, c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p],
                                c_int),
    "fj_flow_eval": ([POINTER(FjFlowfield), c_void_p, c_int, c_double, c_void_p, c_void_p], c_int),
    "fj_beam_workspace_bytes": ([c_int], c_size_t),
    "fj_beam_system": ([c_void_p, c_int] + [c_void_p] * 6 + [c_double, c_double, c_void_p, c_int] + [c_void_p] * 5,
                       c_int),
    "fj_beam_element_jacobians": ([c_void_p, c_int] + [c_void_p] * 10 + [c_void_p], c_int),
    "fj_beam_update": ([c_int, c_int, c_void_p, c_int, c_void_p, c_double, c_double, c_double] + [c_void_p] * 7,
                       c_int),
    "fj_beam_subnorm": ([c_void_p, c_void_p, c_int, c_void_p, c_void_p], c_int),
    "fj_beam_ga": ([c_int, c_int] + [c_double] * 5 + [c_void_p] * 11, c_int),
    "fj_beam_fsi_norms": ([c_int] + [c_void_p] * 7, c_int),
    "fj_beam_ga_nodes": ([c_int, c_int, c_double, c_double] + [c_void_p] * 15, c_int),
    "fj_beam_newton_dynamic": ([c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_double, c_double, c_double, c_void_p, c_int, c_int, c_int, c_void_p]
                               + [c_void_p] * 11, c_int),
    "fj_beam_nodes": ([c_int, c_int] + [c_void_p] * 7, c_int),
    "fj_beam_step_end": ([c_int, c_double, c_double, c_double, c_double, c_double] + [c_void_p] * 7
                         + [c_int, c_int] + [c_void_p] * 14 + [c_void_p], c_int),
    "fj_beam_project_forces": ([c_int, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                c_void_p, c_void_p], c_int),
}

BEAM_NODE_UPDATE, BEAM_NODE_REL, BEAM_NODE_RATE, BEAM_NODE_LOCAL, BEAM_NODE_GLOBAL = 0, 1, 2, 3, 4
GA_PREDICT, GA_ALPHA, GA_FINISH, GA_FORCE = 0, 1, 2, 3


class FjConvergence(ctypes.Structure):
    """fj_convergence: a Newton loop's tolerances (NaN = unset)."""

    _fields_ = [("max_n_iter", c_int), ("rel_disp_tol", c_double), ("abs_disp_tol", c_double),
                ("rel_force_tol", c_double), ("abs_force_tol", c_double)]


class FjBeamModel(ctypes.Structure):
    """fj_beam_model: device element tables of a beam."""

    _fields_ = [("n_nodes", c_int), ("n_elem", c_int)] + [
        (name, c_void_p) for name in ("conn", "o0", "l0", "d0", "k_cs", "m_cs", "node_off", "node_elem")]



class FjStepCase(ctypes.Structure):
    """fj_step_case: one active case of a batched serving tick."""

    _fields_ = [("m", c_int), ("n", c_int), ("m_star", c_int), ("dt", c_double), ("time", c_double)] + [
        (name, c_void_p) for name in (
            "hg", "hg_dot", "dof", "zeta0", "zeta_b_nm1", "zeta_b", "zeta_b_dot", "gamma_b_nm1", "gamma_w_nm1",
            "zeta_w_nm1", "gamma_w", "zeta_w", "aic", "aic_inv", "gamma_b", "gamma_b_dot", "f_steady", "lift",
            "refine_ws", "iters")]


def _step_case_dtype():
    """NumPy structured dtype with the layout of fj_step_case (pointers as
    uint64), so a tick's case records are filled with array operations."""
    names, formats, offsets = [], [], []
    for name, ct in FjStepCase._fields_:
        names.append(name)
        formats.append(np.int32 if ct is c_int else np.float64 if ct is c_double else np.uint64)
        offsets.append(getattr(FjStepCase, name).offset)
    return np.dtype({"names": names, "formats": formats, "offsets": offsets,
                     "itemsize": ctypes.sizeof(FjStepCase)})


STEP_CASE_DTYPE = _step_case_dtype()


class FjLatticeRef(ctypes.Structure):
    """fj_lattice_ref: one lattice (its filaments' first row and size) in a
    filament table."""

    _fields_ = [("seg_off", c_int), ("m", c_int), ("n", c_int)]


class FjGridRef(ctypes.Structure):
    """fj_grid_ref: one source grid of fj_segments_from_grids."""

    _fields_ = [("zeta", c_void_p), ("gamma", c_void_p), ("m", c_int), ("n", c_int), ("mirror", c_int)]


_lib = None


def exported_symbols():
    """Names the library must export (every entry point of the header)."""
    return list(_SIGS)


def load(require_gpu: bool = False):
    """Load the library (once).  Raises if it is not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FlapjaxLibError(
                f"{LIB_PATH} is not built; run `python -m flapjax_amd.build` (there is no CPU fallback)"
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    if require_gpu and not torch.cuda.is_available():
        raise FlapjaxLibError("flapjax_amd needs a ROCm GPU; none is visible (there is no CPU fallback)")
    return _lib


_gpu_checked = False


def lib():
    """The loaded library; raises if it is not built or no GPU is visible
    (checked once per process)."""
    global _gpu_checked
    if _lib is not None and _gpu_checked:
        return _lib
    out = load(require_gpu=True)
    _gpu_checked = True
    return out


def check(rc: int, what: str) -> None:
    if rc != FJ_OK:
        raise FlapjaxLibError(f"{what}: {_ERRS.get(rc, 'error')} (status {rc})")


def ptr(t):
    """Device pointer of a tensor (or None)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("flapjax_amd expects device tensors")
    if not t.is_contiguous():
        raise ValueError("flapjax_amd expects contiguous tensors")
    return t.data_ptr()  # ctypes passes an int as the c_void_p argument


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_cur_device = getattr(torch._C, "_cuda_getDevice", None)


def stream():
    """torch's current HIP stream on the current device, as a raw handle
    (through torch's raw-stream accessor where present: the public
    current_stream() re-validates the device on every call)."""
    if _raw_stream is not None and _cur_device is not None:
        return ctypes.c_void_p(_raw_stream(_cur_device()))
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def host3(v):
    """Host pointer to 3 doubles (or None)."""
    if v is None:
        return None
    arr = (c_double * 3)(*[float(x) for x in v])
    return ctypes.cast(arr, c_void_p), arr
