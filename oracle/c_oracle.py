"""ctypes loader of oracle/liboracle_bs.so (the C restatement of the FP64
pair sum).  Test and CPU-baseline infrastructure only."""

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_bs.so"


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def _lib():
    if not LIB.exists():
        build()
    lib = ctypes.CDLL(str(LIB))
    p = ctypes.POINTER(ctypes.c_double)
    lib.fj_oracle_pair_sum_eps.argtypes = [p, ctypes.c_long, p, p, p, ctypes.c_long, ctypes.c_double, p]
    lib.fj_oracle_set_threads.argtypes = [ctypes.c_int]
    return lib


def set_threads(n: int) -> None:
    """OpenMP threads of the following pair sums."""
    _lib().fj_oracle_set_threads(int(n))


def pair_sum_eps(targets, y0, y1, g, eps=1e-7):
    """FP64 epsilon-kernel pair sum, OpenMP over targets (OMP_NUM_THREADS)."""
    x = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
    a = np.ascontiguousarray(y0, dtype=np.float64).reshape(-1, 3)
    b = np.ascontiguousarray(y1, dtype=np.float64).reshape(-1, 3)
    g = np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
    out = np.empty_like(x)
    p = lambda arr: arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    _lib().fj_oracle_pair_sum_eps(p(x), x.shape[0], p(a), p(b), p(g), a.shape[0], float(eps), p(out))
    return out


def oracle_pair_sum(targets, y0, y1, g, kernel="epsilon", chunk=256):
    """Drop-in for uvlm_oracle.pair_sum (install with uvlm_oracle.set_pair_sum):
    epsilon-kernel sums go to the C restatement, the other kernels to NumPy."""
    if kernel != "epsilon":
        from oracle import uvlm_oracle

        return uvlm_oracle.pair_sum(targets, y0, y1, g, kernel, chunk)
    return pair_sum_eps(targets, y0, y1, g)
