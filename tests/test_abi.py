"""The C-ABI library builds/loads on the CPU and exports every entry point of
include/flapjax_amd.h (no compute calls: no GPU here)."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def header_functions():
    text = (ROOT / "include" / "flapjax_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fj_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = header_functions()
    for required in ("fj_pair_sum", "fj_aic_assemble", "fj_lu_factor", "fj_lu_solve", "fj_wake_convect",
                     "fj_steady_forces", "fj_segments_from_grid"):
        assert required in names


def test_library_exports_every_header_symbol():
    from flapjax_amd import _lib

    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built (run python -m flapjax_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    # the Python binding declares a signature for every header entry point
    assert set(header_functions()) <= set(_lib.exported_symbols())
    assert _lib.load().fj_version().startswith(b"flapjax_amd")


def test_no_gpu_raises_loudly():
    import torch

    from flapjax_amd import _lib

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    with pytest.raises(_lib.FlapjaxLibError):
        _lib.lib()
