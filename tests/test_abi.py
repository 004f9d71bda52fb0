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


def test_workspace_queries_host_only():
    """The size queries are host functions (no device work): the lattice
    pair-sum workspace grows with the targets and holds the partials and the
    strip-column records; the batched step workspace of two cases is at
    least that of either case alone."""
    from flapjax_amd import _kernels as K
    from flapjax_amd import _lib as L

    if not L.LIB_PATH.exists():
        pytest.skip("library not built (run python -m flapjax_amd.build)")
    lib = L.load()
    lats = [(0, 16, 64), (16 * 65 + 17 * 64, 200, 64)]
    refs = K.lattice_refs(lats)
    sizes = [lib.fj_pair_sum_lattices_workspace_bytes(nt, 2, refs) for nt in (1, 64, 1000, 13000)]
    assert all(a <= b for a, b in zip(sizes, sizes[1:]))
    # records alone: (16/2 + 200/2) strips x 65 columns x 36 doubles
    assert sizes[0] >= 108 * 65 * 36 * 8
    assert lib.fj_pair_sum_lattices_workspace_bytes(100, 17, refs) == 256  # more than 16 lattices: rejected

    def cases(specs):
        arr = (L.FjStepCase * len(specs))()
        for c, (m, n, ms) in zip(arr, specs):
            c.m, c.n, c.m_star, c.dt = m, n, ms, 0.1
        return arr

    one = lib.fj_step_batched_workspace_bytes(1, cases([(8, 32, 20)]))
    two = lib.fj_step_batched_workspace_bytes(2, cases([(8, 32, 20), (4, 16, 10)]))
    assert one > 0 and two >= one
    assert lib.fj_step_batched_table_bytes(2) >= lib.fj_step_batched_table_bytes(1)
