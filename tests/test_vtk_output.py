"""VTK XML output (plotting/aerogrid.py:42-134, plotting/pvd.py:9-64): files
parse as VTK XML, carry the grid's dimensions, and order points and data with
the first grid axis fastest, as the reference's axis swap does."""

import xml.etree.ElementTree as Et

import numpy as np
import pytest
import torch


def _arrays(path):
    root = Et.parse(path).getroot()
    assert root.get("type") == "StructuredGrid"
    piece = root.find("StructuredGrid/Piece")
    out = {"extent": piece.get("Extent")}
    for da in piece.iter("DataArray"):
        out[da.get("Name")] = (int(da.get("NumberOfComponents")), np.array(da.text.split(), dtype=float))
    return out


def test_planar_grid_layout(tmp_path):
    from flapjax_amd.vtk_output import plot_grid_to_vtk

    rng = np.random.default_rng(0)
    m, n = 4, 6
    grid = rng.normal(size=(m, n, 3))
    f = rng.normal(size=(m, n, 3))
    g = rng.normal(size=(m - 1, n - 1))
    p = plot_grid_to_vtk(grid, tmp_path / "wing", 7, node_vector_data={"f": f}, cell_scalar_data={"gamma": g})
    assert p.name == "wing_ts_7.vts"
    a = _arrays(p)
    assert a["extent"] == f"0 {m - 1} 0 {n - 1} 0 0"
    nc, pts = a["Points"]
    pts = pts.reshape(-1, 3)
    # point index = i + j * m (first axis fastest)
    for i, j in ((0, 0), (3, 0), (1, 2), (3, 5)):
        assert np.array_equal(pts[i + j * m], grid[i, j])
    assert np.array_equal(a["f"][1].reshape(-1, 3)[2 + 4 * m], f[2, 4])
    assert np.array_equal(a["gamma"][1][1 + 3 * (m - 1)], g[1, 3])


def test_volume_grid_and_errors(tmp_path):
    from flapjax_amd.vtk_output import plot_grid_to_vtk, write_pvd

    grid = np.arange(2 * 3 * 4 * 3, dtype=float).reshape(2, 3, 4, 3)
    p = plot_grid_to_vtk(grid, tmp_path / "vol")
    a = _arrays(p)
    assert a["extent"] == "0 1 0 2 0 3"
    pts = a["Points"][1].reshape(-1, 3)
    assert np.array_equal(pts[1 + 2 * 2 + 3 * 6], grid[1, 2, 3])
    with pytest.raises(ValueError):
        plot_grid_to_vtk(np.zeros((3, 3)), tmp_path / "bad")
    with pytest.raises(ValueError):
        plot_grid_to_vtk(np.zeros((3, 3, 3)), tmp_path / "bad", node_vector_data={"v": np.zeros((3, 3, 2))})
    with pytest.raises(ValueError):
        write_pvd(tmp_path, "x", [], [])
    with pytest.raises(ValueError):
        write_pvd(tmp_path, "x", [p], [0.0, 1.0])


def test_case_plot_history_and_pvd(tmp_path):
    from flapjax_amd.data_structures import AeroCase, GridDiscretisation

    case = AeroCase.zeros([GridDiscretisation(2, 3, 2)], 3, torch.device("cpu"))
    for i in range(3):
        case.zeta_b[0][i] = torch.rand(3, 4, 3, dtype=torch.float64)
        case.zeta_w[0][i] = torch.rand(3, 4, 3, dtype=torch.float64)
        case.gamma_b[0][i] = float(i)
    case.t[:] = [0.0, 0.125, 0.25]
    pvds = case.plot(tmp_path)
    assert [p.name for p in pvds] == ["aero_dynamic_surf_0_bound_ts.pvd", "aero_dynamic_surf_0_wake_ts.pvd"]
    root = Et.parse(pvds[0]).getroot()
    ds = root.findall("Collection/DataSet")
    assert [d.get("timestep") for d in ds] == ["0.0000", "0.1250", "0.2500"]
    assert [d.get("file") for d in ds] == [f"surf_0_bound_ts_{i}.vts" for i in range(3)]
    a = _arrays(tmp_path / "surf_0_bound_ts_2.vts")
    assert np.all(a["gamma"][1] == 2.0)
    assert np.array_equal(a["Points"][1].reshape(-1, 3)[1 + 2 * 3], case.zeta_b[0][2][1, 2].numpy())
    # a single-step case: one file per surface and kind, step 0
    single = AeroCase.zeros([GridDiscretisation(2, 3, 0)], 1, torch.device("cpu"))
    paths = single.plot(tmp_path / "s")
    assert [p.name for p in paths] == ["surf_0_bound_ts_0.vts"]
