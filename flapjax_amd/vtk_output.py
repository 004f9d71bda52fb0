"""VTK XML output of aerodynamic grids (plotting/aerogrid.py:42-134,
plotting/pvd.py:9-64).

The reference builds a vtkStructuredGrid with the vtk package and writes it
with vtkXMLStructuredGridWriter; here the same file kind is written directly
(VTK XML StructuredGrid, ``.vts``, ASCII data arrays with 17 significant
digits), so no vtk installation is needed.  Layout follows the reference:

* a planar grid (m, n, 3) becomes a structured grid of dimensions (m, n, 1);
  a volume grid (m, n, p, 3) of dimensions (m, n, p);
* points and data are ordered with the FIRST grid axis fastest (the
  reference swaps axis 0 with the last grid axis before ravelling);
* the file is ``<filename>_ts_<i_ts>.vts`` when a time step is given;
* a ``.pvd`` collection lists (timestep, file) pairs with paths relative to
  its directory and times printed with four decimals.
"""

from __future__ import annotations

import os
import xml.etree.ElementTree as Et
from collections.abc import Sequence
from pathlib import Path

import numpy as np


def _np(a):
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.float64)


def _vtk_order(arr, i_swap):
    """Swap grid axis 0 with grid axis i_swap and ravel (first axis fastest)."""
    return np.swapaxes(arr, 0, i_swap).reshape(-1)


def _data_array(parent, name, values, ncomp):
    el = Et.SubElement(parent, "DataArray", {"type": "Float64", "Name": name, "NumberOfComponents": str(ncomp),
                                             "format": "ascii"})
    el.text = " ".join(f"{v:.17g}" for v in values)
    return el


def _vectors(arr, i_swap):
    comps = [_vtk_order(arr[..., k], i_swap) for k in range(3)]
    return np.stack(comps, axis=-1).reshape(-1)


def plot_grid_to_vtk(grid_arr, filename: str | os.PathLike, i_ts: int | None = None,
                     node_scalar_data: dict | None = None, node_vector_data: dict | None = None,
                     cell_scalar_data: dict | None = None, cell_vector_data: dict | None = None) -> Path:
    """Write one time step of a structured grid (plotting/aerogrid.py:42-134)."""
    grid = _np(grid_arr)
    if grid.ndim == 3:
        is_planar = True
    elif grid.ndim == 4:
        is_planar = False
    else:
        raise ValueError(f"grid_arr must have 3 or 4 dimensions, got {grid.ndim}-D array")
    i_swap = 2 - int(is_planar)
    dims = list(grid.shape[:-1]) + ([1] if is_planar else [])
    extent = " ".join(f"0 {d - 1}" for d in dims)

    root = Et.Element("VTKFile", {"type": "StructuredGrid", "version": "0.1", "byte_order": "LittleEndian"})
    sg = Et.SubElement(root, "StructuredGrid", {"WholeExtent": extent})
    piece = Et.SubElement(sg, "Piece", {"Extent": extent})

    pd = Et.SubElement(piece, "PointData")
    for name, arr in (node_scalar_data or {}).items():
        _data_array(pd, name, _vtk_order(_np(arr), i_swap), 1)
    for name, arr in (node_vector_data or {}).items():
        a = _np(arr)
        if a.shape[-1] != 3:
            raise ValueError(f"Node vector data '{name}' must have trailing dimension of size 3, got {a.shape[-1]}")
        _data_array(pd, name, _vectors(a, i_swap), 3)
    cd = Et.SubElement(piece, "CellData")
    for name, arr in (cell_scalar_data or {}).items():
        _data_array(cd, name, _vtk_order(_np(arr), i_swap), 1)
    for name, arr in (cell_vector_data or {}).items():
        a = _np(arr)
        if a.shape[-1] != 3:
            raise ValueError(f"Node vector data '{name}' must have trailing dimension of size 3, got {a.shape[-1]}")
        _data_array(cd, name, _vectors(a, i_swap), 3)
    pts = Et.SubElement(piece, "Points")
    _data_array(pts, "Points", _vectors(grid, i_swap), 3)

    path = Path(filename)
    name = path.name + (f"_ts_{i_ts}" if i_ts is not None else "")
    full = path.parent.joinpath(name).with_suffix(".vts")
    full.parent.mkdir(parents=True, exist_ok=True)
    Et.ElementTree(root).write(full, encoding="utf-8", xml_declaration=True)
    return full


def write_pvd(directory: os.PathLike | str, name: str, file_dirs: Sequence[Path], times: Sequence[float]) -> Path:
    """PVD collection of per-step files (plotting/pvd.py:9-64)."""
    dirpath = Path(directory)
    dirpath.mkdir(parents=True, exist_ok=True)
    if len(file_dirs) == 0:
        raise ValueError("filenames must be a non-empty sequence")
    if len(file_dirs) != len(times):
        raise ValueError("filenames and times must have the same length")
    root = Et.Element("VTKFile", {"type": "Collection", "version": "0.1", "byte_order": "LittleEndian"})
    coll = Et.SubElement(root, "Collection")
    if not name.endswith(".pvd"):
        name += ".pvd"
    pvd = dirpath.joinpath(name)
    for fd, t in zip(file_dirs, times):
        Et.SubElement(coll, "DataSet", {"timestep": f"{float(t):.04f}", "group": "", "part": "0",
                                        "file": Path(fd).relative_to(dirpath).as_posix()})
    Et.indent(root, space="  ")
    Et.ElementTree(root).write(pvd, encoding="utf-8", xml_declaration=True)
    return pvd


def plot_surface_snapshot(case, i_surf: int, i_ts: int, directory, plot_bound=True, plot_wake=True,
                          warn_empty=True) -> list[Path]:
    """One surface at one step (aero/data_structures.py:649-697)."""
    paths = []
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    k = i_ts if case.n_tstep > 1 else 0
    if plot_bound:
        gb = _np(case.gamma_b[i_surf][k])
        paths.append(plot_grid_to_vtk(
            case.zeta_b[i_surf][k], d.joinpath(case.surf_b_names[i_surf]), i_ts,
            node_vector_data={"f_steady": case.f_steady[i_surf][k], "f_unsteady": case.f_unsteady[i_surf][k],
                              "zeta_dot": case.zeta_b_dot[i_surf][k]},
            cell_scalar_data={"gamma": gb, "gamma_dot": case.gamma_b_dot[i_surf][k],
                              "alpha": np.broadcast_to(_np(case.alpha[i_surf][k]), gb.shape)}))
    if plot_wake:
        gw = _np(case.gamma_w[i_surf][k])
        if not gw.shape[0]:
            if warn_empty:
                print("Warning: No wake panels to plot, skipping.")
        else:
            paths.append(plot_grid_to_vtk(case.zeta_w[i_surf][k], d.joinpath(case.surf_w_names[i_surf]), i_ts,
                                          cell_scalar_data={"gamma": gw}))
    return paths


def plot_case(case, directory, plot_bound=True, plot_wake=True, index=None) -> list[Path]:
    """AeroCase.plot (aero/data_structures.py:401-461): per-surface files for
    a single-step case; per-step files plus one PVD per surface and kind for a
    history."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    if case.n_tstep == 1:
        out = []
        for s in range(len(case.grid_disc)):
            out += plot_surface_snapshot(case, s, 0, d, plot_bound, plot_wake)
        return out
    if index is None:
        idx = np.arange(case.n_tstep)
    elif isinstance(index, slice):
        idx = np.arange(case.n_tstep)[index]
    else:
        idx = np.atleast_1d(np.asarray(index, dtype=int))
    pvds = []
    for s in range(len(case.grid_disc)):
        per_ts = [plot_surface_snapshot(case, s, int(i), d, plot_bound, plot_wake) for i in idx]
        times = [float(case.t[int(i)]) for i in idx]
        if plot_bound:
            pvds.append(write_pvd(d, f"aero_dynamic_{case.surf_b_names[s]}_ts", [p[0] for p in per_ts], times))
        if plot_wake and all(len(p) == (2 if plot_bound else 1) for p in per_ts):
            pvds.append(write_pvd(d, f"aero_dynamic_{case.surf_w_names[s]}_ts", [p[-1] for p in per_ts], times))
    return pvds
