"""MPAS netCDF grid/static reader and the dycore restart file.

Grid files: ``x1.N.grid.nc`` / ``x1.N.static.nc`` in netCDF classic or
64-bit-offset form (read with ``scipy.io.netcdf_file``).  HDF5-based netCDF-4
files are refused with a message (convert them first, e.g. ``nccopy -k
64-bit-offset``).  The loader mirrors the
reference's ``Mesh.from_netcdf`` (``mesh.py:1373-1450``): variables of the
grid file, overlaid by same-named variables of an optional static file;
1-based connectivity with 0 as "absent" becomes 0-based with -1 padding
(``mesh.py:331-366``).  A grid on the unit sphere (``sphere_radius`` 1) is
scaled to the Earth radius: lengths by R, areas by R^2 (as MPAS init does).

Restart files carry the dycore state in MPAS-style names and dimensions
``(Time, nCells|nEdges, nVertLevels|nVertLevelsP1)``: the prognostic fields,
the carried diagnostics of the last step and, as output only, the last
acoustic ru_p/rw_p words (``cuda_driver.py:4707-4711``).  Reading one back
into a driver with ``set_state(..., rebuild=False)`` continues bit for bit.
"""

from __future__ import annotations

from pathlib import Path
from typing import Dict, Mapping, Optional, Tuple

import numpy as np

from .mesh import EARTH_RADIUS, Mesh, _kite_for_cell

HDF5_MAGIC = b"\x89HDF\r\n\x1a\n"

#: connectivity variables (1-based in files, 0 = absent)
CONNECTIVITY = ("cellsOnEdge", "verticesOnEdge", "edgesOnCell", "cellsOnCell", "verticesOnCell",
                "cellsOnVertex", "edgesOnVertex", "edgesOnEdge")
#: per-entity counts that are not indices
COUNTS = ("nEdgesOnCell", "nEdgesOnEdge")
LENGTHS = ("xCell", "yCell", "zCell", "xEdge", "yEdge", "zEdge", "xVertex", "yVertex", "zVertex",
           "dcEdge", "dvEdge")
AREAS = ("areaCell", "areaTriangle", "kiteAreasOnVertex")
REQUIRED = CONNECTIVITY + COUNTS + ("xCell", "yCell", "zCell", "xVertex", "yVertex", "zVertex", "dcEdge",
                                    "dvEdge", "areaCell", "areaTriangle", "kiteAreasOnVertex", "weightsOnEdge",
                                    "angleEdge", "fEdge", "fVertex")


class MpasFileError(ValueError):
    pass


def _read(path: Path) -> Tuple[Dict[str, np.ndarray], Dict[str, object]]:
    from scipy.io import netcdf_file

    with open(path, "rb") as fh:
        if fh.read(8) == HDF5_MAGIC:
            raise MpasFileError(f"{path}: netCDF-4/HDF5 files are not readable here; "
                                "convert to classic or 64-bit-offset netCDF first")
    arrays: Dict[str, np.ndarray] = {}
    with netcdf_file(str(path), "r", mmap=False) as f:
        attrs = {k: (v.decode() if isinstance(v, bytes) else v) for k, v in f._attributes.items()}
        for name, var in f.variables.items():
            a = np.array(var.data, copy=True)
            if a.ndim >= 1 and var.dimensions and var.dimensions[0] == "Time":
                a = a[-1] if a.shape[0] else a[0:0]
            arrays[name] = a
    return arrays, attrs


def mesh_from_mpas(grid_path, static_path=None, *, radius: float = EARTH_RADIUS) -> Mesh:
    """Mesh from an MPAS grid file, optionally overlaid by a static file."""
    arrays, attrs = _read(Path(grid_path))
    if static_path is not None:
        s_arrays, s_attrs = _read(Path(static_path))
        arrays.update(s_arrays)
        attrs = {**attrs, **s_attrs}
    missing = [n for n in REQUIRED if n not in arrays]
    if missing:
        raise MpasFileError(f"missing MPAS variables: {missing}")
    on_sphere = str(attrs.get("on_a_sphere", "YES")).strip().upper() == "YES"
    if not on_sphere:
        raise MpasFileError("planar meshes are not supported")
    out: Dict[str, np.ndarray] = {}
    for name, a in arrays.items():
        if name in CONNECTIVITY:
            a = a.astype(np.int64) - 1          # 0 (absent) becomes -1
            a[a < -1] = -1
        elif name in COUNTS:
            a = a.astype(np.int64)
        elif a.dtype.kind == "f":
            a = a.astype(np.float64)
        out[name] = a
    file_radius = float(attrs.get("sphere_radius", 0.0) or 0.0)
    if file_radius == 0.0:
        r = np.sqrt(out["xCell"] ** 2 + out["yCell"] ** 2 + out["zCell"] ** 2)
        file_radius = float(np.median(r))
    if abs(file_radius - 1.0) < 1e-6:
        for n in LENGTHS:
            if n in out:
                out[n] = out[n] * radius
        for n in AREAS:
            if n in out:
                out[n] = out[n] * (radius * radius)
        sphere = radius
    else:
        sphere = file_radius
    if "kiteForCell" not in out:
        out["kiteForCell"] = _kite_for_cell(Mesh(arrays=out, sphere_radius=sphere))
    if "nominalMinDc" not in out:
        out["nominalMinDc"] = np.array([out["dcEdge"].min()])
    return Mesh(arrays=out, sphere_radius=sphere)


def write_mpas_grid(mesh: Mesh, path, *, unit_sphere: bool = False) -> None:
    """Write the mesh tables as an MPAS grid file (64-bit offset netCDF)."""
    from scipy.io import netcdf_file

    a = mesh.arrays
    R = mesh.sphere_radius
    dims = {"nCells": mesh.n_cells, "nEdges": mesh.n_edges, "nVertices": mesh.n_vertices,
            "maxEdges": mesh.max_edges, "maxEdges2": a["edgesOnEdge"].shape[1], "vertexDegree": 3}
    shape_dims = {
        "cell": ("nCells",), "edge": ("nEdges",), "vertex": ("nVertices",),
        "cellME": ("nCells", "maxEdges"), "edgeME2": ("nEdges", "maxEdges2"), "edge2": ("nEdges", "TWO"),
        "vertex3": ("nVertices", "vertexDegree"),
    }
    layout = {
        "cellsOnEdge": "edge2", "verticesOnEdge": "edge2", "edgesOnCell": "cellME", "cellsOnCell": "cellME",
        "verticesOnCell": "cellME", "cellsOnVertex": "vertex3", "edgesOnVertex": "vertex3",
        "edgesOnEdge": "edgeME2", "weightsOnEdge": "edgeME2", "nEdgesOnCell": "cell", "nEdgesOnEdge": "edge",
        "kiteAreasOnVertex": "vertex3",
    }
    with netcdf_file(str(path), "w", version=2) as f:
        f.on_a_sphere = "YES"
        f.sphere_radius = 1.0 if unit_sphere else float(R)
        for name, n in dims.items():
            f.createDimension(name, n)
        f.createDimension("TWO", 2)
        for name, arr in a.items():
            if name in ("nominalMinDc",):
                continue
            if name in layout:
                d = shape_dims[layout[name]]
            elif arr.ndim == 1 and arr.size == mesh.n_cells:
                d = ("nCells",)
            elif arr.ndim == 1 and arr.size == mesh.n_edges:
                d = ("nEdges",)
            elif arr.ndim == 1 and arr.size == mesh.n_vertices:
                d = ("nVertices",)
            elif name == "kiteForCell":
                continue
            else:
                continue
            if name in CONNECTIVITY:
                v = f.createVariable(name, "i4", d)
                v[:] = np.where(arr >= 0, arr + 1, 0).astype(np.int32)
            elif name in COUNTS:
                v = f.createVariable(name, "i4", d)
                v[:] = arr.astype(np.int32)
            else:
                x = np.asarray(arr, dtype=np.float64)
                if unit_sphere and name in LENGTHS:
                    x = x / R
                elif unit_sphere and name in AREAS:
                    x = x / (R * R)
                v = f.createVariable(name, "f8", d)
                v[:] = x


# ------------------------------------------------------------------ restart
#: ABI state field -> (MPAS-style name, entity dimension, level dimension)
RESTART_FIELDS = {
    "rho_zz": ("rho_zz", "nCells", "nVertLevels"),
    "rtheta_m": ("rtheta_m", "nCells", "nVertLevels"),
    "ru": ("ru", "nEdges", "nVertLevels"),
    "rw": ("rw", "nCells", "nVertLevelsP1"),
    "theta_m": ("theta_m", "nCells", "nVertLevels"),
    "exner": ("exner", "nCells", "nVertLevels"),
    "rho_p": ("rho_p", "nCells", "nVertLevels"),
    "rtheta_p": ("rtheta_p", "nCells", "nVertLevels"),
    "pressure_p": ("pressure_p", "nCells", "nVertLevels"),
    "u": ("u", "nEdges", "nVertLevels"),
    "w": ("w", "nCells", "nVertLevelsP1"),
}
ACOUSTIC_FIELDS = {"ru_p": ("ru_p", "nEdges", "nVertLevels"), "rw_p": ("rw_p", "nCells", "nVertLevelsP1")}


def write_restart(path, state: Mapping[str, np.ndarray], *, xtime: str, model_time_s: float,
                  acoustic: Optional[Mapping[str, np.ndarray]] = None) -> None:
    """Write a dycore restart: arrays in the ABI's logical [L][entity] order
    are stored as MPAS (Time, entity, level) binary32 variables."""
    from scipy.io import netcdf_file

    rho = np.asarray(state["rho_zz"])
    L, C = rho.shape
    E = np.asarray(state["ru"]).shape[1]
    with netcdf_file(str(path), "w", version=2) as f:
        f.createDimension("Time", None)
        f.createDimension("StrLen", 64)
        f.createDimension("nCells", C)
        f.createDimension("nEdges", E)
        f.createDimension("nVertLevels", L)
        f.createDimension("nVertLevelsP1", L + 1)
        f.model_time_s = float(model_time_s)
        xt = f.createVariable("xtime", "c", ("Time", "StrLen"))
        xt[0, :] = np.array(list(xtime.ljust(64)[:64]), dtype="S1")
        fields = dict(RESTART_FIELDS)
        if acoustic:
            fields.update(ACOUSTIC_FIELDS)
        src = {**state, **(acoustic or {})}
        for key, (name, ent, lev) in fields.items():
            a = np.asarray(src[key], dtype=np.float32)
            v = f.createVariable(name, "f4", ("Time", ent, lev))
            v[0] = a.T
        q = state.get("scalars")
        if q is not None and np.asarray(q).size:
            q = np.asarray(q, dtype=np.float32)
            f.createDimension("num_scalars", q.shape[0])
            v = f.createVariable("scalars", "f4", ("Time", "nCells", "nVertLevels", "num_scalars"))
            v[0] = np.transpose(q, (2, 1, 0))


def read_restart(path) -> Tuple[Dict[str, np.ndarray], Dict[str, object]]:
    """Restart arrays in the ABI's logical order and the file's metadata."""
    arrays, attrs = _read(Path(path))
    out: Dict[str, np.ndarray] = {}
    for key, (name, _ent, _lev) in {**RESTART_FIELDS, **ACOUSTIC_FIELDS}.items():
        if name in arrays:
            out[key] = np.ascontiguousarray(arrays[name].T.astype(np.float32))
    if "scalars" in arrays:
        out["scalars"] = np.ascontiguousarray(np.transpose(arrays["scalars"], (2, 1, 0)).astype(np.float32))
    meta = dict(attrs)
    if "xtime" in arrays:
        meta["xtime"] = b"".join(arrays["xtime"].tolist()).decode().strip()
    return out, meta
