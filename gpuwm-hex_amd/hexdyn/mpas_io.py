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
        if "deriv_two" in out:  # second derivatives: 1 / length^2
            out["deriv_two"] = out["deriv_two"] / (radius * radius)
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

    with netcdf_file(str(path), "w", version=2) as f:
        _write_mesh(f, mesh, unit_sphere=unit_sphere)


def _write_mesh(f, mesh: Mesh, *, unit_sphere: bool = False) -> None:
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
    f.on_a_sphere = "YES"
    f.sphere_radius = 1.0 if unit_sphere else float(R)
    for name, n in dims.items():
        f.createDimension(name, n)
    f.createDimension("TWO", 2)
    if "deriv_two" in a:
        d2 = np.asarray(a["deriv_two"], dtype=np.float64)
        f.createDimension("FIFTEEN", d2.shape[2])
        v = f.createVariable("deriv_two", "f8", ("nEdges", "TWO", "FIFTEEN"))
        v[:] = d2 * (R * R) if unit_sphere else d2
    for name, arr in a.items():
        if name in ("nominalMinDc", "deriv_two"):
            continue
        if name in layout:
            d = shape_dims[layout[name]]
        elif arr.ndim == 1 and arr.size == mesh.n_cells:
            d = ("nCells",)
        elif arr.ndim == 1 and arr.size == mesh.n_edges:
            d = ("nEdges",)
        elif arr.ndim == 1 and arr.size == mesh.n_vertices:
            d = ("nVertices",)
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


# ------------------------------------------------------------------ init files
#: MPAS names of the scalar slots, in the dycore's order (qv first: the moist
#: coefficients sum the leading moisture species)
SCALAR_NAMES = ("qv", "qc", "qr", "qi", "qs", "qg")
#: vertical metrics an initialized file must carry (driver.py:1456-1459);
#: its dss is not read: the v8.4.1 core recomputes the Rayleigh profile from
#: zgrid at start-up (driver.py:2416-2421)
VERTICAL_REQUIRED = ("zgrid", "rdzw", "dzu", "rdzu", "fzm", "fzp", "zz", "zxu", "zb", "zb3",
                     "cf1", "cf2", "cf3")


def write_mpas_init(case, path, *, unit_sphere: bool = False, public_w=None) -> None:
    """Write a case as an MPAS-style initialized file (mesh + vertical metrics +
    reference + state), 64-bit offset netCDF.

    The MPAS variables of an init file are written under their MPAS names and
    dimensions: ``zgrid (nCells, nVertLevelsP1)``, ``zz``/``dss (nCells,
    nVertLevels)``, ``zxu (nEdges, nVertLevels)``, 1-D ``rdzw dzu rdzu fzm fzp``,
    ``zb``/``zb3 (nEdges, TWO, nVertLevelsP1)`` (zb3 unscaled), scalars
    ``cf1 cf2 cf3``, ``rho_base theta_base rho_zz theta_m u w`` and one variable
    per scalar species, with a leading ``Time`` record.  The momenta ``ru``,
    ``rw`` and the binary32 ``rtheta_m``, ``rtheta_base``, ``exner_base`` are
    added as a model dump carries them, so a file read back reproduces the
    case bit for bit.

    With ``public_w`` (the ``[L+1][C]`` vertical velocity of the state) the
    file is written the way init_atmosphere writes one instead: physical
    ``rho = rho_zz * zz`` and dry ``theta = theta_m / (1 + (R_v/R_d) qv)``,
    ``u`` and ``w``, and none of ``rho_zz theta_m rtheta_m ru rw``.
    """
    from scipy.io import netcdf_file

    g, ref, st, m = case.grid, case.reference, case.state, case.mesh
    if g.zb_edge is None or g.zb3_edge is None or g.dzu is None:
        raise MpasFileError("the case's vertical grid carries no per-edge zb/zb3 (rebuild the case)")
    L = case.n_levels
    c1, c2 = m.cellsOnEdge[:, 0], m.cellsOnEdge[:, 1]
    rho = st.rho_zz.astype(np.float32)
    u = (st.ru / (np.float32(0.5) * (rho[:, c1] + rho[:, c2]))).astype(np.float32)
    with netcdf_file(str(path), "w", version=2) as f:
        f.createDimension("Time", None)  # the record dimension comes first in this format
        _write_mesh(f, m, unit_sphere=unit_sphere)
        f.createDimension("nVertLevels", L)
        f.createDimension("nVertLevelsP1", L + 1)

        def var(name, dims, data, kind="f4"):
            v = f.createVariable(name, kind, dims)
            if dims and dims[0] == "Time":
                v[0] = data
            elif dims:
                v[:] = data
            else:
                v[...] = data  # a 0-d variable

        var("zgrid", ("nCells", "nVertLevelsP1"), g.zgrid.T)
        var("zz", ("nCells", "nVertLevels"), g.zz.T)
        var("zxu", ("nEdges", "nVertLevels"), g.zxu.T)
        var("dss", ("nCells", "nVertLevels"), np.asarray(case.dss, dtype=np.float32).T)
        for name, a in (("rdzw", g.rdzw), ("dzu", g.dzu), ("rdzu", g.rdzu), ("fzm", g.fzm), ("fzp", g.fzp),
                        ("u_init", case.u_init), ("v_init", case.v_init)):
            var(name, ("nVertLevels",), np.asarray(a, dtype=np.float32))
        var("zb", ("nEdges", "TWO", "nVertLevelsP1"), np.transpose(g.zb_edge, (2, 1, 0)))
        var("zb3", ("nEdges", "TWO", "nVertLevelsP1"), np.transpose(g.zb3_edge, (2, 1, 0)))
        for name, x in (("cf1", g.cf1), ("cf2", g.cf2), ("cf3", g.cf3)):
            var(name, (), np.float32(x))
        cl = ("Time", "nCells", "nVertLevels")
        var("rho_base", cl, ref.rho_base.T)
        var("theta_base", cl, ref.theta_base.T)
        var("rtheta_base", cl, ref.rtheta_base.T)
        var("exner_base", cl, ref.exner_base.T)
        theta_m = (st.rtheta_m / rho).astype(np.float32)
        if public_w is None:
            var("rho_zz", cl, rho.T)
            var("rtheta_m", cl, st.rtheta_m.T)
            var("theta_m", cl, theta_m.T)
            var("ru", ("Time", "nEdges", "nVertLevels"), st.ru.T)
            var("rw", ("Time", "nCells", "nVertLevelsP1"), st.rw.T)
        else:
            f32 = np.float32
            rvord = f32(461.6) / f32(287.0)
            qv = st.scalars[0] if case.n_scalars else np.zeros_like(rho)
            var("rho", cl, (rho * g.zz).astype(f32).T)
            var("theta", cl, (theta_m / (f32(1.0) + rvord * qv)).astype(f32).T)
            var("w", ("Time", "nCells", "nVertLevelsP1"), np.asarray(public_w, dtype=f32).T)
        var("u", ("Time", "nEdges", "nVertLevels"), u.T)
        for i in range(case.n_scalars):
            var(SCALAR_NAMES[i] if i < len(SCALAR_NAMES) else f"scalar{i}", cl, st.scalars[i].T)
        f.moist = 1 if case.n_moist > 0 else 0
        f.config_dt = float(case.config.config_dt)


def _levels_first(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a).T)


def w_to_rw(mesh, grid, rho_zz: np.ndarray, ru: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Vertical momentum of a vertical velocity on the terrain-following grid,
    the inverse of the recovery of w (atm_init_coupled_diagnostics; the
    reference's driver.py:1791-1854 with its correct_vertical_momentum):

        rw(k) = (w(k) * rho_i(k)) * zz_i(k)
                - sum_s ((sign_s * (zb_s(k) + sign(flux) zb3_s(k))) * flux) * zz_i(k)

    for interior interfaces 1 <= k < L, with rho_i = fzm rho(k) + fzp rho(k-1),
    zz_i = fzm zz(k) + fzp zz(k-1), flux = fzm ru(k) + fzp ru(k-1) on the
    cell's edge in slot s, sign_s = +1 when the cell is the edge's first cell,
    and zb3 the grid's slot metric (scaled by the third-order coefficient, as
    the recovery reads it); rw = 0 at the surface and the top.  Binary32,
    slots in edgesOnCell order.  The reference's correct_vertical_momentum
    is a compiled engine that is not in the reference tree, so its operation
    order is not pinned here; the conversion is pinned by its round trip
    through the recovery of w."""
    f32 = np.float32
    L, C = rho_zz.shape
    fzm = np.asarray(grid.fzm, dtype=f32)[:, None]
    fzp = np.asarray(grid.fzp, dtype=f32)[:, None]
    zz = np.asarray(grid.zz, dtype=f32)
    rw = np.zeros((L + 1, C), dtype=f32)
    rho_i = fzm[1:] * rho_zz[1:] + fzp[1:] * rho_zz[:-1]
    zz_i = fzm[1:] * zz[1:] + fzp[1:] * zz[:-1]
    rw[1:L] = (w[1:L] * rho_i) * zz_i
    coe = np.asarray(mesh.cellsOnEdge)
    eoc = np.asarray(mesh.edgesOnCell)
    n = np.asarray(mesh.nEdgesOnCell)
    cells = np.arange(C)
    for s in range(eoc.shape[1]):
        live = n > s
        e = np.where(live, eoc[:, s], 0)
        sign = np.where(coe[e, 0] == cells, f32(1.0), f32(-1.0)).astype(f32)
        flux = fzm[1:] * ru[1:, e] + fzp[1:] * ru[:-1, e]
        zb = np.asarray(grid.zb_cell[1:L, :, s], dtype=f32)
        zb3 = np.asarray(grid.zb3_cell[1:L, :, s], dtype=f32)
        term = ((sign[None, :] * (zb + np.copysign(f32(1.0), flux) * zb3)) * flux) * zz_i
        rw[1:L] = np.where(live[None, :], rw[1:L] - term, rw[1:L]).astype(f32)
    return rw


def case_from_mpas(init_path, *, grid_path=None, invariant_path=None, config=None, name: str = "mpas",
                   cold_start: bool = False):
    """A dycore case from MPAS files: mesh (grid file, or the init file's own
    mesh variables), vertical metrics (the invariant file when given, the
    reference's authority order, driver.py:1418-1455), reference state and
    prognostic state (driver.py:1583-1700).

    The per-edge ``zb``/``zb3`` become per-cell-slot metrics with zb3 scaled
    by ``config_coef_3rd_order`` in binary32.  ``rho_zz`` and ``theta_m`` are
    used as stored (``rtheta_m`` when the file carries it, else
    ``rho_zz * theta_m`` in binary32).  An init_atmosphere file's public
    fields are converted as the reference does (driver.py:1642-1681): physical
    ``rho`` to ``rho_zz = rho / zz`` and dry ``theta`` to
    ``theta_m = theta * (1 + rvord qv)`` with ``rvord = 461.6f / 287.0f``;
    with ``cold_start`` they are used even where the file also holds
    ``rho_zz``/``theta_m``/``ru``/``rw`` (atm_init_coupled_diagnostics
    rebuilds the coupled state; the saved diagnostics are always recomputed
    from the prognostics when a driver takes the state).  ``ru``/``rw`` are
    read when present, else ``ru = (0.5 u) (rho_zz[c1] + rho_zz[c2])``
    (driver.py:1760-1766, the half binding to u first) and ``rw`` from ``w``
    by ``w_to_rw`` (driver.py:1791-1854).
    ``rtheta_base``/``exner_base`` are read when present, else formed as
    ``rho_base * theta_base`` and ``(zz R_d/p0 rtheta_base)^(R_d/c_v)``
    (binary64, rounded once: parity unpinned against MPAS's powf).
    """
    from .config import DycoreConfig
    from .case import Case, _len_disp
    from . import initial as hinit
    from . import setup as hsetup

    init_path = Path(init_path)
    mesh = mesh_from_mpas(grid_path if grid_path is not None else init_path)
    arrays, attrs = _read(init_path)
    vert = dict(arrays)
    if invariant_path is not None:
        inv, _ = _read(Path(invariant_path))
        vert.update({k: v for k, v in inv.items() if k in VERTICAL_REQUIRED})
    missing = [n for n in VERTICAL_REQUIRED if n not in vert]
    if missing:
        raise MpasFileError(f"native vertical metrics incomplete: missing {missing}")
    cfg = config or DycoreConfig(config_dt=float(attrs.get("config_dt", 720.0)))
    f32 = np.float32
    zgrid = _levels_first(vert["zgrid"]).astype(f32)
    zz = _levels_first(vert["zz"]).astype(f32)
    L = zz.shape[0]
    if zgrid.shape != (L + 1, mesh.n_cells):
        raise MpasFileError("zgrid dimensions disagree with zz and the mesh")
    rdzw = np.asarray(vert["rdzw"], dtype=f32)
    dzw = 1.0 / rdzw.astype(np.float64)
    zw = np.concatenate([[0.0], np.cumsum(dzw)])
    zb_edge = np.transpose(np.asarray(vert["zb"], dtype=f32), (2, 1, 0)).copy()
    zb3_edge = np.transpose(np.asarray(vert["zb3"], dtype=f32), (2, 1, 0)).copy()
    if zb_edge.shape != (L + 1, 2, mesh.n_edges):
        raise MpasFileError("zb/zb3 dimensions disagree with the mesh")
    zb_cell, zb3_cell = hinit.terrain_slot_metrics(mesh, zb_edge, zb3_edge, cfg.config_coef_3rd_order)
    grid = hinit.VerticalGrid(
        zw=zw, zgrid=zgrid, zz=zz, zxu=_levels_first(vert["zxu"]).astype(f32), rdzw=rdzw,
        rdzu=np.asarray(vert["rdzu"], dtype=f32), fzm=np.asarray(vert["fzm"], dtype=f32),
        fzp=np.asarray(vert["fzp"], dtype=f32), cf1=float(f32(vert["cf1"])), cf2=float(f32(vert["cf2"])),
        cf3=float(f32(vert["cf3"])), zb_cell=zb_cell, zb3_cell=zb3_cell, dzw=dzw, zb_edge=zb_edge,
        zb3_edge=zb3_edge, dzu=np.asarray(vert["dzu"], dtype=f32))

    def cell_field(n):
        if n not in arrays:
            raise MpasFileError(f"missing MPAS variable {n}")
        return _levels_first(arrays[n]).astype(f32)

    rho_base, theta_base = cell_field("rho_base"), cell_field("theta_base")
    if "rtheta_base" in arrays:
        rtheta_base = cell_field("rtheta_base")
    else:
        rtheta_base = (rho_base * theta_base).astype(f32)
    if "exner_base" in arrays:
        exner_base = cell_field("exner_base")
    else:
        rgas, cp = 287.0, 1004.5
        arg = zz.astype(np.float64) * (rgas / 1.0e5) * rtheta_base.astype(np.float64)
        exner_base = np.power(arg, rgas / (cp - rgas)).astype(f32)
    reference = hinit.ReferenceState(rho_base=rho_base, rtheta_base=rtheta_base, exner_base=exner_base,
                                     theta_base=theta_base)
    f32 = np.float32
    if "rho_zz" in arrays and not (cold_start and "rho" in arrays):
        rho = cell_field("rho_zz")
    elif "rho" in arrays:
        rho = (cell_field("rho") / zz).astype(f32)
    else:
        raise MpasFileError("neither rho_zz nor the public rho is present")
    if "theta_m" in arrays and not (cold_start and "theta" in arrays):
        rtheta = cell_field("rtheta_m") if ("rtheta_m" in arrays and not cold_start) else \
            (rho * cell_field("theta_m")).astype(f32)
    elif "theta" in arrays:
        if "qv" not in arrays:
            raise MpasFileError("theta_m from the public theta needs qv")
        rvord = f32(461.6) / f32(287.0)
        theta_m = (cell_field("theta") * (f32(1.0) + rvord * cell_field("qv"))).astype(f32)
        rtheta = (rho * theta_m).astype(f32)
    else:
        raise MpasFileError("neither theta_m nor the public theta is present")
    c1, c2 = mesh.cellsOnEdge[:, 0], mesh.cellsOnEdge[:, 1]
    if "ru" in arrays and not cold_start:
        ru = _levels_first(arrays["ru"]).astype(f32)
    elif "u" in arrays:
        u = _levels_first(arrays["u"]).astype(f32)
        ru = ((f32(0.5) * u) * (rho[:, c1] + rho[:, c2])).astype(f32)
    else:
        raise MpasFileError("neither ru nor u is present")
    if "rw" in arrays and not cold_start:
        rw = _levels_first(arrays["rw"]).astype(f32)
    elif "w" in arrays:
        rw = w_to_rw(mesh, grid, rho, ru, _levels_first(arrays["w"]).astype(f32))
    else:
        raise MpasFileError("neither rw nor w is present")
    names = [n for n in SCALAR_NAMES if n in arrays]
    names += sorted((n for n in arrays if n.startswith("scalar") and n[6:].isdigit()), key=lambda n: int(n[6:]))
    scalars = (np.stack([cell_field(n) for n in names]) if names
               else np.zeros((0, L, mesh.n_cells), dtype=f32))
    moist = bool(int(attrs.get("moist", 1 if "qv" in arrays else 0)))
    state = hinit.InitialState(rho_zz=rho, rtheta_m=rtheta, ru=ru, rw=rw, scalars=scalars)

    # the file's quadratic-fit second derivatives when it carries them (the
    # reference builds its advection stencils from the static file's
    # deriv_two), else this package's fit
    deriv = mesh.arrays["deriv_two"] if "deriv_two" in mesh.arrays else hsetup.deriv_two(mesh)
    adv, adv3, nadv, advc = hsetup.advection_coefficients(mesh, deriv)
    etp, etm, ewp, ewm = hsetup.acoustic_offcentering(
        grid.rdzw, minimum=cfg.config_epssm_minimum, maximum=cfg.config_epssm_maximum,
        transition_bottom_z=cfg.config_epssm_transition_bottom_z,
        transition_top_z=cfg.config_epssm_transition_top_z)
    cos_angle, sin_angle = hsetup.angle_cos_sin(mesh.angleEdge)
    return Case(
        name=name, mesh=mesh, grid=grid, reference=reference, state=state, config=cfg,
        n_scalars=scalars.shape[0], n_moist=scalars.shape[0] if moist else 0,
        adv_coefs=adv, adv_coefs_3rd=adv3, n_adv=nadv, adv_cells=advc,
        inv_area_cell=hsetup.stored_inverse(mesh.areaCell),
        inv_area_triangle=hsetup.stored_inverse(mesh.areaTriangle),
        inv_dc_edge=hsetup.stored_inverse(mesh.dcEdge),
        inv_dv_edge=hsetup.stored_inverse(mesh.dvEdge),
        cos_angle_edge=cos_angle, sin_angle_edge=sin_angle,
        etp=etp, etm=etm, ewp=ewp, ewm=ewm,
        dss=hsetup.vertical_velocity_damping(grid.zgrid, xnutr=cfg.config_xnutr, zd=cfg.config_zd),
        u_init=np.asarray(arrays.get("u_init", np.zeros(L)), dtype=f32),
        v_init=np.asarray(arrays.get("v_init", np.zeros(L)), dtype=f32),
        len_disp=_len_disp(mesh, cfg),
    )
