"""Adapter from the reference's own host types to this dycore's boundary.

``case_from_reference`` takes exactly the arguments of the reference's
``CudaDryDycoreDriver.from_host(mesh, state, vertical, reference, config, *,
saved_diagnostics, terrain_metrics, advection_coefficients,
reference_wind_profiles)`` (cuda_driver.py:2826-2842), read by attribute
name (duck typing: nothing from the reference is imported), and returns the
``Case`` that ``capi.build_structs`` turns into the ``hexdyn_mesh`` /
``hexdyn_vertical`` / ``hexdyn_config`` structs, plus the prognostic state
and saved diagnostics in this package's names.  ``HexDycoreDriver.
from_reference`` is the drop-in for ``from_host``.

Mapping (reference type -> field here):
  Mesh (mesh.py:1290; MPAS names as attributes or in ``arrays``, 0-based)
      -> Mesh.arrays (stored inverses and cos/sin of angleEdge formed here
         as the reference's v8.4.1 context does)
  PrognosticState (state.py:14-27) rho, rho_theta, rho_u, rho_w, scalars
      -> rho_zz, rtheta_m, ru, rw, scalars
  VerticalGrid (vertical.py:83-106) zw dzw rdzw dzu rdzu fzm fzp zgrid zz
      zxu cf1-3 -> VerticalGrid (dss is rebuilt from zgrid, as from_host
      does, cuda_driver.py:2917-2926)
  DryReferenceState (driver.py:609) rho_base, rho_theta_base, exner_base
      -> ReferenceState (theta_base = rho_theta_base / rho_base)
  TerrainMetrics (driver.py:636) zb_cell, zb3_cell (zb3 already scaled by
      config_coef_3rd_order) -> VerticalGrid.zb_cell / zb3_cell
  AdvectionCoefficients (transport.py:38) adv_coefs, adv_coefs_3rd,
      n_adv_cells_for_edge, adv_cells_for_edge -> the case's stencils (built
      from the mesh's deriv_two when None, as from_host does)
  V841ReferenceWindProfiles (dynamics_v841.py:49) u_init, v_init
  V841DryDycoreConfig (config_v841.py) config_* -> DycoreConfig; knobs this
      dycore does not implement are refused with ConfigurationRefusal
  DrySavedDiagnostics (driver.py:661) theta_m, exner, density_perturbation,
      rho_theta_perturbation, pressure_perturbation, normal_velocity,
      vertical_velocity -> theta_m, exner, rho_p, rtheta_p, pressure_p, u, w
"""

from __future__ import annotations

from dataclasses import fields
from typing import Any, Dict, Optional, Tuple

import numpy as np

from .config import ConfigurationRefusal, DycoreConfig

F32 = np.float32

MESH_FIELDS = (
    "xCell", "yCell", "zCell", "latCell", "lonCell", "xVertex", "yVertex", "zVertex", "latVertex", "lonVertex",
    "xEdge", "yEdge", "zEdge", "latEdge", "lonEdge", "cellsOnEdge", "verticesOnEdge", "edgesOnCell", "cellsOnCell",
    "verticesOnCell", "nEdgesOnCell", "cellsOnVertex", "edgesOnVertex", "kiteAreasOnVertex", "areaTriangle",
    "areaCell", "dcEdge", "dvEdge", "angleEdge", "fCell", "fEdge", "fVertex", "edgesOnEdge", "weightsOnEdge",
    "nEdgesOnEdge", "kiteForCell", "nominalMinDc", "meshDensity", "deriv_two")
REQUIRED_MESH = ("cellsOnEdge", "verticesOnEdge", "edgesOnCell", "cellsOnCell", "verticesOnCell", "nEdgesOnCell",
                 "cellsOnVertex", "edgesOnVertex", "kiteAreasOnVertex", "areaTriangle", "areaCell", "dcEdge",
                 "dvEdge", "angleEdge", "fEdge", "fVertex", "edgesOnEdge", "weightsOnEdge", "nEdgesOnEdge",
                 "kiteForCell", "xCell", "yCell", "zCell")

# reference knobs outside DycoreConfig and the only values this dycore runs
FIXED_KNOBS = {
    "config_positive_definite": False,
    "config_les_model": "none",
    "config_epssm": 0.0,
    "config_terrain_following": True,
    "config_mix_scalars": False,
    "config_scalar_adv_order": 3,
    "config_theta_adv_order": 3,
}


def _mesh_array(mesh: Any, name: str) -> Optional[np.ndarray]:
    """An MPAS field of a reference mesh (attribute, then ``arrays``)."""
    if hasattr(mesh, name):
        return np.asarray(getattr(mesh, name))
    arrays = getattr(mesh, "arrays", None)
    if isinstance(arrays, dict) and name in arrays:
        return np.asarray(arrays[name])
    return None


def dycore_config(config: Any) -> DycoreConfig:
    """DycoreConfig from a reference config object (same namelist names)."""
    for knob, value in FIXED_KNOBS.items():
        if hasattr(config, knob):
            got = getattr(config, knob)
            if got is None and knob == "config_terrain_following":
                continue
            if got != value:
                raise ConfigurationRefusal(knob, got, "this dycore implements only the default branch",
                                           f"{knob}={value!r}")
    release = getattr(config, "source_release", "v8.4.1")
    if release != "v8.4.1":
        raise ConfigurationRefusal("source_release", release, "only the v8.4.1 numerics are built",
                                   "V841DryDycoreConfig(source_release='v8.4.1')")
    kw = {}
    for f in fields(DycoreConfig):
        if f.name.startswith("config_") and hasattr(config, f.name):
            v = getattr(config, f.name)
            if v is not None:
                kw[f.name] = type(f.default)(v) if isinstance(f.default, (int, float, bool, str)) else v
    if hasattr(config, "numerics"):
        kw["numerics"] = str(config.numerics)
    cfg = DycoreConfig(**kw)
    cfg.validate()
    return cfg


def case_from_reference(mesh: Any, state: Any, vertical: Any, reference: Any, config: Any, *,
                        saved_diagnostics: Any = None, terrain_metrics: Any,
                        advection_coefficients: Any = None, reference_wind_profiles: Any = None,
                        name: str = "reference") -> Tuple[Any, Dict[str, np.ndarray], Optional[Dict[str, np.ndarray]]]:
    """(Case, state, saved) from the reference's from_host arguments."""
    from .case import Case, _len_disp
    from . import initial as hinit
    from . import setup as hsetup
    from .mesh import Mesh

    cfg = dycore_config(config)
    if reference_wind_profiles is None:
        # v8.4.1 always subtracts the reference Coriolis flow (cuda_driver.py:2869-2875)
        raise ConfigurationRefusal("u_init", None, "v8.4.1 subtracts the initialized reference Coriolis flow",
                                   "reference_wind_profiles with u_init and v_init")
    if terrain_metrics is None:
        raise ConfigurationRefusal("terrain_metrics", None, "the omega conversion needs zb_cell/zb3_cell",
                                   "TerrainMetrics(zb_cell=..., zb3_cell=...)")
    arrays = {}
    for n in MESH_FIELDS:
        a = _mesh_array(mesh, n)
        if a is not None:
            arrays[n] = a
    missing = [n for n in REQUIRED_MESH if n not in arrays]
    if missing:
        raise ConfigurationRefusal("mesh", missing, "the mesh lacks MPAS fields the dycore reads", "a full MPAS mesh")
    for n in ("cellsOnEdge", "verticesOnEdge", "edgesOnCell", "cellsOnCell", "verticesOnCell", "cellsOnVertex",
              "edgesOnVertex", "edgesOnEdge", "nEdgesOnCell", "nEdgesOnEdge", "kiteForCell"):
        arrays[n] = np.asarray(arrays[n], dtype=np.int64)
    radius = float(getattr(mesh, "sphere_radius", getattr(mesh, "sphereRadius", 6371229.0)))
    m = Mesh(arrays=arrays, sphere_radius=radius)

    rho = np.asarray(state.rho, dtype=F32)
    L, C = rho.shape
    zz = np.asarray(vertical.zz, dtype=F32)
    rdzw = np.asarray(vertical.rdzw, dtype=F32)
    zgrid = np.asarray(vertical.zgrid, dtype=F32)
    dzw = np.asarray(getattr(vertical, "dzw", 1.0 / rdzw.astype(np.float64)), dtype=np.float64)
    zw = np.asarray(getattr(vertical, "zw", np.concatenate([[0.0], np.cumsum(dzw)])), dtype=np.float64)
    zb_cell = np.asarray(terrain_metrics.zb_cell, dtype=F32)
    zb3_cell = np.asarray(terrain_metrics.zb3_cell, dtype=F32)
    if zb_cell.shape != (L + 1, C, m.max_edges) or zb3_cell.shape != zb_cell.shape:
        raise ValueError(f"terrain metrics shape {zb_cell.shape} != {(L + 1, C, m.max_edges)}")
    grid = hinit.VerticalGrid(
        zw=zw, zgrid=zgrid, zz=zz, zxu=np.asarray(vertical.zxu, dtype=F32), rdzw=rdzw,
        rdzu=np.asarray(vertical.rdzu, dtype=F32), fzm=np.asarray(vertical.fzm, dtype=F32),
        fzp=np.asarray(vertical.fzp, dtype=F32), cf1=float(F32(vertical.cf1)), cf2=float(F32(vertical.cf2)),
        cf3=float(F32(vertical.cf3)), zb_cell=zb_cell, zb3_cell=zb3_cell, dzw=dzw,
        dzu=np.asarray(vertical.dzu, dtype=F32) if hasattr(vertical, "dzu") else None)
    rb = np.asarray(reference.rho_base, dtype=F32)
    rtb = np.asarray(reference.rho_theta_base, dtype=F32)
    ref = hinit.ReferenceState(rho_base=rb, rtheta_base=rtb, exner_base=np.asarray(reference.exner_base, dtype=F32),
                               theta_base=(rtb / rb).astype(F32))
    scalars = np.asarray(state.scalars, dtype=F32)
    if scalars.ndim != 3:
        scalars = scalars.reshape(-1, L, C)
    init = hinit.InitialState(rho_zz=rho, rtheta_m=np.asarray(state.rho_theta, dtype=F32),
                              ru=np.asarray(state.rho_u, dtype=F32), rw=np.asarray(state.rho_w, dtype=F32),
                              scalars=scalars)
    if advection_coefficients is None:
        deriv = arrays["deriv_two"] if "deriv_two" in arrays else hsetup.deriv_two(m)
        adv, adv3, nadv, advc = hsetup.advection_coefficients(m, deriv)
    else:
        a = advection_coefficients
        adv = np.asarray(a.adv_coefs, dtype=F32)
        adv3 = np.asarray(a.adv_coefs_3rd, dtype=F32)
        nadv = np.asarray(a.n_adv_cells_for_edge, dtype=np.int64)
        advc = np.asarray(a.adv_cells_for_edge, dtype=np.int64)
    etp, etm, ewp, ewm = hsetup.acoustic_offcentering(
        grid.rdzw, minimum=cfg.config_epssm_minimum, maximum=cfg.config_epssm_maximum,
        transition_bottom_z=cfg.config_epssm_transition_bottom_z, transition_top_z=cfg.config_epssm_transition_top_z)
    cos_angle, sin_angle = hsetup.angle_cos_sin(m.angleEdge)
    n_scalars = scalars.shape[0]
    moist = bool(getattr(config, "config_moist_physics", cfg.config_moist_physics))
    case = Case(
        name=name, mesh=m, grid=grid, reference=ref, state=init, config=cfg, n_scalars=n_scalars,
        n_moist=n_scalars if (moist and n_scalars) else 0, adv_coefs=adv, adv_coefs_3rd=adv3, n_adv=nadv,
        adv_cells=advc, inv_area_cell=hsetup.stored_inverse(m.areaCell),
        inv_area_triangle=hsetup.stored_inverse(m.areaTriangle), inv_dc_edge=hsetup.stored_inverse(m.dcEdge),
        inv_dv_edge=hsetup.stored_inverse(m.dvEdge), cos_angle_edge=cos_angle, sin_angle_edge=sin_angle,
        etp=etp, etm=etm, ewp=ewp, ewm=ewm,
        dss=hsetup.vertical_velocity_damping(zgrid, xnutr=cfg.config_xnutr, zd=cfg.config_zd),
        u_init=np.asarray(reference_wind_profiles.u_init, dtype=F32),
        v_init=np.asarray(reference_wind_profiles.v_init, dtype=F32),
        len_disp=_len_disp(m, cfg))
    st = {"rho_zz": init.rho_zz, "rtheta_m": init.rtheta_m, "ru": init.ru, "rw": init.rw, "scalars": init.scalars}
    saved = None
    if saved_diagnostics is not None:
        s = saved_diagnostics
        saved = {"theta_m": s.theta_m, "exner": s.exner, "rho_p": s.density_perturbation,
                 "rtheta_p": s.rho_theta_perturbation, "pressure_p": s.pressure_perturbation,
                 "u": s.normal_velocity, "w": s.vertical_velocity}
        saved = {k: np.asarray(v, dtype=F32) for k, v in saved.items()}
    return case, st, saved
