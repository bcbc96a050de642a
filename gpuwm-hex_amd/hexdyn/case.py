"""Named synthetic workloads: mesh + vertical grid + reference + initial state.

Workloads (sizes, seeds):
  x1.2562 dry   : icosahedron subdivided 4x, 20 seeded Lloyd iterations, L55 to
                  30 km, random terrain 0-4 km, JW jet + theta bubble,
                  dt 1800 s, RK3, 2 acoustic substeps, 3 dynamics subcycles.
  x1.40962 moist: subdivided 6x, same recipe, qv from RH ~ U(0.3, 0.9), five
                  log-normal condensates where RH > 0.8, FCT transport, dt 720 s.
  x1.10242 moist: subdivided 5x, dt 1440 s (the ensemble mesh).
  x1.642, x1.162: subdivided 3x and 2x (small test cases), dt 3600 s and 7200 s.
  r642 moist    : 642 seeded random generators, 31 Lloyd iterations: a ragged
                  mesh with 110 pentagons, 434 hexagons and 98 heptagons
                  (maxEdges 7), dt 3600 s.
  x1.163842 qv  : subdivided 7x, water vapour only (seeded RH field), dt 360 s
                  (the scale-up mesh; about 2.5 min of host time to build).
Override reference_wind=(u0, v0) gives a non-zero u_init/v_init profile.
Seeds: mesh 1, terrain 2, theta perturbation 3, RH 4, condensates 5.
"""

from __future__ import annotations

import hashlib
import os
import pickle
from dataclasses import dataclass, replace
from pathlib import Path
from typing import Optional

import numpy as np

from .config import DycoreConfig
from .mesh import Mesh, build_mesh, icosahedral_mesh, random_points
from . import initial as hinit
from . import setup as hsetup

F32 = np.float32


@dataclass
class Case:
    name: str
    mesh: Mesh
    grid: hinit.VerticalGrid
    reference: hinit.ReferenceState
    state: hinit.InitialState
    config: DycoreConfig
    n_scalars: int
    n_moist: int
    adv_coefs: np.ndarray
    adv_coefs_3rd: np.ndarray
    n_adv: np.ndarray
    adv_cells: np.ndarray
    inv_area_cell: np.ndarray
    inv_area_triangle: np.ndarray
    inv_dc_edge: np.ndarray
    inv_dv_edge: np.ndarray
    cos_angle_edge: np.ndarray
    sin_angle_edge: np.ndarray
    etp: np.ndarray
    etm: np.ndarray
    ewp: np.ndarray
    ewm: np.ndarray
    dss: np.ndarray
    u_init: np.ndarray
    v_init: np.ndarray
    len_disp: float

    @property
    def n_levels(self) -> int:
        return int(self.grid.zz.shape[0])


WORKLOADS = {
    "x1.2562-dry": dict(level=4, moist=False, n_scalars=0, dt=1800.0),
    "x1.2562-moist": dict(level=4, moist=True, n_scalars=6, dt=1800.0),
    "x1.10242-moist": dict(level=5, moist=True, n_scalars=6, dt=1440.0),
    "x1.40962-moist": dict(level=6, moist=True, n_scalars=6, dt=720.0),
    "x1.40962-dry": dict(level=6, moist=False, n_scalars=0, dt=720.0),
    "x1.163842-qv": dict(level=7, moist=True, n_scalars=1, dt=360.0),
    "x1.162-moist": dict(level=2, moist=True, n_scalars=6, dt=7200.0),
    "r642-moist": dict(level=0, random_points=642, moist=True, n_scalars=6, dt=3600.0),
    "x1.642-dry": dict(level=3, moist=False, n_scalars=0, dt=3600.0),
    "x1.642-moist": dict(level=3, moist=True, n_scalars=6, dt=3600.0),
}


def _cache_dir() -> Path:
    root = os.environ.get("HEXDYN_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "hexdyn")
    p = Path(root)
    try:
        p.mkdir(parents=True, exist_ok=True)
    except OSError:
        import tempfile

        p = Path(tempfile.gettempdir()) / "hexdyn-cache"
        p.mkdir(parents=True, exist_ok=True)
    return p


def build_case(name: str = "x1.2562-dry", *, nlev: int = 55, config: Optional[DycoreConfig] = None,
               use_cache: bool = True, **overrides) -> Case:
    spec = dict(WORKLOADS[name])
    spec.update(overrides)
    cfg = config or DycoreConfig(config_dt=spec["dt"])
    key = hashlib.sha256(repr((name, nlev, sorted(spec.items()), _generator_digest())).encode()).hexdigest()[:16]
    path = _cache_dir() / f"case-{name}-{key}.pkl"
    if use_cache and path.exists():
        with open(path, "rb") as f:
            case = pickle.load(f)
        return replace(case, config=cfg, len_disp=_len_disp(case.mesh, cfg))
    say = _progress(name)
    if spec.get("random_points"):
        mesh = build_mesh(random_points(spec["random_points"], seed=1, lloyd_iterations=31))
    else:
        mesh = icosahedral_mesh(spec["level"], seed=1, lloyd_iterations=20)
    say(f"mesh {mesh.n_cells} cells")
    deriv = hsetup.deriv_two(mesh)
    terrain = hinit.random_terrain(mesh, seed=2, hmax=spec.get("hmax", 4000.0)) * spec.get("terrain_scale", 1.0)
    grid = hinit.build_vertical_grid(mesh, terrain, nlev=nlev, coef_3rd_order=cfg.config_coef_3rd_order, deriv=deriv)
    reference, state = hinit.build_reference_and_state(mesh, grid, moist=spec["moist"], n_scalars=spec["n_scalars"],
                                                       theta_amplitude=spec.get("theta_amplitude", 0.5))
    say("vertical grid, reference and initial state")
    adv, adv3, nadv, advc = hsetup.advection_coefficients(mesh, deriv)
    etp, etm, ewp, ewm = hsetup.acoustic_offcentering(
        grid.rdzw, minimum=cfg.config_epssm_minimum, maximum=cfg.config_epssm_maximum,
        transition_bottom_z=cfg.config_epssm_transition_bottom_z,
        transition_top_z=cfg.config_epssm_transition_top_z)
    dss = hsetup.vertical_velocity_damping(grid.zgrid, xnutr=cfg.config_xnutr, zd=cfg.config_zd)
    cos_angle, sin_angle = hsetup.angle_cos_sin(mesh.angleEdge)
    L = nlev
    u_init, v_init = np.zeros(L, dtype=F32), np.zeros(L, dtype=F32)
    if spec.get("reference_wind"):
        # a non-zero reference wind profile (u0, v0): u_init ramps from u0/2 at
        # the surface to u0 at the top, v_init from v0 down to 0
        u0, v0 = spec["reference_wind"]
        z = np.arange(L, dtype=np.float64) / max(L - 1, 1)
        u_init = (u0 * (0.5 + 0.5 * z)).astype(F32)
        v_init = (v0 * (1.0 - z)).astype(F32)
    case = Case(
        name=name, mesh=mesh, grid=grid, reference=reference, state=state, config=cfg,
        n_scalars=spec["n_scalars"], n_moist=spec["n_scalars"] if spec["moist"] else 0,
        adv_coefs=adv, adv_coefs_3rd=adv3, n_adv=nadv, adv_cells=advc,
        inv_area_cell=hsetup.stored_inverse(mesh.areaCell),
        inv_area_triangle=hsetup.stored_inverse(mesh.areaTriangle),
        inv_dc_edge=hsetup.stored_inverse(mesh.dcEdge),
        inv_dv_edge=hsetup.stored_inverse(mesh.dvEdge),
        cos_angle_edge=cos_angle, sin_angle_edge=sin_angle,
        etp=etp, etm=etm, ewp=ewp, ewm=ewm, dss=dss,
        u_init=u_init, v_init=v_init,
        len_disp=_len_disp(mesh, cfg),
    )
    if use_cache:
        tmp = path.with_suffix(".tmp")
        with open(tmp, "wb") as f:
            pickle.dump(case, f, protocol=pickle.HIGHEST_PROTOCOL)
        os.replace(tmp, path)
    return case


def _generator_digest() -> str:
    """sha256 of the source of every module that builds a case, so a cached
    case is never served after the generator changed."""
    here = Path(__file__).resolve().parent
    h = hashlib.sha256()
    for mod in ("case.py", "mesh.py", "initial.py", "setup.py", "config.py"):
        h.update((here / mod).read_bytes())
    return h.hexdigest()


def _progress(name: str):
    """Stage messages on stderr while a case is built (HEXDYN_VERBOSE=1)."""
    if not os.environ.get("HEXDYN_VERBOSE"):
        return lambda msg: None
    import sys
    import time

    t0 = time.perf_counter()

    def say(msg: str) -> None:
        print(f"# case {name}: {msg} ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)

    return say


def _len_disp(mesh: Mesh, cfg: DycoreConfig) -> float:
    if cfg.config_len_disp > 0.0:
        return float(F32(cfg.config_len_disp))
    return float(F32(mesh.arrays["nominalMinDc"][0]))


def initial_state_arrays(case: Case) -> dict:
    """Prognostic initial fields (binary32, logical layout)."""
    s = case.state
    return {"rho_zz": s.rho_zz, "rtheta_m": s.rtheta_m, "ru": s.ru, "rw": s.rw, "scalars": s.scalars}
