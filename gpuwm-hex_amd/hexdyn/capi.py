"""ctypes mirrors of the POD structs in include/hexdyn.h and their builders.

The same structs feed libhexdyn.so (the product) and, in the test suite, the
CPU oracle, so both consume identical static arrays.
"""

from __future__ import annotations

import ctypes as ct
from typing import Dict

import numpy as np

i32p = ct.POINTER(ct.c_int32)
f32p = ct.POINTER(ct.c_float)
f64p = ct.POINTER(ct.c_double)

HEXDYN_ROW = 64
HEXDYN_MAX_EDGES = 7

HEXDYN_OK = 0
HEXDYN_E_INVALID = 1
HEXDYN_E_NONFINITE = 2
HEXDYN_E_REFUSED = 3
HEXDYN_E_HIP = -1

FIELDS = [
    "RHO_ZZ", "RTHETA_M", "RU", "RW", "SCALARS",
    "THETA_M", "EXNER", "RHO_P", "RTHETA_P", "PRESSURE_P", "U", "W",
    "V", "PV_EDGE", "KE", "DIVERGENCE", "VORTICITY",
    "TEND_RU", "TEND_RT", "TEND_RW", "TEND_RHO",
    "RU_P", "RW_P", "RHO_PP", "RTHETA_PP", "RU_AVG", "WW_AVG",
    "A_TRI", "ALPHA_TRI", "GAMMA_TRI", "COFWZ", "COFWT", "COFTZ",
    "EULER_RU", "EULER_RW",
    "FLUX_U", "FLUX_W", "SCALARS_STAGE",
    "QTOT", "CQW", "CQU",
    "KDIFF", "EULER_THETA",
]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}


class HexdynMesh(ct.Structure):
    _fields_ = [
        ("n_cells", ct.c_int32), ("n_edges", ct.c_int32), ("n_vertices", ct.c_int32),
        ("max_edges", ct.c_int32), ("vertex_degree", ct.c_int32), ("max_edges2", ct.c_int32),
        ("max_adv", ct.c_int32),
        ("cells_on_edge", i32p), ("vertices_on_edge", i32p), ("edges_on_cell", i32p),
        ("cells_on_cell", i32p), ("vertices_on_cell", i32p), ("n_edges_on_cell", i32p),
        ("kite_for_cell", i32p), ("edges_on_vertex", i32p), ("cells_on_vertex", i32p),
        ("edges_on_edge", i32p), ("n_edges_on_edge", i32p),
        ("weights_on_edge", f32p), ("dc_edge", f32p), ("dv_edge", f32p),
        ("inv_dc_edge", f32p), ("inv_dv_edge", f32p), ("area_cell", f32p), ("inv_area_cell", f32p),
        ("area_triangle", f32p), ("inv_area_triangle", f32p), ("kite_areas_on_vertex", f32p),
        ("f_edge", f32p), ("f_vertex", f32p), ("cos_angle_edge", f32p), ("sin_angle_edge", f32p),
        ("n_adv_cells_for_edge", i32p), ("adv_cells_for_edge", i32p),
        ("adv_coefs", f32p), ("adv_coefs_3rd", f32p),
        ("x_cell", f64p), ("y_cell", f64p), ("z_cell", f64p),
        ("mesh_density", f32p), ("deformation_coef_c2", f32p), ("deformation_coef_s2", f32p),
        ("deformation_coef_cs", f32p),
    ]


class HexdynVertical(ct.Structure):
    _fields_ = [
        ("n_levels", ct.c_int32),
        ("zz", f32p), ("zxu", f32p), ("rdzw", f32p), ("rdzu", f32p), ("fzm", f32p), ("fzp", f32p),
        ("cf1", ct.c_float), ("cf2", ct.c_float), ("cf3", ct.c_float),
        ("zb_cell", f32p), ("zb3_cell", f32p), ("dss", f32p),
        ("etp", f32p), ("etm", f32p), ("ewp", f32p), ("ewm", f32p),
        ("u_init", f32p), ("v_init", f32p),
        ("rho_base", f32p), ("rtheta_base", f32p), ("exner_base", f32p),
        ("zgrid", f32p),
    ]


class HexdynConfig(ct.Structure):
    _fields_ = [
        ("config_dt", ct.c_double),
        ("config_time_integration_order", ct.c_int32),
        ("config_number_of_sub_steps", ct.c_int32),
        ("config_dynamics_split_steps", ct.c_int32),
        ("config_split_dynamics_transport", ct.c_int32),
        ("config_scalar_advection", ct.c_int32),
        ("config_monotonic", ct.c_int32),
        ("config_divergence_damping", ct.c_int32),
        ("config_coef_3rd_order", ct.c_float),
        ("config_apvm_upwinding", ct.c_float),
        ("config_smdiv", ct.c_float),
        ("config_len_disp", ct.c_float),
        ("n_scalars", ct.c_int32),
        ("n_moist", ct.c_int32),
        ("numerics", ct.c_int32),
        ("config_horiz_mixing", ct.c_int32),
        ("config_visc4_2dsmag", ct.c_float),
        ("config_smagorinsky_coef", ct.c_float),
        ("config_del4u_div_factor", ct.c_float),
        ("config_h_ScaleWithMesh", ct.c_int32),
        ("config_mpas_cam_coef", ct.c_float),
        ("config_number_cam_damping_levels", ct.c_int32),
        ("config_rayleigh_damp_u", ct.c_int32),
        ("config_number_rayleigh_damp_u_levels", ct.c_int32),
        ("config_rayleigh_damp_u_timescale_days", ct.c_float),
    ]


class HexdynPhysicsView(ct.Structure):
    _fields_ = [("rho", ct.c_void_p), ("rho_u", ct.c_void_p), ("rho_theta", ct.c_void_p), ("scalars", ct.c_void_p)]


class HexdynStateView(ct.Structure):
    _fields_ = [(name, ct.c_void_p) for name in (
        "rho_zz", "rtheta_m", "ru", "rw", "scalars",
        "theta_m", "exner", "rho_p", "rtheta_p", "pressure_p", "u", "w")]


STATE_FIELDS = ("rho_zz", "rtheta_m", "ru", "rw", "scalars",
                "theta_m", "exner", "rho_p", "rtheta_p", "pressure_p", "u", "w")


class HostBundle:
    """Keeps the contiguous numpy arrays alive behind the ctypes structs."""

    def __init__(self) -> None:
        self.keep: Dict[str, np.ndarray] = {}

    def i32(self, name: str, a) -> ct.POINTER(ct.c_int32):
        arr = np.ascontiguousarray(np.asarray(a), dtype=np.int32)
        self.keep[name] = arr
        return arr.ctypes.data_as(i32p)

    def f32(self, name: str, a) -> ct.POINTER(ct.c_float):
        arr = np.ascontiguousarray(np.asarray(a), dtype=np.float32)
        self.keep[name] = arr
        return arr.ctypes.data_as(f32p)

    def f64(self, name: str, a) -> ct.POINTER(ct.c_double):
        arr = np.ascontiguousarray(np.asarray(a), dtype=np.float64)
        self.keep[name] = arr
        return arr.ctypes.data_as(f64p)


def build_structs(case) -> tuple:
    """(HexdynMesh, HexdynVertical, HexdynConfig, HostBundle) for a Case."""
    m = case.mesh
    g = case.grid
    b = HostBundle()
    mesh = HexdynMesh()
    mesh.n_cells = m.n_cells
    mesh.n_edges = m.n_edges
    mesh.n_vertices = m.n_vertices
    mesh.max_edges = m.max_edges
    mesh.vertex_degree = 3
    mesh.max_edges2 = m.edgesOnEdge.shape[1]
    mesh.max_adv = case.adv_cells.shape[1]
    mesh.cells_on_edge = b.i32("coe", m.cellsOnEdge)
    mesh.vertices_on_edge = b.i32("voe", m.verticesOnEdge)
    mesh.edges_on_cell = b.i32("eoc", m.edgesOnCell)
    mesh.cells_on_cell = b.i32("coc", m.cellsOnCell)
    mesh.vertices_on_cell = b.i32("voc", m.verticesOnCell)
    mesh.n_edges_on_cell = b.i32("neoc", m.nEdgesOnCell)
    mesh.kite_for_cell = b.i32("kfc", m.kiteForCell)
    mesh.edges_on_vertex = b.i32("eov", m.edgesOnVertex)
    mesh.cells_on_vertex = b.i32("cov", m.cellsOnVertex)
    mesh.edges_on_edge = b.i32("eoe", m.edgesOnEdge)
    mesh.n_edges_on_edge = b.i32("neoe", m.nEdgesOnEdge)
    mesh.weights_on_edge = b.f32("woe", m.weightsOnEdge)
    mesh.dc_edge = b.f32("dc", m.dcEdge)
    mesh.dv_edge = b.f32("dv", m.dvEdge)
    mesh.inv_dc_edge = b.f32("idc", case.inv_dc_edge)
    mesh.inv_dv_edge = b.f32("idv", case.inv_dv_edge)
    mesh.area_cell = b.f32("area", m.areaCell)
    mesh.inv_area_cell = b.f32("iarea", case.inv_area_cell)
    mesh.area_triangle = b.f32("tri", m.areaTriangle)
    mesh.inv_area_triangle = b.f32("itri", case.inv_area_triangle)
    mesh.kite_areas_on_vertex = b.f32("kite", m.kiteAreasOnVertex)
    mesh.f_edge = b.f32("fe", m.fEdge)
    mesh.f_vertex = b.f32("fv", m.fVertex)
    mesh.cos_angle_edge = b.f32("cosa", case.cos_angle_edge)
    mesh.sin_angle_edge = b.f32("sina", case.sin_angle_edge)
    mesh.n_adv_cells_for_edge = b.i32("nadv", case.n_adv)
    mesh.adv_cells_for_edge = b.i32("advc", case.adv_cells)
    mesh.adv_coefs = b.f32("adv", case.adv_coefs)
    mesh.adv_coefs_3rd = b.f32("adv3", case.adv_coefs_3rd)
    mesh.x_cell = b.f64("xc", m.xCell)
    mesh.y_cell = b.f64("yc", m.yCell)
    mesh.z_cell = b.f64("zc", m.zCell)
    mesh.mesh_density = b.f32("density", m.arrays.get("meshDensity", np.ones(m.n_cells)))
    if case.config.config_horiz_mixing != "off":
        from .setup import deformation_weights
        c2, s2, cs = deformation_weights(m)
        mesh.deformation_coef_c2 = b.f32("defc2", c2)
        mesh.deformation_coef_s2 = b.f32("defs2", s2)
        mesh.deformation_coef_cs = b.f32("defcs", cs)

    vert = HexdynVertical()
    vert.n_levels = g.zz.shape[0]
    vert.zz = b.f32("zz", g.zz)
    vert.zxu = b.f32("zxu", g.zxu)
    vert.rdzw = b.f32("rdzw", g.rdzw)
    vert.rdzu = b.f32("rdzu", g.rdzu)
    vert.fzm = b.f32("fzm", g.fzm)
    vert.fzp = b.f32("fzp", g.fzp)
    vert.cf1, vert.cf2, vert.cf3 = g.cf1, g.cf2, g.cf3
    vert.zb_cell = b.f32("zb", g.zb_cell)
    vert.zb3_cell = b.f32("zb3", g.zb3_cell)
    vert.dss = b.f32("dss", case.dss)
    vert.etp = b.f32("etp", case.etp)
    vert.etm = b.f32("etm", case.etm)
    vert.ewp = b.f32("ewp", case.ewp)
    vert.ewm = b.f32("ewm", case.ewm)
    vert.u_init = b.f32("uinit", case.u_init)
    vert.v_init = b.f32("vinit", case.v_init)
    vert.rho_base = b.f32("rb", case.reference.rho_base)
    vert.rtheta_base = b.f32("rtb", case.reference.rtheta_base)
    vert.exner_base = b.f32("exb", case.reference.exner_base)
    vert.zgrid = b.f32("zgrid", g.zgrid)

    cfg = HexdynConfig()
    c = case.config
    cfg.config_dt = c.config_dt
    cfg.config_time_integration_order = c.config_time_integration_order
    cfg.config_number_of_sub_steps = c.config_number_of_sub_steps
    cfg.config_dynamics_split_steps = c.config_dynamics_split_steps
    cfg.config_split_dynamics_transport = int(c.config_split_dynamics_transport)
    cfg.config_scalar_advection = int(c.config_scalar_advection)
    cfg.config_monotonic = int(c.config_monotonic)
    cfg.config_divergence_damping = int(c.config_divergence_damping)
    cfg.config_coef_3rd_order = c.config_coef_3rd_order
    cfg.config_apvm_upwinding = c.config_apvm_upwinding
    cfg.config_smdiv = c.config_smdiv
    cfg.config_len_disp = case.len_disp
    cfg.n_scalars = case.n_scalars
    cfg.n_moist = case.n_moist
    cfg.numerics = 0 if c.numerics == "strict" else 1
    cfg.config_horiz_mixing = 0 if c.config_horiz_mixing == "off" else 1
    cfg.config_visc4_2dsmag = c.config_visc4_2dsmag
    cfg.config_smagorinsky_coef = c.config_smagorinsky_coef
    cfg.config_del4u_div_factor = c.config_del4u_div_factor
    cfg.config_h_ScaleWithMesh = int(c.config_h_ScaleWithMesh)
    cfg.config_mpas_cam_coef = c.config_mpas_cam_coef
    cfg.config_number_cam_damping_levels = c.config_number_cam_damping_levels
    cfg.config_rayleigh_damp_u = int(c.config_rayleigh_damp_u)
    cfg.config_number_rayleigh_damp_u_levels = c.config_number_rayleigh_damp_u_levels
    cfg.config_rayleigh_damp_u_timescale_days = c.config_rayleigh_damp_u_timescale_days
    return mesh, vert, cfg, b


def state_view_from_ptrs(ptrs: Dict[str, int]) -> HexdynStateView:
    view = HexdynStateView()
    for name in STATE_FIELDS:
        setattr(view, name, ptrs.get(name))
    return view
