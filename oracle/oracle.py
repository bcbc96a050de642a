"""TEST INFRASTRUCTURE — Python binding of the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's CPU leg (the CPU
baseline timing and the FP64 error of one GPU step) use this module, as the
checker.  The product (libhexdyn.so and hexdyn/) never imports
it.
"""

from __future__ import annotations

import ctypes as ct
import os
import subprocess
import sys
from pathlib import Path
from typing import Dict

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
sys.path.insert(0, str(ROOT / "gpuwm-hex_amd"))

from hexdyn import capi  # noqa: E402

# planted mistakes (the PLANT_* enum of hexdyn_oracle.cpp)
PLANT_CORIOLIS_TARGET_ANGLE = 1
PLANT_CORIOLIS_NEIGHBOUR_F = 2
PLANT_BLEND_SIGN = 3
PLANT_ADV_SLOT = 4
PLANT_CORRECTION_SIGN = 5
PLANT_VERTICAL_CORRECTION = 6
PLANT_W_INTERP = 7
PLANT_OMEGA_SIGN = 8

LIB = HERE / "liboracle.so"


def build(force: bool = False) -> Path:
    if force or not LIB.exists():
        subprocess.run(["make", "-C", str(HERE), "-s"], check=True)
    return LIB


_lib = None


def lib() -> ct.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ct.CDLL(str(LIB))
        _lib.oracle_create.restype = ct.c_void_p
        _lib.oracle_create.argtypes = [ct.POINTER(capi.HexdynMesh), ct.POINTER(capi.HexdynVertical),
                                       ct.POINTER(capi.HexdynConfig), ct.c_int]
        _lib.oracle_destroy.argtypes = [ct.c_void_p]
        for name in ("oracle_set_state",):
            getattr(_lib, name).argtypes = [ct.c_void_p, ct.POINTER(capi.HexdynStateView), ct.c_int]
        _lib.oracle_get_state.argtypes = [ct.c_void_p, ct.POINTER(capi.HexdynStateView)]
        _lib.oracle_export_field.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_int64]
        _lib.oracle_import_field.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_int64]
        _lib.oracle_field_size.argtypes = [ct.c_void_p, ct.c_int]
        _lib.oracle_field_size.restype = ct.c_int64
        _lib.oracle_step.argtypes = [ct.c_void_p, ct.c_int]
        for name in ("oracle_begin_step", "oracle_end_step"):
            getattr(_lib, name).argtypes = [ct.c_void_p]
        for name in ("oracle_begin_subcycle", "oracle_dyn_tend", "oracle_divergence_damping", "oracle_recover",
                     "oracle_solve_diagnostics", "oracle_end_subcycle", "oracle_scalars"):
            getattr(_lib, name).argtypes = [ct.c_void_p, ct.c_int]
        _lib.oracle_acoustic_substep.argtypes = [ct.c_void_p, ct.c_int, ct.c_int]
        _lib.oracle_powf.argtypes = [ct.c_float, ct.c_float]
        _lib.oracle_powf.restype = ct.c_float
        _lib.oracle_set_physics.argtypes = [ct.c_void_p] * 5
        _lib.oracle_set_post_transport_clamp.argtypes = [ct.c_void_p, ct.c_int]
        _lib.oracle_model_time.argtypes = [ct.c_void_p]
        _lib.oracle_model_time.restype = ct.c_double
        _lib.oracle_schedule.argtypes = [ct.c_double, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p,
                                         ct.c_void_p]
        _lib.oracle_export_btri_ctri_f64.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int64]
        _lib.oracle_export_field_f64.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_int64]
        _lib.oracle_internal_size.argtypes = [ct.c_void_p, ct.c_char_p]
        _lib.oracle_internal_size.restype = ct.c_int64
        _lib.oracle_export_internal.argtypes = [ct.c_void_p, ct.c_char_p, ct.c_void_p, ct.c_int64]
        _lib.oracle_import_internal.argtypes = [ct.c_void_p, ct.c_char_p, ct.c_void_p, ct.c_int64]
        _lib.oracle_vertical_transport_u.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64]
        _lib.oracle_export_internal_f64.argtypes = [ct.c_void_p, ct.c_char_p, ct.c_void_p, ct.c_int64]
        _lib.oracle_set_plant.argtypes = [ct.c_void_p, ct.c_int]
        for name in ("oracle_begin_subcycle_pools", "oracle_horizontal_mixing"):
            getattr(_lib, name).argtypes = [ct.c_void_p]
        _lib.oracle_momentum_terms.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64]
        _lib.oracle_momentum_terms_f64.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64]
        _lib.oracle_omega_conversion.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        for name in ("oracle_libm_powf_many", "oracle_hexdyn_powf_many"):
            getattr(_lib, name).argtypes = [ct.c_void_p, ct.c_float, ct.c_void_p, ct.c_int64]
            getattr(_lib, name).restype = None
    return _lib


def state_shapes(case) -> Dict[str, tuple]:
    L, C, E = case.n_levels, case.mesh.n_cells, case.mesh.n_edges
    cell = (L, C)
    return {"rho_zz": cell, "rtheta_m": cell, "ru": (L, E), "rw": (L + 1, C),
            "scalars": (case.n_scalars, L, C), "theta_m": cell, "exner": cell, "rho_p": cell,
            "rtheta_p": cell, "pressure_p": cell, "u": (L, E), "w": (L + 1, C)}


class Oracle:
    """CPU restatement of the dycore step; precision 32 (strict) or 64."""

    def __init__(self, case, precision: int = 32) -> None:
        self.case = case
        self.mesh_s, self.vert_s, self.cfg_s, self._bundle = capi.build_structs(case)
        self.h = lib().oracle_create(ct.byref(self.mesh_s), ct.byref(self.vert_s), ct.byref(self.cfg_s), precision)
        if not self.h:
            raise ValueError("oracle refused the case (missing mixing inputs)")
        self.shapes = state_shapes(case)

    def __del__(self) -> None:
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def set_state(self, arrays: Dict[str, np.ndarray], rebuild: bool = True) -> None:
        keep = {}
        view = capi.HexdynStateView()
        for name in capi.STATE_FIELDS:
            if name in arrays and arrays[name] is not None:
                a = np.ascontiguousarray(arrays[name], dtype=np.float32)
                keep[name] = a
                setattr(view, name, a.ctypes.data)
        lib().oracle_set_state(self.h, ct.byref(view), 1 if rebuild else 0)

    def get_state(self) -> Dict[str, np.ndarray]:
        out = {name: np.empty(shape, dtype=np.float32) for name, shape in self.shapes.items()}
        view = capi.HexdynStateView()
        for name, a in out.items():
            setattr(view, name, a.ctypes.data)
        lib().oracle_get_state(self.h, ct.byref(view))
        return out

    def export_field(self, name: str) -> np.ndarray:
        fid = capi.FIELD_ID[name]
        n = lib().oracle_field_size(self.h, fid)
        out = np.empty(n, dtype=np.float32)
        rc = lib().oracle_export_field(self.h, fid, out.ctypes.data, n)
        if rc:
            raise KeyError(name)
        return out

    def import_field(self, name: str, values: np.ndarray) -> None:
        fid = capi.FIELD_ID[name]
        a = np.ascontiguousarray(values, dtype=np.float32).ravel()
        rc = lib().oracle_import_field(self.h, fid, a.ctypes.data, a.size)
        if rc:
            raise KeyError(name)

    def export_field_f64(self, name: str) -> np.ndarray:
        fid = capi.FIELD_ID[name]
        n = lib().oracle_field_size(self.h, fid)
        out = np.empty(n, dtype=np.float64)
        if lib().oracle_export_field_f64(self.h, fid, out.ctypes.data, n):
            raise KeyError(name)
        return out

    def export_internal(self, name: str) -> np.ndarray:
        """A working array by its member name (s_theta, s_exner, c_rho, s_rw,
        c_rw, w, cofwr, rtheta_pp_old, ...), flat, logical [level][entity]."""
        n = lib().oracle_internal_size(self.h, name.encode())
        if n < 0:
            raise KeyError(name)
        out = np.empty(n, dtype=np.float32)
        if lib().oracle_export_internal(self.h, name.encode(), out.ctypes.data, n):
            raise KeyError(name)
        return out

    def import_internal(self, name: str, values: np.ndarray) -> None:
        a = np.ascontiguousarray(values, dtype=np.float32).ravel()
        if lib().oracle_import_internal(self.h, name.encode(), a.ctypes.data, a.size):
            raise KeyError(name)

    def export_internal_f64(self, name: str) -> np.ndarray:
        n = lib().oracle_internal_size(self.h, name.encode())
        if n < 0:
            raise KeyError(name)
        out = np.empty(n, dtype=np.float64)
        if lib().oracle_export_internal_f64(self.h, name.encode(), out.ctypes.data, n):
            raise KeyError(name)
        return out

    def set_plant(self, code: int) -> None:
        """Plant a deliberate mistake in one operator (PLANT_* in the C++
        source; 0 removes it).  Test hook: shows that a check catches it."""
        lib().oracle_set_plant(self.h, int(code))

    def begin_subcycle_pools(self):
        """begin_subcycle up to, not including, the horizontal mixing."""
        return lib().oracle_begin_subcycle_pools(self.h)

    def horizontal_mixing(self):
        """The mixing half of begin_subcycle (after begin_subcycle_pools)."""
        return lib().oracle_horizontal_mixing(self.h)

    def momentum_terms(self) -> np.ndarray:
        """rho_edge*(q - dKE/dc) - (u*0.5)*(hdiv0+hdiv1) of every edge and
        level from the current u, rho_zz, pv_edge, KE and ru, flat [L][E]
        (the per-element code dyn_tend runs, without its initial term)."""
        n = self.case.n_levels * self.case.mesh.n_edges
        out = np.empty(n, dtype=np.float32)
        if lib().oracle_momentum_terms(self.h, out.ctypes.data, n):
            raise ValueError("size mismatch")
        return out

    def momentum_terms_f64(self) -> np.ndarray:
        """momentum_terms in the oracle's own precision, as float64."""
        n = self.case.n_levels * self.case.mesh.n_edges
        out = np.empty(n, dtype=np.float64)
        if lib().oracle_momentum_terms_f64(self.h, out.ctypes.data, n):
            raise ValueError("size mismatch")
        return out

    def omega_conversion(self, w_tend: np.ndarray, u_tend: np.ndarray) -> np.ndarray:
        """The w -> omega conversion dyn_tend applies, on a given w tendency
        [L+1][C] and momentum tendency [L][E]."""
        L, C, E = self.case.n_levels, self.case.mesh.n_cells, self.case.mesh.n_edges
        w = np.ascontiguousarray(w_tend, dtype=np.float32).reshape(L + 1, C)
        u = np.ascontiguousarray(u_tend, dtype=np.float32).reshape(L, E)
        out = np.empty((L + 1, C), dtype=np.float32)
        lib().oracle_omega_conversion(self.h, w.ctypes.data, u.ctypes.data, out.ctypes.data)
        return out

    def vertical_transport_u(self) -> np.ndarray:
        """The vertical transport term of the momentum tendency for the
        current u and rw, flat [L][E] (the code dyn_tend runs)."""
        n = self.case.n_levels * self.case.mesh.n_edges
        out = np.empty(n, dtype=np.float32)
        if lib().oracle_vertical_transport_u(self.h, out.ctypes.data, n):
            raise ValueError("size mismatch")
        return out

    def btri_ctri_f64(self):
        n = self.case.n_levels * self.case.mesh.n_cells
        b = np.empty(n, dtype=np.float64)
        c = np.empty(n, dtype=np.float64)
        if lib().oracle_export_btri_ctri_f64(self.h, b.ctypes.data, c.ctypes.data, n):
            raise RuntimeError("btri/ctri size mismatch")
        return b, c

    def step(self, n: int = 1) -> int:
        return lib().oracle_step(self.h, n)

    def set_physics(self, rho=None, rho_u=None, rho_theta=None, scalars=None) -> None:
        """Held physics tendencies (logical [L][entity]; scalars [T][L][C]);
        all None clears them."""
        self._phys = [None if a is None else np.ascontiguousarray(a, dtype=np.float32)
                      for a in (rho, rho_u, rho_theta, scalars)]
        lib().oracle_set_physics(self.h, *[None if a is None else a.ctypes.data for a in self._phys])

    def set_post_transport_clamp(self, enable: bool) -> None:
        """False hands the transported scalars over unclamped (no q < 0 -> 0 at the step's end)."""
        lib().oracle_set_post_transport_clamp(self.h, 1 if enable else 0)

    def model_time(self) -> float:
        return lib().oracle_model_time(self.h)

    # phases (same names and order as the product's C ABI)
    def begin_step(self): return lib().oracle_begin_step(self.h)
    def begin_subcycle(self, sub): return lib().oracle_begin_subcycle(self.h, sub)
    def dyn_tend(self, stage): return lib().oracle_dyn_tend(self.h, stage)
    def acoustic_substep(self, stage, ss): return lib().oracle_acoustic_substep(self.h, stage, ss)
    def divergence_damping(self, stage): return lib().oracle_divergence_damping(self.h, stage)
    def recover(self, stage): return lib().oracle_recover(self.h, stage)
    def solve_diagnostics(self, rk): return lib().oracle_solve_diagnostics(self.h, rk)
    def end_subcycle(self, sub): return lib().oracle_end_subcycle(self.h, sub)
    def scalars(self, stage): return lib().oracle_scalars(self.h, stage)
    def end_step(self): return lib().oracle_end_step(self.h)


def powf(x: float, y: float) -> float:
    return float(lib().oracle_powf(x, y))


def schedule_f32(dt: float, order: int, substeps: int, splits: int):
    large = np.zeros(3, dtype=np.float32)
    acoustic = np.zeros(3, dtype=np.float32)
    steps = np.zeros(3, dtype=np.int32)
    lib().oracle_schedule(dt, order, substeps, splits, large.ctypes.data, acoustic.ctypes.data, steps.ctypes.data)
    return large, acoustic, steps


def powf_many(x: np.ndarray, y: float, method: str) -> np.ndarray:
    """powf(x, y) elementwise in binary32 by the host C library ("libm") or
    by the product's exner method ("hexdyn")."""
    xs = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(xs)
    fn = {"libm": lib().oracle_libm_powf_many, "hexdyn": lib().oracle_hexdyn_powf_many}[method]
    fn(xs.ctypes.data, float(np.float32(y)), out.ctypes.data, xs.size)
    return out
