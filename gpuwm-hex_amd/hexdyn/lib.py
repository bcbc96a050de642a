"""Loader for libhexdyn.so (the HIP dycore).  There is no CPU fallback: if the
library is missing or fails to load, every entry point raises."""

from __future__ import annotations

import ctypes as ct
import subprocess
from pathlib import Path

from . import capi

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libhexdyn.so"
CSRC = HERE.parent / "csrc"

ABI_VERSION = 4  # HEXDYN_ABI_VERSION in include/hexdyn.h

_lib = None


class HexdynLibraryMissing(RuntimeError):
    pass


def build(force: bool = False) -> Path:
    """Compile libhexdyn.so for gfx950 with hipcc (make in csrc/)."""
    args = ["make", "-C", str(CSRC), "-j4"]
    if force:
        subprocess.run(["make", "-C", str(CSRC), "clean"], check=True, capture_output=True)
    subprocess.run(args, check=True)
    return LIB_PATH


def lib() -> ct.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise HexdynLibraryMissing(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C {CSRC}")
    L = ct.CDLL(str(LIB_PATH))
    vp = ct.c_void_p
    L.hexdyn_abi_version.restype = ct.c_int
    L.hexdyn_create.argtypes = [ct.POINTER(capi.HexdynMesh), ct.POINTER(capi.HexdynVertical),
                                ct.POINTER(capi.HexdynConfig), ct.POINTER(vp)]
    L.hexdyn_destroy.argtypes = [vp]
    L.hexdyn_destroy.restype = None
    L.hexdyn_last_error.argtypes = [vp, ct.c_char_p, ct.c_size_t]
    L.hexdyn_last_create_error.argtypes = [ct.c_char_p, ct.c_size_t]
    L.hexdyn_graph_count.argtypes = [vp]
    L.hexdyn_set_state.argtypes = [vp, ct.POINTER(capi.HexdynStateView), ct.c_int, vp]
    L.hexdyn_get_state.argtypes = [vp, ct.POINTER(capi.HexdynStateView), vp]
    L.hexdyn_export_field.argtypes = [vp, ct.c_int, vp, vp]
    L.hexdyn_import_field.argtypes = [vp, ct.c_int, vp, vp]
    L.hexdyn_field_shape.argtypes = [vp, ct.c_int, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64)]
    L.hexdyn_step.argtypes = [vp, ct.c_int, vp]
    L.hexdyn_model_time.argtypes = [vp]
    L.hexdyn_model_time.restype = ct.c_double
    for name in ("hexdyn_begin_step", "hexdyn_end_step"):
        getattr(L, name).argtypes = [vp, vp]
    for name in ("hexdyn_begin_subcycle", "hexdyn_dyn_tend", "hexdyn_divergence_damping", "hexdyn_recover",
                 "hexdyn_solve_diagnostics", "hexdyn_end_subcycle", "hexdyn_scalars"):
        getattr(L, name).argtypes = [vp, ct.c_int, vp]
    L.hexdyn_acoustic_substep.argtypes = [vp, ct.c_int, ct.c_int, vp]
    L.hexdyn_profile_enable.argtypes = [vp, ct.c_int]
    L.hexdyn_profile_read.argtypes = [vp, ct.c_char_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_int64)]
    L.hexdyn_profile_kernels.argtypes = [vp, ct.c_char_p, ct.c_size_t]
    L.hexdyn_powf.argtypes = [ct.c_float, ct.c_float]
    L.hexdyn_powf.restype = ct.c_float
    L.hexdyn_integrals.argtypes = [vp, ct.POINTER(ct.c_double), ct.c_int, vp]
    L.hexdyn_integrals_noalloc.argtypes = [vp, vp, ct.c_int, vp]
    L.hexdyn_copy_state.argtypes = [vp, vp, vp]
    L.hexdyn_set_launch.argtypes = [vp, ct.c_char_p, ct.c_int]
    L.hexdyn_get_launch.argtypes = [vp, ct.c_char_p]
    # learned corrector
    f32p = ct.POINTER(ct.c_float)
    L.hexdyn_net_create.argtypes = [vp, ct.c_int, ct.POINTER(vp)]
    L.hexdyn_net_destroy.argtypes = [vp]
    L.hexdyn_net_destroy.restype = None
    L.hexdyn_net_param_count.argtypes = [vp]
    L.hexdyn_net_param_count.restype = ct.c_int64
    for name in ("hexdyn_net_set_params", "hexdyn_net_get_params", "hexdyn_net_get_grads"):
        getattr(L, name).argtypes = [vp, f32p, ct.c_int64, vp]
    L.hexdyn_net_set_scales.argtypes = [vp, f32p, f32p, f32p, vp]
    L.hexdyn_net_forward.argtypes = [vp, vp, ct.c_int, vp]
    L.hexdyn_net_set_target.argtypes = [vp, vp, vp, vp]
    L.hexdyn_net_loss.argtypes = [vp, ct.POINTER(ct.c_double), vp]
    L.hexdyn_net_backward.argtypes = [vp, vp]
    L.hexdyn_net_adam.argtypes = [vp, ct.c_float, ct.c_float, ct.c_float, ct.c_float, ct.c_int, vp]
    L.hexdyn_net_apply.argtypes = [vp, vp, ct.c_float, vp]
    L.hexdyn_net_checks.argtypes = [vp, vp, f32p, ct.POINTER(ct.c_double), vp]
    L.hexdyn_net_features.argtypes = [vp, vp, vp, vp]
    L.hexdyn_net_export.argtypes = [vp, ct.c_int, vp, vp]
    L.hexdyn_net_last_error.argtypes = [vp, ct.c_char_p, ct.c_size_t]
    L.hexdyn_set_graphs.argtypes = [vp, ct.c_int]
    L.hexdyn_set_physics_tendencies.argtypes = [vp, ct.POINTER(capi.HexdynPhysicsView), vp]
    L.hexdyn_set_post_transport_clamp.argtypes = [vp, ct.c_int]
    if L.hexdyn_abi_version() != ABI_VERSION:
        raise HexdynLibraryMissing("libhexdyn.so ABI version mismatch")
    _lib = L
    return L


EXPORTED = [
    "hexdyn_abi_version", "hexdyn_create", "hexdyn_destroy", "hexdyn_last_error", "hexdyn_set_state",
    "hexdyn_get_state", "hexdyn_export_field", "hexdyn_import_field", "hexdyn_field_shape", "hexdyn_step",
    "hexdyn_model_time", "hexdyn_begin_step", "hexdyn_begin_subcycle", "hexdyn_dyn_tend",
    "hexdyn_acoustic_substep", "hexdyn_divergence_damping", "hexdyn_recover", "hexdyn_solve_diagnostics",
    "hexdyn_end_subcycle", "hexdyn_scalars", "hexdyn_end_step", "hexdyn_profile_enable", "hexdyn_profile_read",
    "hexdyn_profile_kernels", "hexdyn_powf", "hexdyn_integrals", "hexdyn_set_graphs",
    "hexdyn_set_physics_tendencies", "hexdyn_set_post_transport_clamp", "hexdyn_last_create_error",
    "hexdyn_graph_count", "hexdyn_integrals_noalloc", "hexdyn_copy_state", "hexdyn_set_launch",
    "hexdyn_get_launch",
    "hexdyn_net_create", "hexdyn_net_destroy", "hexdyn_net_param_count", "hexdyn_net_set_params",
    "hexdyn_net_get_params", "hexdyn_net_get_grads", "hexdyn_net_set_scales", "hexdyn_net_forward",
    "hexdyn_net_set_target", "hexdyn_net_loss", "hexdyn_net_backward", "hexdyn_net_adam", "hexdyn_net_apply",
    "hexdyn_net_checks", "hexdyn_net_features", "hexdyn_net_export", "hexdyn_net_last_error",
]
