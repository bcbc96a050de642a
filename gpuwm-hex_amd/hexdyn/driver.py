"""Host-side dycore driver over libhexdyn.so, shaped like the reference's
``CudaDryDycoreDriver`` (``cuda_driver.py:2658``):

* ``HexDycoreDriver.from_host(case, state=None, saved=None)`` uploads the
  mesh, vertical grid, reference state and initial state once
  (``from_host`` at ``cuda_driver.py:2826-3025``);
* ``step_device()`` advances one large step on the device and returns a
  receipt (``step_device`` at ``:5751``); a set validation flag raises
  ``FloatingPointError`` and nothing is published (``:5390-5403``);
* ``step()`` also downloads the state (``:6112-6149``).

Fields live in PyTorch-ROCm tensors in MPAS logical layout ([L][entity]);
the library keeps its own level-fastest padded layout internally.
"""

from __future__ import annotations

import ctypes as ct
import time
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

from . import capi
from .config import ConfigurationRefusal
from .lib import lib

STATE_FIELDS = capi.STATE_FIELDS


def _torch():
    import torch

    return torch


@dataclass
class StepReceipt:
    start_time_seconds: float
    end_time_seconds: float
    wall_seconds: float
    stage_acoustic_steps: tuple
    dynamics_split_steps: int


@dataclass
class HostStepResult:
    state: Dict[str, np.ndarray]
    receipt: StepReceipt


def _check(h, rc: int, what: str) -> None:
    if rc == 0:
        return
    buf = ct.create_string_buffer(512)
    if h:
        lib().hexdyn_last_error(h, buf, 512)
    msg = f"{what} failed (rc={rc}): {buf.value.decode(errors='replace')}"
    if rc == capi.HEXDYN_E_NONFINITE:
        raise FloatingPointError(msg)
    if rc == capi.HEXDYN_E_REFUSED:
        raise ConfigurationRefusal(what, rc, msg, "a supported configuration")
    raise RuntimeError(msg)


class HexDycoreDriver:
    """One dycore instance on the current HIP device."""

    def __init__(self, case, handle: int, bundle) -> None:
        self.case = case
        self._h = ct.c_void_p(handle)
        self._bundle = bundle
        self.shapes = self._state_shapes(case)
        self.tuning = None

    @staticmethod
    def _state_shapes(case) -> Dict[str, tuple]:
        L, C, E = case.n_levels, case.mesh.n_cells, case.mesh.n_edges
        return {"rho_zz": (L, C), "rtheta_m": (L, C), "ru": (L, E), "rw": (L + 1, C),
                "scalars": (case.n_scalars, L, C), "theta_m": (L, C), "exner": (L, C), "rho_p": (L, C),
                "rtheta_p": (L, C), "pressure_p": (L, C), "u": (L, E), "w": (L + 1, C)}

    @classmethod
    def from_host(cls, case, state: Optional[Dict[str, np.ndarray]] = None,
                  saved: Optional[Dict[str, np.ndarray]] = None) -> "HexDycoreDriver":
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("HexDycoreDriver needs a HIP device")
        mesh_s, vert_s, cfg_s, bundle = capi.build_structs(case)
        h = ct.c_void_p()
        L = lib()
        rc = L.hexdyn_create(ct.byref(mesh_s), ct.byref(vert_s), ct.byref(cfg_s), ct.byref(h))
        if rc != 0:
            buf = ct.create_string_buffer(512)
            L.hexdyn_last_create_error(buf, len(buf))
            why = buf.value.decode(errors="replace") or f"hexdyn_create returned {rc}"
            if rc == capi.HEXDYN_E_REFUSED:
                raise ConfigurationRefusal("hexdyn_create", rc, why, "a supported mesh and configuration")
            raise RuntimeError(f"hexdyn_create failed ({rc}): {why}")
        drv = cls(case, h.value, bundle)
        try:  # a stored launch tuning of this device and mesh (hexdyn.autotune)
            from .autotune import apply_saved

            drv.tuning = apply_saved(drv)
        except (OSError, ValueError):
            drv.tuning = None
        if state is None:
            s = case.state
            state = {"rho_zz": s.rho_zz, "rtheta_m": s.rtheta_m, "ru": s.ru, "rw": s.rw, "scalars": s.scalars}
        drv.set_state(state, saved)
        return drv

    @classmethod
    def from_reference(cls, mesh, state, vertical, reference, config, *, saved_diagnostics=None,
                       terrain_metrics, advection_coefficients=None,
                       reference_wind_profiles=None) -> "HexDycoreDriver":
        """Drop-in for the reference's CudaDryDycoreDriver.from_host
        (cuda_driver.py:2826-2842): the same arguments, read by attribute name
        (hexdyn.reference_types), uploaded through hexdyn_create."""
        from .reference_types import case_from_reference

        case, st, saved = case_from_reference(
            mesh, state, vertical, reference, config, saved_diagnostics=saved_diagnostics,
            terrain_metrics=terrain_metrics, advection_coefficients=advection_coefficients,
            reference_wind_profiles=reference_wind_profiles)
        return cls.from_host(case, state=st, saved=saved)

    def close(self) -> None:
        if self._h:
            lib().hexdyn_destroy(self._h)
            self._h = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> ct.c_void_p:
        return self._h

    @staticmethod
    def stream_ptr() -> int:
        torch = _torch()
        return torch.cuda.current_stream().cuda_stream

    def _view(self, tensors: Dict[str, object]) -> capi.HexdynStateView:
        return capi.state_view_from_ptrs({k: (t.data_ptr() if t is not None else None) for k, t in tensors.items()})

    def set_state(self, state: Dict[str, np.ndarray], saved: Optional[Dict[str, np.ndarray]] = None) -> None:
        torch = _torch()
        tens = {}
        for name in ("rho_zz", "rtheta_m", "ru", "rw", "scalars"):
            a = np.ascontiguousarray(state[name], dtype=np.float32)
            if a.shape != self.shapes[name]:
                raise ValueError(f"{name} shape {a.shape} != {self.shapes[name]}")
            if not np.isfinite(a).all():
                raise FloatingPointError(f"{name} contains non-finite values")
            tens[name] = torch.from_numpy(a).to("cuda") if a.size else None
        if np.any(np.asarray(state["rho_zz"]) <= 0.0):
            raise ValueError("rho must be strictly positive")
        rebuild = 1
        if saved is not None:
            rebuild = 0
            for name in ("theta_m", "exner", "rho_p", "rtheta_p", "pressure_p", "u", "w"):
                tens[name] = torch.from_numpy(np.ascontiguousarray(saved[name], dtype=np.float32)).to("cuda")
        view = self._view(tens)
        torch.cuda.synchronize()
        _check(self._h, lib().hexdyn_set_state(self._h, ct.byref(view), rebuild, ct.c_void_p(self.stream_ptr())),
               "hexdyn_set_state")

    def state_tensors(self) -> Dict[str, object]:
        torch = _torch()
        out = {name: torch.empty(shape, dtype=torch.float32, device="cuda") for name, shape in self.shapes.items()}
        view = self._view({k: (v if v.numel() else None) for k, v in out.items()})
        _check(self._h, lib().hexdyn_get_state(self._h, ct.byref(view), ct.c_void_p(self.stream_ptr())),
               "hexdyn_get_state")
        return out

    def get_state(self) -> Dict[str, np.ndarray]:
        return {k: v.cpu().numpy() for k, v in self.state_tensors().items()}

    def export_field(self, name: str):
        torch = _torch()
        fid = capi.FIELD_ID[name]
        rows, cols = ct.c_int64(), ct.c_int64()
        _check(self._h, lib().hexdyn_field_shape(self._h, fid, ct.byref(rows), ct.byref(cols)), "hexdyn_field_shape")
        out = torch.empty((rows.value, cols.value), dtype=torch.float32, device="cuda")
        _check(self._h, lib().hexdyn_export_field(self._h, fid, ct.c_void_p(out.data_ptr()),
                                                  ct.c_void_p(self.stream_ptr())), "hexdyn_export_field")
        return out

    def step_device(self, n_steps: int = 1) -> StepReceipt:
        t0 = lib().hexdyn_model_time(self._h)
        w0 = time.perf_counter()
        _check(self._h, lib().hexdyn_step(self._h, n_steps, ct.c_void_p(self.stream_ptr())), "hexdyn_step")
        from .setup import rk_schedule

        c = self.case.config
        sched = rk_schedule(c.config_dt, order=c.config_time_integration_order,
                            acoustic_substeps=c.config_number_of_sub_steps,
                            dynamics_splits=c.config_dynamics_split_steps)
        return StepReceipt(t0, lib().hexdyn_model_time(self._h), time.perf_counter() - w0,
                           tuple(s.acoustic_steps for s in sched), c.config_dynamics_split_steps)

    def step(self, n_steps: int = 1) -> HostStepResult:
        receipt = self.step_device(n_steps)
        return HostStepResult(self.get_state(), receipt)

    def write_restart(self, path, xtime: str) -> None:
        """MPAS-style restart of the published state (hexdyn.mpas_io.write_restart), with the
        last acoustic ru_p/rw_p words as output-only fields."""
        from .mpas_io import write_restart

        acoustic = {"ru_p": self.export_field("RU_P").cpu().numpy(), "rw_p": self.export_field("RW_P").cpu().numpy()}
        write_restart(path, self.get_state(), xtime=xtime, model_time_s=lib().hexdyn_model_time(self._h),
                      acoustic=acoustic)

    def read_restart(self, path) -> Dict[str, object]:
        """Load a restart written by write_restart; the next step continues bit for bit."""
        from .mpas_io import read_restart

        state, meta = read_restart(path)
        if "scalars" not in state:
            state["scalars"] = np.zeros(self.shapes["scalars"], dtype=np.float32)
        self.set_state(state, saved=state)
        return meta

    def set_physics_tendencies(self, rho=None, rho_u=None, rho_theta=None, scalars=None) -> None:
        """Hold physics tendencies (logical [L][entity] arrays; scalars [T][L][C],
        the transport's scalar_tendency) through the following steps; all None
        clears them (hexdyn_set_physics_tendencies)."""
        torch = _torch()
        if rho is None and rho_u is None and rho_theta is None and scalars is None:
            _check(self._h, lib().hexdyn_set_physics_tendencies(self._h, None, ct.c_void_p(self.stream_ptr())),
                   "hexdyn_set_physics_tendencies")
            return
        keep = {}
        view = capi.HexdynPhysicsView()
        for name, a, shape in (("rho", rho, self.shapes["rho_zz"]), ("rho_u", rho_u, self.shapes["ru"]),
                               ("rho_theta", rho_theta, self.shapes["rtheta_m"]),
                               ("scalars", scalars, self.shapes["scalars"])):
            if a is None:
                continue
            arr = np.ascontiguousarray(a, dtype=np.float32)
            if arr.shape != shape:
                raise ValueError(f"{name} shape {arr.shape} != {shape}")
            keep[name] = torch.from_numpy(arr).to("cuda")
            setattr(view, name, keep[name].data_ptr())
        torch.cuda.synchronize()
        _check(self._h, lib().hexdyn_set_physics_tendencies(self._h, ct.byref(view), ct.c_void_p(self.stream_ptr())),
               "hexdyn_set_physics_tendencies")

    def set_post_transport_clamp(self, enable: bool) -> None:
        """False hands the transported scalars to the caller unclamped (the reference's
        physics_post_transport_callback path, cuda_driver.py:5299-5311)."""
        _check(self._h, lib().hexdyn_set_post_transport_clamp(self._h, 1 if enable else 0),
               "hexdyn_set_post_transport_clamp")

    def graph_count(self) -> int:
        """Captured step graphs held (one per buffer layout)."""
        return int(lib().hexdyn_graph_count(self._h))

    def use_graphs(self, enable: bool) -> None:
        """Step through captured hipGraphs (default) or kernel by kernel."""
        _check(self._h, lib().hexdyn_set_graphs(self._h, 1 if enable else 0), "hexdyn_set_graphs")

    def model_time(self) -> float:
        """Model time of the published state in seconds (hexdyn_model_time)."""
        return float(lib().hexdyn_model_time(self._h))

    def integrals(self) -> Dict[str, float]:
        """FP64 global integrals of the published state (hexdyn_integrals)."""
        nq = 3 + self.case.n_scalars
        out = (ct.c_double * nq)()
        _check(self._h, lib().hexdyn_integrals(self._h, out, nq, ct.c_void_p(self.stream_ptr())),
               "hexdyn_integrals")
        res = {"dry_mass": out[0], "rho_theta": out[1], "total_energy": out[2]}
        for s in range(self.case.n_scalars):
            res[f"tracer_mass_{s}"] = out[3 + s]
        return res

    # phase entry points (parity tests walk a step phase by phase)
    def phase(self, name: str, *args: int) -> None:
        fn = getattr(lib(), f"hexdyn_{name}")
        _check(self._h, fn(self._h, *[ct.c_int(a) for a in args], ct.c_void_p(self.stream_ptr())), f"hexdyn_{name}")

    def profile(self, enable: bool) -> None:
        lib().hexdyn_profile_enable(self._h, 1 if enable else 0)

    def profile_read(self, kernel: str):
        avg = ct.c_double()
        n = ct.c_int64()
        lib().hexdyn_profile_read(self._h, kernel.encode(), ct.byref(avg), ct.byref(n))
        return avg.value, n.value

    def profile_kernels(self):
        buf = ct.create_string_buffer(8192)
        lib().hexdyn_profile_kernels(self._h, buf, 8192)
        names = buf.value.decode()
        return [n for n in names.split(",") if n]
