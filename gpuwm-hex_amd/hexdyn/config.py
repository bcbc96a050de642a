"""Dycore configuration with MPAS namelist names and refusal of unknown knobs.

Mirrors ``driver.py:172-230`` (``DryDycoreConfig``) and
``config_v841.py:114-140`` (``V841DryDycoreConfig``): a frozen dataclass whose
``from_mapping`` refuses any knob it does not implement, and whose
``validate`` refuses values outside the implemented branches with
``ConfigurationRefusal`` (``errors.py``).
"""

from __future__ import annotations

from dataclasses import dataclass, fields
from typing import Any, Mapping

import math


class ConfigurationRefusal(ValueError):
    """Raised at admission for a knob value this dycore does not implement."""

    def __init__(self, knob: str, value: object, reason: str, remedy: str) -> None:
        super().__init__(f"{knob}={value!r} refused: {reason}; use {remedy}")
        self.knob = knob
        self.value = value
        self.reason = reason
        self.remedy = remedy


@dataclass(frozen=True)
class DycoreConfig:
    config_dt: float = 720.0
    config_time_integration_order: int = 3
    config_number_of_sub_steps: int = 2
    config_dynamics_split_steps: int = 3
    config_split_dynamics_transport: bool = True
    config_scalar_advection: bool = True
    config_monotonic: bool = True
    config_coef_3rd_order: float = 0.25
    config_apvm_upwinding: float = 0.5
    config_divergence_damping: bool = True
    config_smdiv: float = 0.1
    config_len_disp: float = 0.0          # 0 resolves to the mesh's nominal minimum dc
    config_xnutr: float = 0.2
    config_zd: float = 22000.0
    config_epssm_minimum: float = 0.1
    config_epssm_maximum: float = 0.5
    config_epssm_transition_bottom_z: float = 30000.0
    config_epssm_transition_top_z: float = 50000.0
    config_horiz_mixing: str = "off"      # or "2d_smagorinsky" (mixing_v841.py:71-85)
    config_visc4_2dsmag: float = 0.05
    config_smagorinsky_coef: float = 0.125
    config_del4u_div_factor: float = 10.0
    config_h_ScaleWithMesh: bool = True
    config_mpas_cam_coef: float = 0.0
    config_number_cam_damping_levels: int = 4
    config_rayleigh_damp_u: bool = False
    config_number_rayleigh_damp_u_levels: int = 6
    config_rayleigh_damp_u_timescale_days: float = 5.0
    config_moist_physics: bool = False
    numerics: str = "strict"

    @classmethod
    def from_mapping(cls, raw: Mapping[str, Any]) -> "DycoreConfig":
        known = {f.name for f in fields(cls)}
        unknown = sorted(set(raw) - known)
        if unknown:
            knob = unknown[0]
            raise ConfigurationRefusal(knob, raw[knob], "this option is not implemented by the hexdyn step",
                                       f"one of {sorted(known)}")
        cfg = cls(**raw)
        cfg.validate()
        return cfg

    def validate(self) -> None:
        if not math.isfinite(self.config_dt) or self.config_dt <= 0.0:
            raise ConfigurationRefusal("config_dt", self.config_dt, "the timestep must be finite and positive",
                                       "config_dt>0")
        if self.config_time_integration_order not in (2, 3):
            raise ConfigurationRefusal("config_time_integration_order", self.config_time_integration_order,
                                       "only the order-2 and order-3 schedules exist", "2 or 3")
        if self.config_number_of_sub_steps < 1:
            raise ConfigurationRefusal("config_number_of_sub_steps", self.config_number_of_sub_steps,
                                       "the acoustic loop needs a positive count", "a positive integer")
        if self.config_dynamics_split_steps not in (1, 3):
            raise ConfigurationRefusal("config_dynamics_split_steps", self.config_dynamics_split_steps,
                                       "one or three dynamics subcycles are implemented", "1 or 3")
        if self.config_dynamics_split_steps == 3 and not self.config_split_dynamics_transport:
            raise ConfigurationRefusal("config_dynamics_split_steps", 3,
                                       "dynamics subcycles are only active with split transport",
                                       "config_split_dynamics_transport=True")
        if not self.config_split_dynamics_transport and self.config_scalar_advection:
            raise ConfigurationRefusal("config_split_dynamics_transport", False,
                                       "scalar transport runs only in the split (after-dynamics) form",
                                       "config_split_dynamics_transport=True")
        if self.config_apvm_upwinding < 0.0:
            raise ConfigurationRefusal("config_apvm_upwinding", self.config_apvm_upwinding,
                                       "APVM is active only for non-negative coefficients", ">= 0")
        if self.config_smdiv < 0.0:
            raise ConfigurationRefusal("config_smdiv", self.config_smdiv, "damping must be non-negative", ">= 0")
        if not 0.0 <= self.config_xnutr <= 1.0:
            raise ConfigurationRefusal("config_xnutr", self.config_xnutr, "bounded damping coefficient",
                                       "0 <= config_xnutr <= 1")
        if self.config_horiz_mixing not in ("off", "2d_smagorinsky"):
            raise ConfigurationRefusal("config_horiz_mixing", self.config_horiz_mixing,
                                       "only the v8.4.1 2-D Smagorinsky branch is implemented",
                                       "'off' or '2d_smagorinsky'")
        if self.config_horiz_mixing == "2d_smagorinsky":
            for knob in ("config_visc4_2dsmag", "config_smagorinsky_coef"):
                value = getattr(self, knob)
                if not math.isfinite(value) or value < 0.0:
                    raise ConfigurationRefusal(knob, value, "the Registry admits only non-negative values",
                                               f"{knob}>=0")
            if not math.isfinite(self.config_del4u_div_factor) or self.config_del4u_div_factor <= 0.0:
                raise ConfigurationRefusal("config_del4u_div_factor", self.config_del4u_div_factor,
                                           "the Registry requires a positive divergent hyperdiffusion factor",
                                           "config_del4u_div_factor>0")
            if not math.isfinite(self.config_mpas_cam_coef) or self.config_mpas_cam_coef < 0.0:
                raise ConfigurationRefusal("config_mpas_cam_coef", self.config_mpas_cam_coef,
                                           "the CAM viscosity floor needs a non-negative coefficient",
                                           "config_mpas_cam_coef>=0")
            if self.config_mpas_cam_coef > 0.0 and self.config_number_cam_damping_levels < 1:
                raise ConfigurationRefusal("config_number_cam_damping_levels",
                                           self.config_number_cam_damping_levels,
                                           "the CAM floor needs at least one level", "1..nVertLevels")
        if self.config_rayleigh_damp_u:
            if self.config_number_rayleigh_damp_u_levels < 1:
                raise ConfigurationRefusal("config_number_rayleigh_damp_u_levels",
                                           self.config_number_rayleigh_damp_u_levels,
                                           "the damping layer must hold at least one level", "1..nVertLevels")
            days = self.config_rayleigh_damp_u_timescale_days
            if not math.isfinite(days) or days <= 0.0:
                raise ConfigurationRefusal("config_rayleigh_damp_u_timescale_days", days,
                                           "enabled Rayleigh damping requires a finite positive timescale",
                                           "config_rayleigh_damp_u_timescale_days>0")
        if self.numerics not in ("strict", "fast"):
            raise ConfigurationRefusal("numerics", self.numerics, "two numerics modes exist", "'strict' or 'fast'")
