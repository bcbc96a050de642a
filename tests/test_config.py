import pytest

from hexdyn.config import ConfigurationRefusal, DycoreConfig


def test_defaults_valid():
    DycoreConfig().validate()


@pytest.mark.parametrize("raw", [
    {"config_unknown_knob": 1},
    {"config_dt": -5.0},
    {"config_time_integration_order": 4},
    {"config_dynamics_split_steps": 2},
    {"config_dynamics_split_steps": 3, "config_split_dynamics_transport": False},
    {"config_apvm_upwinding": -0.1},
    {"config_horiz_mixing": "2d_fixed"},
    {"config_horiz_mixing": "2d_smagorinsky", "config_mpas_cam_coef": -0.1},
    {"config_horiz_mixing": "2d_smagorinsky", "config_mpas_cam_coef": 0.1, "config_number_cam_damping_levels": 0},
    {"config_rayleigh_damp_u": True, "config_number_rayleigh_damp_u_levels": 0},
    {"config_rayleigh_damp_u": True, "config_rayleigh_damp_u_timescale_days": 0.0},
    {"config_horiz_mixing": "2d_smagorinsky", "config_visc4_2dsmag": -1.0},
    {"config_horiz_mixing": "2d_smagorinsky", "config_del4u_div_factor": 0.0},
    {"numerics": "loose"},
])
def test_refusals(raw):
    with pytest.raises(ConfigurationRefusal):
        DycoreConfig.from_mapping(raw)


def test_upper_damping_knobs_admitted():
    """The MPAS upper-damping namelist knobs are implemented: the CAM
    viscosity floor and Rayleigh damping of u (mixing_v841.py:746-750,
    cuda_driver.py:4427-4433)."""
    cfg = DycoreConfig.from_mapping({"config_horiz_mixing": "2d_smagorinsky", "config_mpas_cam_coef": 0.1,
                                     "config_number_cam_damping_levels": 4, "config_rayleigh_damp_u": True,
                                     "config_number_rayleigh_damp_u_levels": 6,
                                     "config_rayleigh_damp_u_timescale_days": 5.0})
    assert cfg.config_rayleigh_damp_u and cfg.config_mpas_cam_coef == 0.1
