"""The adapter from the reference's from_host argument types (duck-typed
objects carrying the reference's attribute names) to this dycore's case and
C structs: every struct array equals the one built from the same case
directly, the oracle steps the adapted case bit for bit like the original,
and unimplemented reference knobs are refused."""

from dataclasses import fields
from types import SimpleNamespace

import numpy as np
import pytest

from hexdyn import capi
from hexdyn.config import ConfigurationRefusal, DycoreConfig
from hexdyn.reference_types import case_from_reference


def _reference_objects(case, **config_overrides):
    g, r, s = case.grid, case.reference, case.state
    mesh = SimpleNamespace(arrays=dict(case.mesh.arrays), sphere_radius=case.mesh.sphere_radius)
    state = SimpleNamespace(rho=s.rho_zz, rho_theta=s.rtheta_m, rho_u=s.ru, rho_w=s.rw, scalars=s.scalars,
                            time_seconds=0.0)
    vertical = SimpleNamespace(zw=g.zw, dzw=g.dzw, rdzw=g.rdzw, dzu=g.dzu, rdzu=g.rdzu, fzm=g.fzm, fzp=g.fzp,
                               zgrid=g.zgrid, zz=g.zz, zxu=g.zxu, dss=np.zeros_like(case.dss), cf1=g.cf1, cf2=g.cf2,
                               cf3=g.cf3, first_height_level=0)
    reference = SimpleNamespace(rho_base=r.rho_base, rho_theta_base=r.rtheta_base, exner_base=r.exner_base,
                                pressure_base=np.ones_like(r.rho_base))
    terrain = SimpleNamespace(zb_cell=g.zb_cell, zb3_cell=g.zb3_cell)
    adv = SimpleNamespace(adv_coefs=case.adv_coefs, adv_coefs_3rd=case.adv_coefs_3rd,
                          n_adv_cells_for_edge=case.n_adv, adv_cells_for_edge=case.adv_cells, horizontal_order=3)
    winds = SimpleNamespace(u_init=case.u_init, v_init=case.v_init)
    cfg = {f.name: getattr(case.config, f.name) for f in fields(DycoreConfig) if f.name.startswith("config_")}
    cfg.update(source_release="v8.4.1", config_positive_definite=False, config_les_model="none",
               config_epssm=0.0, config_scalar_adv_order=3, config_moist_physics=case.n_moist > 0)
    cfg.update(config_overrides)
    return mesh, state, vertical, reference, SimpleNamespace(**cfg), terrain, adv, winds


def test_adapter_builds_the_same_structs(get_case):
    case = get_case("x1.642-moist")
    mesh, state, vertical, reference, config, terrain, adv, winds = _reference_objects(case)
    back, st, saved = case_from_reference(mesh, state, vertical, reference, config, terrain_metrics=terrain,
                                          advection_coefficients=adv, reference_wind_profiles=winds)
    assert saved is None
    a = capi.build_structs(case)[3].keep
    b = capi.build_structs(back)[3].keep
    assert set(a) == set(b)
    for k in a:
        assert a[k].shape == b[k].shape and np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k
    for k in ("config_dt", "config_number_of_sub_steps", "config_smdiv", "config_len_disp"):
        assert getattr(capi.build_structs(case)[2], k) == getattr(capi.build_structs(back)[2], k)


def test_adapted_case_steps_bitwise_on_the_oracle(get_case):
    from hexdyn.case import initial_state_arrays
    from oracle.oracle import Oracle

    case = get_case("x1.642-moist")
    mesh, state, vertical, reference, config, terrain, adv, winds = _reference_objects(case)
    back, st, _ = case_from_reference(mesh, state, vertical, reference, config, terrain_metrics=terrain,
                                      advection_coefficients=None, reference_wind_profiles=winds)
    outs = []
    for c, s in ((case, initial_state_arrays(case)), (back, st)):
        o = Oracle(c, 32)
        o.set_state(s, rebuild=True)
        assert o.step(1) == 0
        outs.append(o.get_state())
    for k in outs[0]:
        assert np.array_equal(outs[0][k].view(np.uint32), outs[1][k].view(np.uint32)), k


@pytest.mark.parametrize("knob,value", [("config_positive_definite", True), ("config_les_model", "smagorinsky"),
                                        ("config_epssm", 0.1), ("source_release", "v8.2.3")])
def test_unimplemented_reference_knobs_are_refused(get_case, knob, value):
    case = get_case("x1.642-moist")
    mesh, state, vertical, reference, config, terrain, adv, winds = _reference_objects(case, **{knob: value})
    with pytest.raises(ConfigurationRefusal):
        case_from_reference(mesh, state, vertical, reference, config, terrain_metrics=terrain,
                            advection_coefficients=adv, reference_wind_profiles=winds)


def test_missing_reference_winds_are_refused(get_case):
    case = get_case("x1.642-moist")
    mesh, state, vertical, reference, config, terrain, adv, winds = _reference_objects(case)
    with pytest.raises(ConfigurationRefusal):
        case_from_reference(mesh, state, vertical, reference, config, terrain_metrics=terrain,
                            advection_coefficients=adv, reference_wind_profiles=None)
