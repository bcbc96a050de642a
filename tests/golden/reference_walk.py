"""The oracle walk at whose checkpoints the reference vectors are taken.

Both `make_reference_vectors.py` (which feeds the operands found at each
checkpoint to the reference's own operator) and `tests/test_reference_vectors.py`
(which runs the oracle's phase from the same checkpoint) walk one large step
of the seeded x1.162 moist case at 10 levels, after one developed step so
that w, rw and every perturbation are non-zero.  The schedule is RK3 with
acoustic substep counts (1, 1, 2) and 3 dynamics subcycles; the walk covers
the first subcycle.
"""
from __future__ import annotations

import hashlib

import numpy as np

CASE, NLEV = "x1.162-moist", 10
F32 = np.float32


def build_case():
    from hexdyn.case import build_case as build

    return build(CASE, nlev=NLEV, use_cache=False)


def schedule(case):
    from hexdyn.setup import rk_schedule

    c = case.config
    return rk_schedule(c.config_dt, order=c.config_time_integration_order,
                       acoustic_substeps=c.config_number_of_sub_steps, dynamics_splits=c.config_dynamics_split_steps)


def walk(case):
    """Yield (checkpoint, oracle) along the first subcycle of a large step.
    A checkpoint named X is followed by X_out after the phase it names ran."""
    from hexdyn.case import initial_state_arrays
    from oracle.oracle import Oracle

    sched = schedule(case)
    assert [s.acoustic_steps for s in sched] == [1, 1, 2]
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(1) == 0
    o.begin_step()
    o.begin_subcycle(1)
    yield "vcoef_1", o
    o.dyn_tend(1)
    o.acoustic_substep(1, 1)
    o.divergence_damping(1)
    assert o.recover(1) == 0
    o.solve_diagnostics(1)
    o.dyn_tend(2)
    yield "vcoef_2", o
    yield "acoustic_2_1", o
    o.acoustic_substep(2, 1)
    yield "acoustic_2_1_out", o
    yield "damping_2", o
    o.divergence_damping(2)
    yield "damping_2_out", o
    yield "recover_2", o
    assert o.recover(2) == 0
    yield "recover_2_out", o
    o.solve_diagnostics(2)
    o.dyn_tend(3)
    yield "acoustic_3_1", o
    o.acoustic_substep(3, 1)
    yield "acoustic_3_1_out", o
    o.divergence_damping(3)
    yield "acoustic_3_2", o
    o.acoustic_substep(3, 2)
    yield "acoustic_3_2_out", o
    yield "damping_3", o
    o.divergence_damping(3)
    yield "damping_3_out", o
    yield "recover_3", o
    assert o.recover(3) == 0
    yield "recover_3_out", o


def scalar_tendency(case) -> np.ndarray:
    """A seeded held scalar tendency [T][L][C] (kg/kg/s) that moves each
    mixing ratio by up to +-50 % per hour, so some columns go negative before
    the limiter and the final clamp."""
    q = np.asarray(case.state.scalars, dtype=F32)
    r = np.random.default_rng(11).uniform(-0.5, 0.5, q.shape).astype(F32)
    return ((q * r) / F32(3600.0)).astype(F32)


def walk_transport(case, source=None):
    """Yield (checkpoint, oracle) around the three transport stages of a
    large step (after its three dynamics subcycles), with `source` as the
    held scalar tendency when given."""
    from hexdyn.case import initial_state_arrays
    from oracle.oracle import Oracle

    sched = schedule(case)
    c = case.config
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(1) == 0
    if source is not None:
        o.set_physics(scalars=source)
    o.begin_step()
    for sub in range(1, c.config_dynamics_split_steps + 1):
        o.begin_subcycle(sub)
        for st in (1, 2, 3):
            o.dyn_tend(st)
            for ss in range(1, sched[st - 1].acoustic_steps + 1):
                o.acoustic_substep(st, ss)
                o.divergence_damping(st)
            assert o.recover(st) == 0
            o.solve_diagnostics(st)
        o.end_subcycle(sub)
        yield f"subcycle_{sub}", o
    for st in (1, 2, 3):
        yield f"scalars_{st}", o
        o.scalars(st)
        yield f"scalars_{st}_out", o


TRANSPORT = [("scalar_old", "I:o_q", "tracer"), ("rho_zz_old", "I:o_rho", "cell"), ("rho_zz_new", "I:c_rho", "cell")]
TRANSPORT_FLUX = [("uh_avg", "I:split_u", "edge"), ("ww_avg", "I:split_w", "w")]


def shapes(case):
    L, C, E = case.n_levels, case.mesh.n_cells, case.mesh.n_edges
    return {"cell": (L, C), "edge": (L, E), "w": (L + 1, C), "tracer": (case.n_scalars, L, C),
            "vertex": (L, case.mesh.n_vertices)}


def mixing_density(mesh) -> np.ndarray:
    """A smooth variable mesh density in [1, 4] (a variable-resolution mesh's
    meshDensity), so the del2/del4 mesh scaling is not the identity."""
    lat = np.asarray(mesh.latCell, dtype=np.float64)
    lon = np.asarray(mesh.lonCell, dtype=np.float64)
    return (1.0 + 1.5 * (1.0 + np.sin(lat) * np.cos(lon))).astype(F32)


def build_mixing_case():
    """The x1.162 moist case at 10 levels with 2-D Smagorinsky mixing
    (visc4 0.05) and the variable mesh density above."""
    from hexdyn.case import build_case as build
    from hexdyn.config import DycoreConfig

    case = build(CASE, nlev=NLEV, use_cache=False,
                 config=DycoreConfig(config_dt=7200.0, config_horiz_mixing="2d_smagorinsky"))
    case.mesh.arrays["meshDensity"] = mixing_density(case.mesh)
    return case


def walk_mixing(case):
    """Yield ("mixing", oracle) after the subcycle-start pools of a developed
    step and ("mixing_out", oracle) after the horizontal mixing added to them."""
    from hexdyn.case import initial_state_arrays
    from oracle.oracle import Oracle

    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(1) == 0
    o.begin_step()
    o.begin_subcycle_pools()
    yield "mixing", o
    o.horizontal_mixing()
    yield "mixing_out", o


# what each reference operator reads, as (argument, source, shape); source is a
# published field id ("F:NAME") or an oracle working array ("I:name")
ACOUSTIC_STATE = [("ru_p", "F:RU_P", "edge"), ("rw_p", "F:RW_P", "w"), ("rtheta_pp", "F:RTHETA_PP", "cell"),
                  ("rtheta_pp_old", "I:rtheta_pp_old", "cell"), ("rho_pp", "F:RHO_PP", "cell"),
                  ("ru_avg", "F:RU_AVG", "edge"), ("ww_avg", "F:WW_AVG", "w")]
ACOUSTIC_FORCING = [("rho_zz", "I:c_rho", "cell"), ("theta_m", "I:s_theta", "cell"), ("exner", "I:s_exner", "cell"),
                    ("cqu", "F:CQU", "edge"), ("tend_ru", "F:TEND_RU", "edge"), ("tend_rho", "F:TEND_RHO", "cell"),
                    ("tend_rt", "F:TEND_RT", "cell"), ("tend_rw", "F:TEND_RW", "w"), ("w", "I:w", "w"),
                    ("rw", "I:c_rw", "w"), ("rw_save", "I:s_rw", "w")]
COEFFICIENTS = [("cofwr", "I:cofwr", "cell"), ("cofwz", "F:COFWZ", "cell"), ("coftz", "F:COFTZ", "w"),
                ("cofwt", "F:COFWT", "cell"), ("a_tri", "F:A_TRI", "cell"), ("alpha_tri", "F:ALPHA_TRI", "cell"),
                ("gamma_tri", "F:GAMMA_TRI", "cell")]
RECOVERY_STATE = [("ww_avg", "F:WW_AVG", "w"), ("rw_save", "I:s_rw", "w"), ("w", "I:w", "w"), ("rw", "I:c_rw", "w"),
                  ("rw_p", "F:RW_P", "w"), ("rtheta_p", "I:c_rtheta_p", "cell"),
                  ("rtheta_pp", "F:RTHETA_PP", "cell"), ("rtheta_p_save", "I:s_rtheta_p", "cell"),
                  ("rho_p", "I:c_rho_p", "cell"), ("rho_p_save", "I:s_rho_p", "cell"), ("rho_pp", "F:RHO_PP", "cell"),
                  ("rho_zz", "I:c_rho", "cell"), ("ru_avg", "F:RU_AVG", "edge"), ("ru_save", "I:s_ru", "edge"),
                  ("ru_p", "F:RU_P", "edge"), ("u", "I:u", "edge"), ("ru", "I:c_ru", "edge"),
                  ("exner", "I:exner", "cell"), ("pressure_p", "I:pp", "cell"), ("theta_m", "I:c_theta", "cell")]
# the oracle array holding each recovered output (the time averages are kept
# beside the acoustic sums, in FLUX_U / FLUX_W)
RECOVERY_OUT = [("rho_p", "I:c_rho_p", "cell"), ("rho_zz", "I:c_rho", "cell"), ("w", "I:w", "w"),
                ("ww_avg", "F:FLUX_W", "w"), ("rw", "I:c_rw", "w"), ("rtheta_p", "I:c_rtheta_p", "cell"),
                ("theta_m", "I:c_theta", "cell"), ("exner", "I:exner", "cell"), ("pressure_p", "I:pp", "cell"),
                ("ru_avg", "F:FLUX_U", "edge"), ("ru", "I:c_ru", "edge"), ("u", "I:u", "edge")]


# the v8.2.3-association momentum terms (dynamics.py:273-305) and the omega
# conversion (acoustic.py:246-281) at the stage-2 tendency checkpoint
MOMENTUM = [("normal_velocity", "I:u", "edge"), ("rho_zz", "I:c_rho", "cell"), ("pv_edge", "F:PV_EDGE", "edge"),
            ("kinetic_energy", "F:KE", "cell"), ("rho_u", "I:c_ru", "edge")]
OMEGA = [("w_tendency", "I:euler_rw", "w"), ("u_tendency", "F:TEND_RU", "edge")]
# the mixing's operands at the subcycle start, and the pools it adds to
MIXING = [("divergence", "I:divergence", "cell"), ("vorticity", "I:vorticity", "vertex"),
          ("rho_zz", "I:c_rho", "cell"), ("theta_m", "I:s_theta", "cell"), ("w", "I:w", "w"),
          ("euler_ru", "I:euler_ru", "edge"), ("euler_rw", "I:euler_rw", "w")]

POOLS = [("rho_u", "I:s_ru", "edge"), ("rho_w", "I:s_rw", "w"), ("pressure_perturbation", "I:pp", "cell"),
         ("dpdz", "I:dpdz", "cell"), ("cqu", "F:CQU", "edge")]
VCOEF = {1: [("cqw", "F:CQW", "cell"), ("exner", "I:s_exner", "cell"), ("theta", "I:s_theta", "cell"),
             ("rho_theta_perturbation", "I:s_rtheta_p", "cell"), ("qtot", "F:QTOT", "cell")],
         2: [("cqw", "F:CQW", "cell"), ("exner", "I:s_exner", "cell"), ("theta", "I:c_theta", "cell"),
             ("rho_theta_perturbation", "I:c_rtheta_p", "cell"), ("qtot", "F:QTOT", "cell")]}
VCOEF_OUT = [("cofwr", "I:cofwr", "cell"), ("cofwz", "F:COFWZ", "cell"), ("coftz", "F:COFTZ", "w"),
             ("cofwt", "F:COFWT", "cell"), ("a_tri", "F:A_TRI", "cell"), ("alpha_tri", "F:ALPHA_TRI", "cell"),
             ("gamma_tri", "F:GAMMA_TRI", "cell")]
DAMPING = [("rho_u_perturbation", "F:RU_P", "edge"), ("theta_m", "I:s_theta", "cell"),
           ("rtheta_pp", "F:RTHETA_PP", "cell"), ("rtheta_pp_old", "I:rtheta_pp_old", "cell")]


def read(o, case, source, shape):
    kind, name = source.split(":")
    a = o.export_field(name) if kind == "F" else o.export_internal(name)
    return a.reshape(shapes(case)[shape])


def read_all(o, case, table):
    return {name: read(o, case, src, shp) for name, src, shp in table}


def operand_digest(arrays) -> np.ndarray:
    """sha256 of the operands read from the oracle at a checkpoint, in table
    order (the test checks that its own walk reached the same operands)."""
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, dtype=np.float32).tobytes())
    return np.frombuffer(h.hexdigest().encode(), dtype=np.uint8)


def case_digest(case) -> np.ndarray:
    """sha256 over every binary32/int32 mesh, vertical and initial array the
    operators read."""
    m, g = case.mesh, case.grid
    h = hashlib.sha256()
    for a in (m.cellsOnEdge, m.edgesOnCell, m.verticesOnEdge, m.edgesOnEdge, m.nEdgesOnCell, m.nEdgesOnEdge,
              m.edgesOnVertex, m.verticesOnCell, m.kiteForCell, m.weightsOnEdge, m.dcEdge, m.dvEdge, m.areaCell,
              m.areaTriangle, m.kiteAreasOnVertex, m.fVertex, case.inv_area_cell, case.inv_area_triangle,
              case.inv_dc_edge, case.inv_dv_edge, case.adv_coefs, case.adv_coefs_3rd, case.adv_cells, case.n_adv,
              g.zz, g.zxu, g.rdzw, g.rdzu, g.fzm, g.fzp, g.zb_cell, g.zb3_cell, [g.cf1, g.cf2, g.cf3], case.dss,
              case.etp, case.etm, case.ewp, case.ewm, case.reference.rho_base, case.reference.rtheta_base,
              case.reference.exner_base, case.state.rho_zz, case.state.rtheta_m, case.state.ru, case.state.rw,
              case.state.scalars):
        arr = np.asarray(a)
        arr = arr.astype(F32) if arr.dtype.kind == "f" else arr.astype(np.int32)
        h.update(np.ascontiguousarray(arr).tobytes())
    return np.frombuffer(h.hexdigest().encode(), dtype=np.uint8)
