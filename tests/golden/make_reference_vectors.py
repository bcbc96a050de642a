#!/usr/bin/env python3
"""Write golden vectors of the reference's NumPy operator restatement.

Run in a development checkout where the reference tree is present (never at
test time, never on a GPU host):

    python tests/golden/make_reference_vectors.py /root/reference

Only the reference functions whose binary32 bits nothing else pins are
called (their exact operation order): solve diagnostics, the v8.4.1 implicit
coefficients and acoustic step, large-step recovery, the advection
coefficients and monotonic scalar transport.  Each call's inputs and outputs
are written to `tests/golden/reference_<name>.npz` (data only) together with
a digest of the mesh and vertical arrays they were made on;
`tests/test_reference_vectors.py` feeds the same inputs to the oracle's phase
and compares bits.

Inputs are this package's own x1.162 moist case at 10 levels (seeded, so the
test rebuilds the identical case) and states the oracle produces from it, so
every operand is realistic and non-zero.
"""
from __future__ import annotations

import sys
import types
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT / "gpuwm-hex_amd"))
sys.path.insert(0, str(ROOT))

from reference_walk import (ACOUSTIC_FORCING, ACOUSTIC_STATE, COEFFICIENTS, DAMPING, MIXING, MOMENTUM,  # noqa: E402
                            OMEGA, POOLS, RECOVERY_STATE, TRANSPORT, TRANSPORT_FLUX, VCOEF, build_case as build,
                            build_mixing_case, case_digest, operand_digest, read, read_all, scalar_tendency,
                            schedule, walk, walk_mixing, walk_transport)

F32 = np.float32


def mesh_view(case, geometry=F32):
    """The mesh as the reference's operators read it: MPAS names, 0-based
    connectivity, binary32 geometry (the values the oracle is given)."""
    m = case.mesh
    ns = types.SimpleNamespace()
    names = ["cellsOnEdge", "verticesOnEdge", "edgesOnCell", "cellsOnCell", "verticesOnCell", "nEdgesOnCell",
             "kiteForCell", "edgesOnVertex", "cellsOnVertex", "edgesOnEdge", "nEdgesOnEdge", "weightsOnEdge",
             "dcEdge", "dvEdge", "areaCell", "areaTriangle", "kiteAreasOnVertex", "fEdge", "fVertex",
             "angleEdge", "xCell", "yCell", "zCell", "xEdge", "yEdge", "zEdge", "xVertex", "yVertex", "zVertex",
             "latCell", "lonCell", "latEdge", "lonEdge", "meshDensity"]
    for n in names:
        try:
            a = np.asarray(getattr(m, n))
        except AttributeError:
            if n in m.arrays:
                a = np.asarray(m.arrays[n])
            else:
                continue
        setattr(ns, n, a.astype(geometry) if a.dtype.kind == "f" else a.astype(np.int64))
    ns.nCells, ns.nEdges, ns.nVertices = m.n_cells, m.n_edges, m.n_vertices
    ns.maxEdges = m.max_edges
    ns.arrays = {n: getattr(ns, n) for n in names if hasattr(ns, n)}
    return ns


def oracle_state(case, steps: int):
    from hexdyn.case import initial_state_arrays
    from oracle.oracle import Oracle

    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    if steps:
        assert o.step(steps) == 0
    return o


def solve_diagnostics_vectors(case):
    """diagnostics.compute_solve_diagnostics (diagnostics.py:95-306) at rk
    step 3 (v reconstructed) and rk step 1 (v from the cache)."""
    from hexcore.diagnostics import compute_solve_diagnostics

    L, C, E = case.n_levels, case.mesh.n_cells, case.mesh.n_edges
    o = oracle_state(case, 1)
    u = o.export_field("U").reshape(L, E)
    rho = o.export_field("RHO_ZZ").reshape(L, C)
    # the cache an rk step 1 reads: v of the previous large step's start state
    v_cached = oracle_state(case, 0).export_field("V").reshape(L, E)
    assert np.count_nonzero(v_cached != o.export_field("V").reshape(L, E)) > 0
    mesh = mesh_view(case)
    cfg = case.config
    out = {"u": u, "rho_zz": rho, "v_cached": v_cached}
    for rk in (3, 1):  # kinetic energy, divergence and vorticity do not depend on the rk step
        d = compute_solve_diagnostics(
            mesh, u, rho, dt=float(F32(cfg.config_dt)), apvm_upwinding=float(cfg.config_apvm_upwinding),
            rk_step=rk, cached_tangential_velocity=v_cached if rk != 3 else None,
            f_vertex=mesh.fVertex, inv_area_cell=case.inv_area_cell.astype(F32),
            inv_area_triangle=case.inv_area_triangle.astype(F32), inv_dc_edge=case.inv_dc_edge.astype(F32),
            inv_dv_edge=case.inv_dv_edge.astype(F32))
        fields = [("V", d.tangential_velocity), ("PV_EDGE", d.pv_edge)]
        if rk == 3:
            fields += [("KE", d.kinetic_energy), ("DIVERGENCE", d.divergence), ("VORTICITY", d.vorticity)]
        for name, val in fields:
            assert val.dtype == F32, name
            out[f"rk{rk}_{name}"] = np.ascontiguousarray(val)
    return out


def _offcentering(case):
    from hexcore.offcentering_v841 import AcousticOffcenteringV841

    L = case.n_levels
    return AcousticOffcenteringV841(etp=case.etp.astype(F32), etm=case.etm.astype(F32), ewp=case.ewp.astype(F32),
                                    ewm=case.ewm.astype(F32), height_u_levels=np.zeros(L, F32),
                                    height_w_levels=np.zeros(L + 1, F32))


def _levels(case):
    g = case.grid
    return {n: np.asarray(getattr(g, n), dtype=F32) for n in ("rdzw", "fzm", "fzp", "rdzu")}


def vertical_coefficient_vectors(case):
    """acoustic_v841.compute_vertical_implicit_coefficients_v841
    (acoustic_v841.py:81-258) at stage 1 (subcycle-start theta and
    rtheta_p) and stage 2 (the stage-1 recovered ones), the exner of the
    subcycle start, dts of each stage."""
    from hexcore.acoustic_v841 import compute_vertical_implicit_coefficients_v841

    sched = schedule(case)
    lev = _levels(case)
    zz = case.grid.zz.astype(F32)
    ref = case.reference
    out = {}
    for label, o in walk(case):
        if label not in ("vcoef_1", "vcoef_2"):
            continue
        stage = int(label[-1])
        ops = read_all(o, case, VCOEF[stage])
        res = compute_vertical_implicit_coefficients_v841(
            dts=sched[stage - 1].acoustic_timestep, offcentering=_offcentering(case), zz=zz,
            rho_base=ref.rho_base.astype(F32), rho_theta_base=ref.rtheta_base.astype(F32),
            exner_base=ref.exner_base.astype(F32), **ops, **lev)
        out[f"{label}.operands"] = operand_digest(ops.values())
        for name in ("cofwr", "cofwz", "coftz", "cofwt", "a_tri", "b_tri", "c_tri", "alpha_tri", "gamma_tri"):
            val = np.asarray(getattr(res, name))
            assert val.dtype == F32, name
            out[f"{label}.{name}"] = np.ascontiguousarray(val)
    return out


def _coefficients(o, case):
    from hexcore.acoustic_v841 import VerticalImplicitCoefficientsV841

    L, C = case.n_levels, case.mesh.n_cells
    vals = {name: read(o, case, src, shp) for name, src, shp in COEFFICIENTS}
    b64, c64 = o.btri_ctri_f64()
    vals["b_tri"] = b64.astype(F32).reshape(L, C)
    vals["c_tri"] = c64.astype(F32).reshape(L, C)
    vals["cofrz"] = np.asarray(case.grid.rdzw, dtype=F32)
    return VerticalImplicitCoefficientsV841(**vals), vals


def acoustic_vectors(case):
    """acoustic_v841.advance_acoustic_step_v841 (acoustic_v841.py:299-560)
    for the stage-2 substep and both stage-3 substeps, fed the oracle's own
    coefficients, tendencies and acoustic state at each substep."""
    from hexcore.acoustic import AcousticStepForcing, AcousticStepState
    from hexcore.acoustic_v841 import advance_acoustic_step_v841

    sched = schedule(case)
    lev = _levels(case)
    mesh = mesh_view(case)
    out = {}
    for label, o in walk(case):
        if not label.startswith("acoustic_") or label.endswith("_out"):
            continue
        stage, ss = int(label.split("_")[1]), int(label.split("_")[2])
        state = read_all(o, case, ACOUSTIC_STATE)
        forcing = read_all(o, case, ACOUSTIC_FORCING)
        coefs, cvals = _coefficients(o, case)
        out[f"{label}.operands"] = operand_digest(list(state.values()) + list(forcing.values()) +
                                                  list(cvals.values()))
        forcing["zz"] = case.grid.zz.astype(F32)
        forcing["zxu"] = case.grid.zxu.astype(F32)
        forcing["dss"] = np.asarray(case.dss, dtype=F32).reshape(shapes_cell(case))
        res = advance_acoustic_step_v841(mesh, AcousticStepState(**state), AcousticStepForcing(**forcing), coefs,
                                         dts=sched[stage - 1].acoustic_timestep, small_step=ss,
                                         offcentering=_offcentering(case), fzm=lev["fzm"], fzp=lev["fzp"],
                                         rdzw=lev["rdzw"])
        for name, _, _ in ACOUSTIC_STATE:
            val = np.asarray(getattr(res, name))
            assert val.dtype == F32, name
            out[f"{label}.{name}"] = np.ascontiguousarray(val)
    return out


def shapes_cell(case):
    return (case.n_levels, case.mesh.n_cells)


def damping_vectors(case):
    """mixing.divergence_damping_3d (mixing.py:851-932) with the stored
    inverse of dts, after the stage-2 substep and the second stage-3 one."""
    from hexcore.mixing import divergence_damping_3d

    sched = schedule(case)
    mesh = mesh_view(case)
    cfg = case.config
    out = {}
    for label, o in walk(case):
        if label not in ("damping_2", "damping_3"):
            continue
        stage = int(label[-1])
        args = list(read_all(o, case, DAMPING).values())
        res = divergence_damping_3d(mesh, *args, dts=sched[stage - 1].acoustic_timestep, config_smdiv=cfg.config_smdiv,
                                    config_len_disp=float(case.len_disp), native_stored_inverse=True)
        assert res.dtype == F32
        out[f"{label}.operands"] = operand_digest(args)
        out[f"{label}.ru_p"] = np.ascontiguousarray(res)
    return out


def recovery_vectors(case):
    """integration.recover_large_step_variables (integration.py:198-327)
    at stages 2 and 3 (exner and pressure at stage 3), no diabatic theta
    tendency, the case's terrain metrics zb/zb3."""
    from hexcore.integration import RecoveryBackground, RecoveryState, recover_large_step_variables

    sched = schedule(case)
    lev = _levels(case)
    mesh = mesh_view(case)
    g, ref = case.grid, case.reference
    bg = RecoveryBackground(rho_base=ref.rho_base.astype(F32), exner_base=ref.exner_base.astype(F32),
                            rtheta_base=ref.rtheta_base.astype(F32), zz=g.zz.astype(F32))
    out = {}
    for label, o in walk(case):
        if label not in ("recover_2", "recover_3"):
            continue
        stage = int(label[-1])
        state = read_all(o, case, RECOVERY_STATE)
        res = recover_large_step_variables(
            mesh, RecoveryState(**state), bg, dt=sched[2].large_timestep, acoustic_steps=sched[stage - 1].acoustic_steps,
            rk_step=stage, rt_diabatic_tendency=np.zeros(shapes_cell(case), F32), fzm=lev["fzm"], fzp=lev["fzp"],
            cf1=float(F32(g.cf1)), cf2=float(F32(g.cf2)), cf3=float(F32(g.cf3)),
            zb_cell=np.asarray(g.zb_cell, dtype=F32), zb3_cell=np.asarray(g.zb3_cell, dtype=F32))
        out[f"{label}.operands"] = operand_digest(state.values())
        for name in ("rho_p", "rho_zz", "w", "ww_avg", "rw", "rtheta_p", "theta_m", "exner", "pressure_p",
                     "ru_avg", "ru", "u"):
            val = np.asarray(getattr(res, name))
            assert val.dtype == F32, name
            out[f"{label}.{name}"] = np.ascontiguousarray(val)
        if stage == 3:
            # This NumPy restatement raises to the power with NumPy's binary32
            # `**`, which on this host is not the GNU powf the v8.4.1 path
            # selects (cuda_driver.py:1466-1472).  Exner and pressure are
            # restated here from the function's own argument (integration.py
            # :267-278) with the reference's GNU powf (native_powf.py).
            from hexcore.native_powf import powf_native_gnu

            full = res.rtheta_p + bg.rtheta_base
            arg = (bg.zz * F32(287.0 / 100000.0)) * full
            rcv = F32(287.0 / (1004.5 - 287.0))
            ex = np.asarray(powf_native_gnu(arg, rcv), dtype=F32)
            pp = (bg.zz * F32(287.0)) * (ex * res.rtheta_p + bg.rtheta_base * (ex - bg.exner_base))
            n_pow = int(np.count_nonzero(ex.view(np.uint32) != res.exner.view(np.uint32)))
            print(f"recovery stage 3: NumPy ** differs from the GNU powf on {n_pow} of {ex.size} exner values")
            out[f"{label}.exner"] = np.ascontiguousarray(ex)
            out[f"{label}.pressure_p"] = np.ascontiguousarray(pp.astype(F32))
            out[f"{label}.numpy_pow_differs"] = np.array([n_pow], dtype=np.int64)
    return out


def powf_vectors(case):
    """native_powf.powf_native_gnu (native_powf.py:18-60), the GNU powf the
    v8.4.1 exner recovery uses (cuda_driver.py:1466-1472), at the exner
    exponent R/cv over every binary32 argument in [2^-12, 4): the arguments
    where it differs from the product's correctly rounded binary64 method,
    with the reference's values there.  Also counts where the host C
    library's powf differs from the reference (the oracle and the device
    correction use the host's)."""
    from hexcore.native_powf import powf_native_gnu
    from oracle.oracle import powf_many

    y = F32(287.0) / (F32(1004.5) - F32(287.0))
    lo = int(np.float32(2.0 ** -12).view(np.uint32))
    hi = int(np.float32(4.0).view(np.uint32))
    xs_bits, refs, libm_differs = [], [], 0
    for start in range(lo, hi, 1 << 22):
        x = np.arange(start, min(start + (1 << 22), hi), dtype=np.uint32).view(np.float32)
        ref = np.asarray(powf_native_gnu(x, y), dtype=F32)
        ours = powf_many(x, y, "hexdyn")
        host = powf_many(x, y, "libm")
        d = ref.view(np.uint32) != ours.view(np.uint32)
        xs_bits.append(x.view(np.uint32)[d])
        refs.append(ref[d])
        libm_differs += int(np.count_nonzero(ref.view(np.uint32) != host.view(np.uint32)))
    xb = np.concatenate(xs_bits)
    print(f"powf: {xb.size} of {hi - lo} arguments differ from the binary64 method; "
          f"host libm differs from the reference on {libm_differs}")
    return {"exponent": np.array([y], dtype=F32), "range_bits": np.array([lo, hi], dtype=np.uint32),
            "x_bits": xb, "reference": np.concatenate(refs),
            "host_libm_differs": np.array([libm_differs], dtype=np.int64)}


def advection_vectors(case):
    """The set-up builders of the advection stencils: transport.py
    initialize_deriv_two (:610-738, the quadratic-fit second derivatives)
    and the v8.4.1 compression of adv_coefs/adv_coefs_3rd in native
    insertion order (_native_advection_order, :2264-2403, and
    _native_advection_compression, :2405-2443; build_advection_coefficients
    itself calls the absent Rust engine first), fed this package's stencil
    and deriv_two."""
    from hexcore.transport import _native_advection_compression, _native_advection_order, initialize_deriv_two
    from hexdyn import setup as hsetup

    m = case.mesh
    mesh64 = mesh_view(case, geometry=np.float64)
    mesh64.sphere_radius = float(m.sphere_radius)
    ours_d2 = hsetup.deriv_two(m)
    ref_d2 = np.asarray(initialize_deriv_two(mesh64))
    w = ours_d2.shape[2]
    assert not ref_d2[:, :, w:].any()
    a, b = ref_d2[:, :, :w], ours_d2
    rel = float(np.abs(a - b).max() / np.abs(a).max())
    same32 = float(np.mean(a.astype(F32).view(np.uint32) == b.astype(F32).view(np.uint32)))
    print(f"deriv_two (binary64): max |diff| / max |ref| = {rel:.3e}; equal after rounding to binary32: {same32:.4f}")
    counts = np.asarray(m.nEdgesOnCell, dtype=np.int64)
    coc = np.asarray(m.cellsOnCell, dtype=np.int64)
    coe = np.asarray(m.cellsOnEdge, dtype=np.int64)
    d2 = np.asarray(ours_d2, dtype=F32)
    dc, dv = m.dcEdge.astype(F32), m.dvEdge.astype(F32)
    nadv = np.asarray(case.n_adv, dtype=np.int64)
    cells = np.asarray(case.adv_cells, dtype=np.int64)
    zeros = np.zeros(cells.shape, F32)
    _, _, ordered = _native_advection_order(counts, coc, coe, zeros, zeros, nadv, cells, garbage=-1)
    adv, adv3 = _native_advection_compression(counts, coc, coe, d2, dc, dv, nadv, cells)
    return {"deriv_two_reference": ref_d2[:, :, :w], "deriv_two": d2, "adv_cells": np.asarray(ordered),
            "adv_coefs": np.asarray(adv, F32), "adv_coefs_3rd": np.asarray(adv3, F32)}


def transport_vectors(case, source=None):
    """transport.advance_scalars (:1649-1826, stages 1 and 2, split-transport
    density weights) and advance_scalars_monotonic (:1829-2175, stage 3,
    Zalesak FCT with the split-transport target density), stored-inverse
    area, third-order fluxes with the upwind blend, fed the oracle's own
    time-averaged fluxes and scalars of each stage."""
    from hexcore.transport import advance_scalars, advance_scalars_monotonic
    from hexdyn.setup import rk_schedule

    c = case.config
    sca = rk_schedule(c.config_dt, order=c.config_time_integration_order,
                      acoustic_substeps=c.config_number_of_sub_steps, dynamics_splits=1)
    lev = _levels(case)
    mesh = mesh_view(case)
    common = dict(adv_coefs=case.adv_coefs.astype(F32), adv_coefs_3rd=case.adv_coefs_3rd.astype(F32),
                  n_adv_cells_for_edge=np.asarray(case.n_adv), adv_cells_for_edge=np.asarray(case.adv_cells),
                  fzm=lev["fzm"], fzp=lev["fzp"], rdzw=lev["rdzw"],
                  config_coef_3rd_order=float(c.config_coef_3rd_order), advance_density=True,
                  inv_area_cell=case.inv_area_cell.astype(F32))
    if source is not None:
        common["scalar_tendency"] = source
    out = {}
    held = {}
    for label, o in walk_transport(case, source):
        if not label.startswith("scalars_"):
            continue
        st = int(label.split("_")[1])
        if label.endswith("_out"):
            if st == 1:
                held.update(read_all(o, case, TRANSPORT_FLUX))  # the 1/splits average is taken at stage 1
            ops = held[st]
            stage_in = ops.pop("scalar_stage")
            args = dict(ops, scalar_stage=stage_in, **{k: held[k] for k in ("uh_avg", "ww_avg")})
            if st < 3:
                res = advance_scalars(mesh, args["scalar_old"], args["scalar_stage"], args["rho_zz_old"],
                                      args["rho_zz_new"], args["uh_avg"], args["ww_avg"],
                                      sca[st - 1].large_timestep, rk_step=st,
                                      config_time_integration_order=c.config_time_integration_order, **common)
            else:
                res = advance_scalars_monotonic(mesh, args["scalar_old"], args["scalar_stage"], args["rho_zz_old"],
                                                args["rho_zz_new"], args["uh_avg"], args["ww_avg"],
                                                sca[2].large_timestep, **common)
                plain = advance_scalars(mesh, args["scalar_old"], args["scalar_stage"], args["rho_zz_old"],
                                        args["rho_zz_new"], args["uh_avg"], args["ww_avg"], sca[2].large_timestep,
                                        rk_step=3, config_time_integration_order=c.config_time_integration_order,
                                        **common)
                limited = int(np.count_nonzero(np.asarray(res.scalars) != np.asarray(plain.scalars)))
                print(f"transport stage 3: the limiter changes {limited} of {res.scalars.size} mixing ratios")
                assert limited > 0
                out["scalars_3.limited"] = np.array([limited], dtype=np.int64)
            val = np.asarray(res.scalars)
            assert val.dtype == F32
            out[f"{label[:-4]}.operands"] = operand_digest(
                [args[k] for k in ("scalar_old", "scalar_stage", "rho_zz_old", "rho_zz_new", "uh_avg", "ww_avg")])
            out[f"{label[:-4]}.scalars"] = np.ascontiguousarray(val)
            out[f"{label[:-4]}.density"] = np.ascontiguousarray(np.asarray(res.density, dtype=F32))
        else:
            ops = read_all(o, case, TRANSPORT)
            ops["scalar_stage"] = ops["scalar_old"] if st == 1 else read(o, case, "F:SCALARS_STAGE", "tracer")
            held[st] = ops
    return out


def pool_vectors(case):
    """The subcycle-start pools: dynamics.density_tendency (:122-148, with
    mass_flux_divergence :86-119) and dynamics.pressure_gradient_euler_tendency
    (:188-224), stored inverses, from the subcycle-start rho_u, rho_w,
    pressure perturbation and moist dp/dz."""
    from hexcore.dynamics import density_tendency, pressure_gradient_euler_tendency

    mesh = mesh_view(case)
    g = case.grid
    out = {}
    for label, o in walk(case):
        if label != "vcoef_1":
            continue
        ops = read_all(o, case, POOLS)
        rho = density_tendency(mesh, ops["rho_u"], ops["rho_w"], np.asarray(g.rdzw, F32),
                               inv_area_cell=case.inv_area_cell.astype(F32))
        pgf = pressure_gradient_euler_tendency(mesh, ops["pressure_perturbation"], ops["dpdz"], ops["cqu"],
                                               g.zz.astype(F32), g.zxu.astype(F32),
                                               inv_dc_edge=case.inv_dc_edge.astype(F32))
        assert rho.dtype == F32 and pgf.dtype == F32
        out["pools_1.operands"] = operand_digest(ops.values())
        out["pools_1.tend_rho"] = np.ascontiguousarray(rho)
        out["pools_1.euler_ru"] = np.ascontiguousarray(pgf)
        break
    return out


def damping_profile_vectors(case):
    """damping_v841.build_v841_vertical_velocity_damping (:23-107), the
    binary32 path (the host C library's sinf, the reference's default
    numerics), on the zgrid of the x1.162 case at 55 levels to 30 km with
    the Rayleigh layer from 22 km."""
    from hexcore.damping_v841 import build_v841_vertical_velocity_damping
    from hexdyn.case import build_case as build55

    c = build55("x1.162-moist", nlev=55, use_cache=False)
    z = np.asarray(c.grid.zgrid, dtype=F32)
    cfg = c.config
    dss = build_v841_vertical_velocity_damping(z, xnutr=cfg.config_xnutr, damping_start_z=cfg.config_zd)
    assert dss.dtype == F32 and np.count_nonzero(dss) > 1000
    return {"zgrid": z, "xnutr_zd": np.array([cfg.config_xnutr, cfg.config_zd]), "dss": np.ascontiguousarray(dss)}


def offcentering_vectors(case):
    """offcentering_v841.build_v841_acoustic_offcentering (:45-174) on a
    60-level stretched column to 45 km, so the sine transition between 30
    and 50 km is exercised (the workloads' 30 km tops never reach it)."""
    from hexcore.offcentering_v841 import build_v841_acoustic_offcentering

    k = np.arange(60, dtype=np.float64)
    dz = (150.0 + 1500.0 * (k / 59.0) ** 1.5).astype(F32)
    dz = (dz * F32(45000.0 / float(dz.astype(np.float64).sum()))).astype(F32)
    rdzw = (F32(1.0) / dz).astype(F32)
    res = build_v841_acoustic_offcentering(rdzw)
    assert float(res.height_w_levels[-1]) > 40000.0
    return {"rdzw": rdzw, **{n: np.asarray(getattr(res, n), dtype=F32) for n in ("etp", "etm", "ewp", "ewm")}}


def transport_source_vectors(case):
    """The three transport stages as in transport_vectors, with a held
    scalar tendency (the transport's scalar_tendency, cuda_driver.py:5162-5163):
    added to the flux divergence at stages 1-2 (transport.py:1800-1808) and
    to the old mixing ratio as dt*source/rho_old before the limiter (:1906)."""
    return transport_vectors(case, scalar_tendency(case))


def vertical_transport_vectors(case):
    """dynamics.vertical_transport_u (:227-270), the vertical transport of
    normal momentum that opens the momentum tendency, with the stage-1
    recovered u and rw (the operands of the stage-2 tendencies)."""
    from hexcore.dynamics import vertical_transport_u

    mesh = mesh_view(case)
    lev = _levels(case)
    for label, o in walk(case):
        if label != "vcoef_2":
            continue
        ops = read_all(o, case, [("normal_velocity", "I:u", "edge"), ("rho_w", "I:c_rw", "w")])
        res = vertical_transport_u(mesh, ops["normal_velocity"], ops["rho_w"], fzm=lev["fzm"], fzp=lev["fzp"],
                                   rdzw=lev["rdzw"])
        assert res.dtype == F32
        return {"vtu_2.operands": operand_digest(ops.values()), "vtu_2.result": np.ascontiguousarray(res)}


def split_flux_vectors(case):
    """integration.accumulate_split_flux (:27-55) over the three dynamics
    subcycles' time-averaged mass fluxes, then finish_split_flux (:58-82):
    the transport's uh_avg and ww_avg."""
    from hexcore.integration import accumulate_split_flux, finish_split_flux

    acc_u = acc_w = None
    fluxes = []
    for label, o in walk_transport(case):
        if label.startswith("subcycle_"):
            fu, fw = read(o, case, "F:FLUX_U", "edge"), read(o, case, "F:FLUX_W", "w")
            fluxes += [fu, fw]
            acc_u = accumulate_split_flux(fu, acc_u)
            acc_w = accumulate_split_flux(fw, acc_w)
        elif label == "scalars_1":
            break
    n = case.config.config_dynamics_split_steps
    return {"split.operands": operand_digest(fluxes),
            "split.uh_avg": np.ascontiguousarray(finish_split_flux(acc_u, n)),
            "split.ww_avg": np.ascontiguousarray(finish_split_flux(acc_w, n))}


def moist_coefficient_vectors(case):
    """regional_v841.compute_moist_coefficients (:1180-1219, the
    atm_compute_moist_coefficients restatement): qtot, cqw, cqu from the
    step-start scalars, every carried species loading the air."""
    from hexcore.regional_v841 import compute_moist_coefficients

    for label, o in walk(case):
        if label != "vcoef_1":
            continue
        q = read(o, case, "I:o_q", "tracer")
        qtot, cqw, cqu = compute_moist_coefficients(
            q, moist_indices=list(range(case.n_moist)),
            cells_on_edge_remapped=np.asarray(case.mesh.cellsOnEdge, dtype=np.int64), nlev=case.n_levels)
        assert qtot.dtype == F32 and cqw.dtype == F32 and cqu.dtype == F32
        return {"moist.operands": operand_digest([q]), "moist.qtot": np.ascontiguousarray(qtot),
                "moist.cqw": np.ascontiguousarray(cqw), "moist.cqu": np.ascontiguousarray(cqu)}


def momentum_vectors(case):
    """dynamics.vector_invariant_momentum_tendency (:273-305), the nonlinear
    PV/KE terms, fed the reference's own mass_flux_divergence (:86-121) of
    the current rho_u; reference wind zero (the v8.4.1 form then adds
    nothing)."""
    from hexcore.dynamics import mass_flux_divergence, vector_invariant_momentum_tendency

    assert not np.any(case.u_init) and not np.any(case.v_init)
    mesh = mesh_view(case)
    c0, c1 = np.asarray(case.mesh.cellsOnEdge, dtype=np.int64).T
    for label, o in walk(case):
        if label != "vcoef_2":
            continue
        ops = read_all(o, case, MOMENTUM)
        rho = ops["rho_zz"]
        rho_edge = (F32(0.5) * (rho[:, c0] + rho[:, c1])).astype(F32)
        hdiv = mass_flux_divergence(mesh, ops["rho_u"], inv_area_cell=case.inv_area_cell.astype(F32))
        res = vector_invariant_momentum_tendency(
            mesh, normal_velocity=ops["normal_velocity"], rho_edge=rho_edge, pv_edge=ops["pv_edge"],
            kinetic_energy=ops["kinetic_energy"], horizontal_divergence=hdiv)
        assert hdiv.dtype == F32 and res.dtype == F32
        return {"momentum.operands": operand_digest(ops.values()), "momentum.hdiv": np.ascontiguousarray(hdiv),
                "momentum.result": np.ascontiguousarray(res)}


def omega_vectors(case):
    """acoustic.convert_w_tendency_to_omega (:246-281) of a realistic w
    tendency (the stage's Euler w pool) with the stage's momentum tendency."""
    from hexcore.acoustic import convert_w_tendency_to_omega

    mesh = mesh_view(case)
    lev = _levels(case)
    g = case.grid
    for label, o in walk(case):
        if label != "vcoef_2":
            continue
        ops = read_all(o, case, OMEGA)
        res = convert_w_tendency_to_omega(mesh, ops["w_tendency"], ops["u_tendency"], fzm=lev["fzm"], fzp=lev["fzp"],
                                          zz=np.asarray(g.zz, F32), zb_cell=np.asarray(g.zb_cell, F32),
                                          zb3_cell=np.asarray(g.zb3_cell, F32))
        assert res.dtype == F32
        return {"omega.operands": operand_digest(ops.values()), "omega.result": np.ascontiguousarray(res)}


def mixing_vectors(_case):
    """mixing.compute_mesh_mixing_scaling (:271-323) on a variable mesh
    density, and the del2/del4 filters the v8.4.1 mixing applies with its
    kdiff (mixing_v841.py:709-790): momentum_horizontal_filter_tendency
    (:391-530) onto the subcycle-start u pool,
    vertical_momentum_horizontal_filter_tendency (:533-635) and
    theta_horizontal_filter_tendency (:638-733), native stored-inverse
    association.  kdiff is the oracle's (mixing_v841's viscosity needs the
    absent Rust engine)."""
    from hexcore.mixing import (compute_mesh_mixing_scaling, momentum_horizontal_filter_tendency,
                                theta_horizontal_filter_tendency, vertical_momentum_horizontal_filter_tendency)

    case = build_mixing_case()
    mesh = mesh_view(case)
    cfg = case.config
    length = F32(case.len_disp)
    h4 = F32(F32(cfg.config_visc4_2dsmag) * ((length * length) * length))
    sc = compute_mesh_mixing_scaling(mesh, config_h_ScaleWithMesh=True, dtype=np.float32)
    assert sc.del2.dtype == F32 and np.count_nonzero(sc.del2 != F32(1.0)) > 0
    out = {"mixing_case_digest": case_digest(case),
           "mixing_density": np.asarray(case.mesh.arrays["meshDensity"], F32),
           "scaling.del2": np.ascontiguousarray(sc.del2), "scaling.del4": np.ascontiguousarray(sc.del4)}
    c0, c1 = np.asarray(case.mesh.cellsOnEdge, dtype=np.int64).T
    ops = None
    for label, o in walk_mixing(case):
        if label == "mixing":
            ops = read_all(o, case, MIXING)
            continue
        kdiff = read(o, case, "I:kdiff", "cell")
        rho = ops["rho_zz"]
        rho_edge = (F32(0.5) * (rho[:, c0] + rho[:, c1])).astype(F32)
        common = dict(rho_edge=rho_edge, kdiff=kdiff, mesh_scaling_del2=sc.del2, mesh_scaling_del4=sc.del4,
                      native_stored_inverse=True)
        mom = momentum_horizontal_filter_tendency(
            mesh, divergence=ops["divergence"], vorticity=ops["vorticity"], h_mom_eddy_visc4=float(h4),
            config_del4u_div_factor=float(F32(cfg.config_del4u_div_factor)), initial_tendency=ops["euler_ru"],
            **common)
        wf = vertical_momentum_horizontal_filter_tendency(mesh, vertical_velocity=ops["w"],
                                                          h_mom_eddy_visc4=float(h4), **common)
        th = theta_horizontal_filter_tendency(mesh, theta_m=ops["theta_m"], h_theta_eddy_visc4=float(h4), **common)
        for a in (mom.tendency, wf.tendency, th.tendency):
            assert a.dtype == F32
        out.update({"mixing.operands": operand_digest(list(ops.values()) + [kdiff]),
                    "mixing.euler_ru": np.ascontiguousarray(mom.tendency),
                    "mixing.w_tendency": np.ascontiguousarray(wf.tendency),
                    "mixing.theta_tendency": np.ascontiguousarray(th.tendency)})
    return out


VECTORS = {
    "solve_diagnostics": solve_diagnostics_vectors,
    "vertical_coefficients": vertical_coefficient_vectors,
    "acoustic": acoustic_vectors,
    "divergence_damping": damping_vectors,
    "recovery": recovery_vectors,
    "powf": powf_vectors,
    "advection": advection_vectors,
    "transport": transport_vectors,
    "pools": pool_vectors,
    "damping_profile": damping_profile_vectors,
    "offcentering": offcentering_vectors,
    "transport_source": transport_source_vectors,
    "vertical_transport": vertical_transport_vectors,
    "split_flux": split_flux_vectors,
    "moist_coefficients": moist_coefficient_vectors,
    "momentum": momentum_vectors,
    "omega": omega_vectors,
    "mixing": mixing_vectors,
}


def main(argv):
    ref_root = Path(argv[1] if len(argv) > 1 else "/root/reference")
    sys.path.insert(0, str(ref_root / "tree" / "src"))
    import hexcore  # noqa: F401  (the reference package)

    case = build()
    digest = case_digest(case)
    import hexcore.diagnostics  # noqa: F401  (fail early if an operator module does not import)
    want = argv[2:] or list(VECTORS)
    for name in want:
        arrays = VECTORS[name](case)
        arrays["case_digest"] = digest
        path = HERE / f"reference_{name}.npz"
        np.savez_compressed(path, **arrays)
        print(f"{path.name}: {len(arrays) - 1} arrays, {path.stat().st_size / 1024:.0f} KiB")


if __name__ == "__main__":
    main(sys.argv)
