"""The CPU oracle pinned by what the mathematics fixes: an exact rest-state
fixed point, the Thomas factorization of the assembled tridiagonal system,
conservation of dry and tracer mass (FP64), Zalesak bounds, and a steady
balanced jet."""
import dataclasses

import numpy as np
import pytest

from hexdyn import lib as hlib
from hexdyn.case import initial_state_arrays
from oracle.oracle import Oracle


def _rest_case(case):
    """Reference state == model state, with exner_base from the model's own
    power function so the rest state is exactly representable."""
    L = hlib.lib()
    zz = case.grid.zz
    rtb = case.reference.rtheta_base
    cfac = np.float32(287.0 / 100000.0)
    rcv = np.float32(287.0 / 717.5)
    arg = (zz * cfac) * rtb
    exb = np.array([L.hexdyn_powf(float(a), float(rcv)) for a in arg.ravel()], dtype=np.float32).reshape(arg.shape)
    ref = dataclasses.replace(case.reference, exner_base=exb)
    st = dataclasses.replace(case.state, rho_zz=case.reference.rho_base.copy(), rtheta_m=rtb.copy(),
                             ru=np.zeros_like(case.state.ru), rw=np.zeros_like(case.state.rw),
                             scalars=np.zeros_like(case.state.scalars))
    return dataclasses.replace(case, reference=ref, state=st)


def test_rest_state_is_exact_fixed_point(get_case):
    case = _rest_case(get_case("x1.642-dry"))
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    before = o.get_state()
    assert o.step(2) == 0
    after = o.get_state()
    for name in ("rho_zz", "rtheta_m", "rw", "exner", "pressure_p", "w"):
        assert np.array_equal(before[name], after[name]), name
    assert np.abs(after["ru"]).max() == 0.0 and np.abs(after["u"]).max() == 0.0


def test_rest_state_mutation_is_caught(get_case):
    # a perturbation of one cell's theta must move the state (the test is not vacuous)
    case = _rest_case(get_case("x1.642-dry"))
    st = initial_state_arrays(case)
    st["rtheta_m"] = st["rtheta_m"].copy()
    st["rtheta_m"][10, 5] *= np.float32(1.001)
    o = Oracle(case, 32)
    o.set_state(st, rebuild=True)
    o.step(1)
    assert np.abs(o.get_state()["w"]).max() > 1e-6


def _dz(case):
    return (1.0 / case.grid.rdzw.astype(np.float64))[:, None]


def test_dry_and_tracer_mass_conserved_fp64(get_case):
    case = get_case("x1.642-moist")
    o = Oracle(case, 64)
    o.set_state(initial_state_arrays(case), rebuild=True)
    # the step's own cell measure: the stored binary32 inverse areas
    area = 1.0 / case.inv_area_cell.astype(np.float64)[None, :]
    s0 = o.get_state()
    rho0 = o.export_field_f64("RHO_ZZ").reshape(s0["rho_zz"].shape)
    q0 = o.export_field_f64("SCALARS").reshape(s0["scalars"].shape)
    m0 = (area * rho0 * _dz(case)).sum()
    t0 = [(area * rho0 * q0[i] * _dz(case)).sum() for i in range(case.n_scalars)]
    assert o.step(1) == 0
    rho1 = o.export_field_f64("RHO_ZZ").reshape(rho0.shape)
    q1 = o.export_field_f64("SCALARS").reshape(q0.shape)
    m1 = (area * rho1 * _dz(case)).sum()
    assert abs(m1 / m0 - 1) < 1e-13
    for i in range(case.n_scalars):
        t1 = (area * rho1 * q1[i] * _dz(case)).sum()
        assert abs(t1 / t0[i] - 1) < 1e-10, i


def test_fct_respects_zalesak_bounds_fp64(get_case):
    case = get_case("x1.642-moist")
    o = Oracle(case, 64)
    o.set_state(initial_state_arrays(case), rebuild=True)
    q0 = o.export_field_f64("SCALARS").reshape(case.n_scalars, case.n_levels, case.mesh.n_cells)
    o.step(1)
    q1 = o.export_field_f64("SCALARS").reshape(q0.shape)
    L = case.n_levels
    coc, n = case.mesh.cellsOnCell, case.mesh.nEdgesOnCell
    for sp in range(case.n_scalars):
        lo = q0[sp].copy()
        hi = q0[sp].copy()
        lo[1:] = np.minimum(lo[1:], q0[sp, :-1]); hi[1:] = np.maximum(hi[1:], q0[sp, :-1])
        lo[:-1] = np.minimum(lo[:-1], q0[sp, 1:]); hi[:-1] = np.maximum(hi[:-1], q0[sp, 1:])
        for j in range(case.mesh.max_edges):
            c = np.flatnonzero(n > j)
            lo[:, c] = np.minimum(lo[:, c], q0[sp][:, coc[c, j]])
            hi[:, c] = np.maximum(hi[:, c], q0[sp][:, coc[c, j]])
        scale = max(np.abs(q0[sp]).max(), 1e-30)
        assert (q1[sp] - hi).max() <= 1e-9 * scale, sp
        assert (lo - q1[sp]).max() <= 1e-9 * scale, sp
        assert q1[sp].min() >= 0.0


def test_thomas_factors_match_textbook_and_solve(get_case):
    case = get_case("x1.642-dry")
    o = Oracle(case, 64)
    o.set_state(initial_state_arrays(case), rebuild=True)
    o.begin_step()
    o.begin_subcycle(1)
    L, C = case.n_levels, case.mesh.n_cells
    a = o.export_field_f64("A_TRI").reshape(L, C)
    al = o.export_field_f64("ALPHA_TRI").reshape(L, C)
    ga = o.export_field_f64("GAMMA_TRI").reshape(L, C)
    b, c = (x.reshape(L, C) for x in o.btri_ctri_f64())
    from hexdyn.setup import rk_schedule
    dts = rk_schedule(case.config.config_dt, acoustic_substeps=case.config.config_number_of_sub_steps,
                      dynamics_splits=case.config.config_dynamics_split_steps)[0].acoustic_timestep
    d2 = dts * dts
    rng = np.random.default_rng(0)
    for cell in (0, 17, C - 1):
        # textbook Thomas on the system d2*a x[k-1] + (1 + d2*b) x[k] + d2*c x[k+1] = r, k=1..L-1
        g_prev = 0.0
        for k in range(1, L):
            alpha = 1.0 / (1.0 + d2 * b[k, cell] - d2 * a[k, cell] * g_prev)
            gamma = d2 * c[k, cell] * alpha
            assert abs(alpha - al[k, cell]) <= 1e-13 * abs(alpha)
            assert abs(gamma - ga[k, cell]) <= 1e-12 * max(abs(gamma), 1e-30)
            g_prev = gamma
        n = L - 1
        M = np.zeros((n, n))
        for i, k in enumerate(range(1, L)):
            M[i, i] = 1.0 + d2 * b[k, cell]
            if i > 0:
                M[i, i - 1] = d2 * a[k, cell]
            if i < n - 1:
                M[i, i + 1] = d2 * c[k, cell]
        r = rng.standard_normal(n)
        x = np.zeros(L + 1)
        x[1:L] = r
        for k in range(1, L):
            x[k] = (x[k] - d2 * a[k, cell] * x[k - 1]) * al[k, cell]
        for k in range(L - 1, -1, -1):
            x[k] = x[k] - ga[k, cell] * x[k + 1]
        assert np.allclose(x[1:L], np.linalg.solve(M, r), rtol=1e-10, atol=1e-12)


def test_flat_balanced_jet_stays_steady(get_case):
    case = get_case("x1.642-dry", terrain_scale=0.0, theta_amplitude=0.0)
    o = Oracle(case, 64)
    o.set_state(initial_state_arrays(case), rebuild=True)
    u0 = o.get_state()["u"]
    assert o.step(4) == 0  # 4 h at dt 3600
    s = o.get_state()
    assert np.abs(s["u"] - u0).max() < 1.0
    assert np.abs(s["w"]).max() < 0.02


def test_fp32_and_fp64_oracles_agree(get_case):
    case = get_case("x1.642-moist")
    o32, o64 = Oracle(case, 32), Oracle(case, 64)
    for o in (o32, o64):
        o.set_state(initial_state_arrays(case), rebuild=True)
        assert o.step(1) == 0
    a, b = o32.get_state(), o64.get_state()
    for name in ("rho_zz", "rtheta_m", "u", "theta_m"):
        rel = np.linalg.norm(a[name] - b[name]) / np.linalg.norm(b[name])
        assert rel < 1e-5, (name, rel)
