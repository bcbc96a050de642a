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


@pytest.mark.parametrize("name", ["x1.642-dry", "r642-moist"])
def test_rest_state_is_exact_fixed_point(get_case, name):
    # r642: the ragged mesh of pentagons, hexagons and heptagons (maxEdges 7)
    case = _rest_case(get_case(name))
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


@pytest.mark.parametrize("name", ["x1.642-moist", "r642-moist"])
def test_dry_and_tracer_mass_conserved_fp64(get_case, name):
    case = get_case(name)
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


# ----------------------------------------------------------- horizontal mixing
def _mixing_config(case, **kw):
    return dataclasses.replace(case.config, config_horiz_mixing="2d_smagorinsky", **kw)


def test_mixing_rest_state_is_exact_fixed_point_without_hyperdiffusion(get_case):
    # at rest u = v = 0, so the Smagorinsky viscosity, the del2 filters and
    # (visc4 = 0) the del4 filters all vanish exactly
    base = get_case("x1.642-dry")
    case = _rest_case(get_case("x1.642-dry", config=_mixing_config(base, config_visc4_2dsmag=0.0)))
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    before = o.get_state()
    assert o.step(1) == 0
    after = o.get_state()
    for name in ("rho_zz", "rtheta_m", "rw", "exner", "pressure_p", "w"):
        assert np.array_equal(before[name], after[name]), name
    assert np.abs(after["u"]).max() == 0.0


def test_mixing_theta_filter_is_in_flux_form_fp64(get_case):
    # the del2 and del4 theta tendencies are divergences of edge fluxes:
    # their area-weighted sum over the sphere vanishes at every level
    base = get_case("x1.2562-dry")
    case = get_case("x1.2562-dry", config=_mixing_config(base))
    o = Oracle(case, 64)
    o.set_state(initial_state_arrays(case), rebuild=True)
    o.begin_step()
    o.begin_subcycle(1)
    L, C = case.n_levels, case.mesh.n_cells
    tend = o.export_field_f64("EULER_THETA").reshape(L, C)
    area = case.mesh.areaCell.astype(np.float32).astype(np.float64)[None, :]
    total = (area * tend).sum(axis=1)
    scale = (area * np.abs(tend)).sum(axis=1)
    assert np.all(scale > 0.0)
    assert np.all(np.abs(total) <= 1e-12 * scale)


@pytest.mark.parametrize("flow,strain", [("shear_y", 1.0), ("shear_xy", 2.0), ("stretch_x", 1.0),
                                         ("rotation", 0.0)])
def test_mixing_viscosity_of_linear_winds_fp64(get_case, flow, strain):
    # linear winds near (lat, lon) = (0, 0) with x east and y north:
    # kdiff = (c_s*len)^2 * sqrt(0.25*(D11-D22)^2 + D12^2), D11 = 2 du/dx,
    # D22 = 2 dv/dy, D12 = du/dy + dv/dx (mpas_atm_dissipation_models.F smagorinsky_2d)
    base = get_case("x1.2562-dry")
    case = _rest_case(get_case("x1.2562-dry", config=_mixing_config(base)))
    m = case.mesh
    R = m.sphere_radius
    a = 2e-6
    y = R * np.arcsin(m.zEdge / R)
    x = R * np.arctan2(m.yEdge, m.xEdge)
    east, north = {"shear_y": (a * y, 0 * y), "shear_xy": (a * y, a * x), "stretch_x": (a * x, 0 * y),
                   "rotation": (a * y, -a * x)}[flow]
    u = (east * np.cos(m.angleEdge) + north * np.sin(m.angleEdge)).astype(np.float32)
    o = Oracle(case, 64)
    o.set_state(initial_state_arrays(case), rebuild=True)
    o.import_field("U", np.broadcast_to(u, (case.n_levels, m.n_edges)))
    o.begin_step()
    o.begin_subcycle(1)
    kd = o.export_field_f64("KDIFF").reshape(case.n_levels, m.n_cells)
    near = (np.abs(np.arcsin(m.zCell / R)) < 0.15) & (np.abs(np.arctan2(m.yCell, m.xCell)) < 0.15)
    assert near.sum() >= 10
    scale = (0.125 * np.float32(case.len_disp)) ** 2 * a
    err = np.median(np.abs(kd[:, near] / scale - strain))
    assert err < 0.05 * max(strain, 1.0), (flow, err)
    assert np.all(kd <= (0.01 * np.float32(case.len_disp) ** 2) / case.config.config_dt * (1 + 1e-6))


@pytest.mark.parametrize("mixing", ["off", "2d_smagorinsky"])
def test_mass_conserved_with_and_without_mixing_fp64(get_case, mixing):
    base = get_case("x1.642-moist")
    case = get_case("x1.642-moist", config=dataclasses.replace(base.config, config_horiz_mixing=mixing))
    o = Oracle(case, 64)
    o.set_state(initial_state_arrays(case), rebuild=True)
    area = 1.0 / case.inv_area_cell.astype(np.float64)[None, :]
    shape = (case.n_levels, case.mesh.n_cells)
    m0 = (area * o.export_field_f64("RHO_ZZ").reshape(shape) * _dz(case)).sum()
    assert o.step(1) == 0
    m1 = (area * o.export_field_f64("RHO_ZZ").reshape(shape) * _dz(case)).sum()
    assert abs(m1 / m0 - 1) < 1e-13


def test_mixing_theta_del2_is_dissipative_fp64(get_case):
    # sum_c A_c theta_c tend_c = -sum_e (dv/dc) rho_e K_e (theta1 - theta0)^2 <= 0 for
    # the second-order filter (visc4 = 0): a sign or index error in any
    # flux term makes it positive somewhere
    base = get_case("x1.2562-dry")
    case = get_case("x1.2562-dry", config=_mixing_config(base, config_visc4_2dsmag=0.0))
    o = Oracle(case, 64)
    o.set_state(initial_state_arrays(case), rebuild=True)
    o.begin_step()
    o.begin_subcycle(1)
    L, C = case.n_levels, case.mesh.n_cells
    tend = o.export_field_f64("EULER_THETA").reshape(L, C)
    theta = o.export_field_f64("THETA_M").reshape(L, C)
    area = case.mesh.areaCell.astype(np.float32).astype(np.float64)[None, :]
    diss = (area * theta * tend).sum(axis=1)
    assert np.all(diss < 0.0)
    # and the filter is active: each level loses a visible amount
    assert np.all(np.abs(diss) > 1e-12 * (area * theta * np.abs(tend)).sum(axis=1))


# ---------------------------------------------------------- physics seam
def test_physics_tendencies_enter_the_budgets_fp64(get_case):
    # a held density and rho*theta tendency adds dt * sum(A dz tendency) to the
    # global dry mass and rho*theta integrals (the dynamics conserve both);
    # a held rho_u tendency moves the momentum but no mass
    case = get_case("x1.642-dry")
    L, C, E = case.n_levels, case.mesh.n_cells, case.mesh.n_edges
    area = (1.0 / case.inv_area_cell.astype(np.float64))[None, :]
    dz = _dz(case)
    rng = np.random.default_rng(7)
    prho = (1e-9 * rng.standard_normal((L, C))).astype(np.float32)
    prt = (1e-6 * rng.standard_normal((L, C))).astype(np.float32)
    pru = (1e-6 * rng.standard_normal((L, E))).astype(np.float32)
    shape = (L, C)

    def totals(o):
        return ((area * o.export_field_f64("RHO_ZZ").reshape(shape) * dz).sum(),
                (area * o.export_field_f64("RTHETA_M").reshape(shape) * dz).sum())

    base = Oracle(case, 64)
    base.set_state(initial_state_arrays(case), rebuild=True)
    m0, t0 = totals(base)
    assert base.step(1) == 0
    mb, tb = totals(base)
    phys = Oracle(case, 64)
    phys.set_state(initial_state_arrays(case), rebuild=True)
    phys.set_physics(rho=prho, rho_u=pru, rho_theta=prt)
    assert phys.step(1) == 0
    mp, tp = totals(phys)
    dt = case.config.config_dt
    assert abs((mp - mb) - dt * (area * prho * dz).sum()) <= 1e-9 * abs(dt * (area * np.abs(prho) * dz).sum())
    assert abs((tp - tb) - dt * (area * prt * dz).sum()) <= 1e-9 * abs(dt * (area * np.abs(prt) * dz).sum())
    # the momentum tendency changes the flow but not the mass budget
    mom = Oracle(case, 64)
    mom.set_state(initial_state_arrays(case), rebuild=True)
    mom.set_physics(rho_u=pru)
    assert mom.step(1) == 0
    mm, _ = totals(mom)
    assert abs(mm / mb - 1) < 1e-13
    assert np.abs(mom.export_field_f64("RU") - base.export_field_f64("RU")).max() > 0


def test_post_transport_handoff_skips_the_clamp(get_case):
    """With unlimited transport some mixing ratios go negative; the default
    end-of-step clamp (atm_srk3's DO_PHYSICS branch) zeroes them, and the
    post-transport handoff leaves them for the caller's microphysics
    (cuda_driver.py:5299-5311).  Everything else is identical."""
    import dataclasses

    base = get_case("x1.642-moist")
    case = get_case("x1.642-moist", config=dataclasses.replace(base.config, config_monotonic=False))
    runs = {}
    for clamp in (True, False):
        o = Oracle(case, 32)
        o.set_state(initial_state_arrays(case), rebuild=True)
        o.set_post_transport_clamp(clamp)
        assert o.step(1) == 0
        runs[clamp] = o.get_state()
    q_on, q_off = runs[True]["scalars"], runs[False]["scalars"]
    assert (q_off < 0).any(), "unlimited transport produced no negative mixing ratio to hand over"
    assert not (q_on < 0).any()
    assert np.array_equal(np.where(q_off < 0, np.float32(0), q_off).view(np.uint32), q_on.view(np.uint32))
    for k in ("rho_zz", "rtheta_m", "ru", "rw"):
        assert np.array_equal(runs[True][k].view(np.uint32), runs[False][k].view(np.uint32)), k


def test_mass_loading_counts_only_the_leading_species(get_case):
    """physics_mass_loading_count (cuda_driver.py:5063-5068): qtot sums the
    leading n_moist carried species in species order; cqw and cqu follow from
    it in closed form (the moist-coefficient kernels, cuda_driver.py:1522-1634).
    Species after the count are carried but weigh nothing."""
    base = get_case("x1.642-moist")
    case = dataclasses.replace(base, n_moist=4)
    L, C = case.n_levels, case.mesh.n_cells
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    o.begin_step()
    q = np.asarray(case.state.scalars, dtype=np.float32)
    t = np.zeros((L, C), dtype=np.float32)
    for s in range(4):
        t = t + q[s]
    assert np.array_equal(o.export_field("QTOT").reshape(L, C).view(np.uint32), t.view(np.uint32))
    cqw = o.export_field("CQW").reshape(L, C)
    half = np.float32(0.5)
    want = np.float32(1.0) / (np.float32(1.0) + half * (t[1:] + t[:-1]))
    assert np.array_equal(cqw[1:].view(np.uint32), want.view(np.uint32))
    c0, c1 = case.mesh.cellsOnEdge[:, 0], case.mesh.cellsOnEdge[:, 1]
    te = np.zeros((L, case.mesh.n_edges), dtype=np.float32)
    for s in range(4):
        te = te + half * (q[s][:, c0] + q[s][:, c1])
    cqu = o.export_field("CQU").reshape(L, -1)
    assert np.array_equal(cqu.view(np.uint32), (np.float32(1.0) / (np.float32(1.0) + te)).view(np.uint32))
    # the full count differs: the last two species are not zero here
    full = Oracle(base, 32)
    full.set_state(initial_state_arrays(base), rebuild=True)
    full.begin_step()
    assert not np.array_equal(full.export_field("QTOT"), o.export_field("QTOT"))


def test_held_scalar_tendency_enters_the_tracer_budget_fp64(get_case):
    """A held scalar tendency (the transport's scalar_tendency) adds
    dt * sum(A dz rho_old-weighted source / rho_old) = dt * sum(A dz source)
    to each tracer mass through the limited transport, where no mixing
    ratio is clipped: the transport itself conserves tracer mass."""
    case = get_case("x1.642-moist")
    L, C = case.n_levels, case.mesh.n_cells
    area = (1.0 / case.inv_area_cell.astype(np.float64))[None, :]
    dz = _dz(case)
    q0 = np.asarray(case.state.scalars, dtype=np.float32)
    src = (q0 * np.random.default_rng(5).uniform(0.0, 0.2, q0.shape).astype(np.float32)
           / np.float32(3600.0)).astype(np.float32)

    def masses(o):
        rho = o.export_field_f64("RHO_ZZ").reshape(L, C)
        q = o.export_field_f64("SCALARS").reshape(case.n_scalars, L, C)
        return np.array([(area * rho * q[s] * dz).sum() for s in range(case.n_scalars)])

    base = Oracle(case, 64)
    base.set_state(initial_state_arrays(case), rebuild=True)
    assert base.step(1) == 0
    mb = masses(base)
    phys = Oracle(case, 64)
    phys.set_state(initial_state_arrays(case), rebuild=True)
    phys.set_physics(scalars=src)
    assert phys.step(1) == 0
    mp = masses(phys)
    dt = case.config.config_dt
    added = np.array([dt * (area * src[s].astype(np.float64) * dz).sum() for s in range(case.n_scalars)])
    for s in range(case.n_scalars):
        if added[s] == 0:
            assert mp[s] == mb[s]
            continue
        assert abs((mp[s] - mb[s]) - added[s]) <= 1e-9 * abs(added[s]) + 1e-12 * abs(mb[s]), s
