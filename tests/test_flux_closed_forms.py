"""The stage tendencies' edge-flux terms and the perturbation-Coriolis term
pinned by closed forms on the FP64 oracle, each shown to fail under a
plausible mistake planted in the oracle (a flipped sign, a wrong neighbour
index, a swapped interpolation weight).

What each check fixes:
* perturbation Coriolis (dynamics_v841.py:128-185): the term equals
  -rho_edge * f_edge * V[u_ref], with V the tangential reconstruction of
  solve_diagnostics (pinned to the reference's bits) applied to the normal
  reference wind u_init*cos(angle) - v_init*sin(angle): the neighbour edge's
  angle, the target edge's f (the reference docstring's subtle point);
* theta edge flux (driver.py:1003-1034): a uniform theta0 gives
  ru * theta0 * dvEdge (the advection weights sum to dvEdge, the
  third-order ones to 0); a smooth field is interpolated to the edge better
  than the two-cell average (the slot pairing of weights and cells);
* theta tendency at stage 2 (driver.py:933-1079 with the RK correction flux,
  cuda_dynamics_v841.py:209-247): with a uniform theta0 it is theta0 times
  the subcycle-start density tendency, which fixes the sign of the
  horizontal and vertical correction terms;
* the third-order upwind blend: its contribution to sum A*theta*tend_rt is
  negative (a fourth-derivative hyperdiffusion), so a flipped blend sign
  turns dissipation into anti-dissipation;
* w edge flux (driver.py:1140-1170): a uniform w0 gives
  (fzm*ru_k + fzp*ru_{k-1}) * w0 * dvEdge;
* w tendency (driver.py:1081-1207, acoustic.py:246-281): with a uniform w0
  it is the omega conversion of w0 times the w-grid mass convergence plus
  the Euler w pool;
* rho*theta is conserved to rounding by a whole step (flux form).
"""
import dataclasses

import numpy as np
import pytest

import oracle.oracle as oo
from hexdyn.case import initial_state_arrays
from oracle.oracle import Oracle

NLEV = 12


@pytest.fixture(scope="module")
def case(get_case):
    return get_case("x1.642-moist", nlev=NLEV)


@pytest.fixture(scope="module")
def developed(case):
    """The published state after one large step (w, rw and every
    perturbation non-zero)."""
    o = Oracle(case, 64)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(1) == 0
    return o.get_state()


def _sizes(case):
    return case.n_levels, case.mesh.n_cells, case.mesh.n_edges


def _f32_as_f64(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def _at_stage2(case, state, plant=0, theta=None, w=None):
    """FP64 oracle from `state`, through stage 1 of the first subcycle; theta
    (both current and subcycle-start) and w replaced when given; then the
    stage-2 tendencies."""
    o = Oracle(case, 64)
    o.set_state(state, rebuild=False)
    o.set_plant(plant)
    o.begin_step()
    o.begin_subcycle(1)
    o.dyn_tend(1)
    o.acoustic_substep(1, 1)
    o.divergence_damping(1)
    assert o.recover(1) == 0
    o.solve_diagnostics(1)
    if theta is not None:
        o.import_internal("c_theta", theta)
        o.import_internal("s_theta", theta)
    if w is not None:
        o.import_internal("w", w)
    o.dyn_tend(2)
    return o


# ------------------------------------------------------------ perturbation Coriolis

def _coriolis_error(case, case_wind, state, plant):
    L, C, E = _sizes(case)
    terms = []
    for c in (case, case_wind):
        o = Oracle(c, 64)
        o.set_state(state, rebuild=False)
        o.set_plant(plant)
        o.begin_step()
        o.begin_subcycle(1)
        terms.append((o.momentum_terms_f64().reshape(L, E), o))
    d = terms[1][0] - terms[0][0]
    o = terms[1][1]
    rho = o.export_internal_f64("c_rho").reshape(L, C)
    c0, c1 = np.asarray(case.mesh.cellsOnEdge).T
    rho_e = 0.5 * (rho[:, c0] + rho[:, c1])
    f = _f32_as_f64(case.mesh.fEdge)
    u_ref = (_f32_as_f64(case_wind.u_init)[:, None] * _f32_as_f64(case_wind.cos_angle_edge)[None]
             - _f32_as_f64(case_wind.v_init)[:, None] * _f32_as_f64(case_wind.sin_angle_edge)[None])
    o.import_field("U", u_ref.astype(np.float32))
    o.solve_diagnostics(3)
    v_ref = o.export_internal_f64("v").reshape(L, E)
    want = -rho_e * f[None] * v_ref
    return np.abs(d - want).max() / np.abs(want).max()


@pytest.mark.parametrize("plant", [0, oo.PLANT_CORIOLIS_TARGET_ANGLE, oo.PLANT_CORIOLIS_NEIGHBOUR_F],
                         ids=["as-built", "planted-target-angle", "planted-neighbour-f"])
def test_perturbation_coriolis_is_minus_f_times_reconstructed_reference_wind(get_case, case, developed, plant):
    case_wind = get_case("x1.642-moist", nlev=NLEV, reference_wind=(12.0, -7.0))
    assert np.any(case_wind.u_init) and np.any(case_wind.v_init)
    err = _coriolis_error(case, case_wind, developed, plant)
    if plant == 0:
        assert err < 1e-6  # the reference wind enters V through one binary32 rounding
    else:
        assert err > 1e-3, "the identity must catch the planted mistake"


# ------------------------------------------------------------ theta fluxes

THETA0 = 300.0


def _flux_scale(case, o, field, theta0):
    """theta0 * max over cells of sum |ru*dv| / area: the size of the flux
    terms whose convergence the tendency is."""
    L, C, E = _sizes(case)
    ru = o.export_internal_f64(field).reshape(L, E)
    dv = _f32_as_f64(case.mesh.dvEdge)
    eoc, n = case.mesh.edgesOnCell, case.mesh.nEdgesOnCell
    s = np.zeros((L, C))
    for j in range(case.mesh.max_edges):
        c = np.flatnonzero(n > j)
        s[:, c] += np.abs(ru[:, eoc[c, j]] * dv[eoc[c, j]])
    return theta0 * (s * case.inv_area_cell.astype(np.float64)[None]).max()


@pytest.mark.parametrize("plant", [0, oo.PLANT_CORRECTION_SIGN, oo.PLANT_VERTICAL_CORRECTION],
                         ids=["as-built", "planted-correction-sign", "planted-vertical-correction"])
def test_uniform_theta_stage2_tendency_is_theta0_times_density_tendency(case, developed, plant):
    L, C, E = _sizes(case)
    th = np.full((L, C), THETA0, np.float32)
    o = _at_stage2(case, developed, plant, theta=th)
    trt = o.export_internal_f64("tend_rt").reshape(L, C)
    trho = o.export_internal_f64("tend_rho").reshape(L, C)
    err = np.abs(trt - THETA0 * trho).max() / _flux_scale(case, o, "s_ru", THETA0)
    if plant == 0:
        assert err < 1e-6  # advection weights rounded to binary32 sum to dvEdge within ~1e-7
    else:
        assert err > 1e-4, "the closed form must catch the planted mistake"


def test_uniform_theta_edge_flux_is_ru_theta0_dv(case, developed):
    L, C, E = _sizes(case)
    o = _at_stage2(case, developed, theta=np.full((L, C), THETA0, np.float32))
    fth = o.export_internal_f64("fth").reshape(L, E)
    want = o.export_internal_f64("c_ru").reshape(L, E) * THETA0 * _f32_as_f64(case.mesh.dvEdge)[None]
    assert np.abs(fth - want).max() / np.abs(want).max() < 1e-6


def _smooth(x, y, z):
    r = np.sqrt(x * x + y * y + z * z)
    return 300.0 + 30.0 * z / r + 20.0 * x / r + 10.0 * (y / r) ** 2


@pytest.mark.parametrize("plant", [0, oo.PLANT_ADV_SLOT], ids=["as-built", "planted-slot-shift"])
def test_smooth_theta_edge_value_beats_two_cell_average(case, developed, plant):
    m = case.mesh
    L, C, E = _sizes(case)
    th = np.ascontiguousarray(np.broadcast_to(_smooth(m.xCell, m.yCell, m.zCell)[None], (L, C)), dtype=np.float32)
    o = _at_stage2(case, developed, plant, theta=th)
    fth = o.export_internal_f64("fth").reshape(L, E)
    ru = o.export_internal_f64("c_ru").reshape(L, E)
    dv = _f32_as_f64(m.dvEdge)
    exact = _smooth(m.xEdge, m.yEdge, m.zEdge)
    ok = np.abs(ru) > 1e-3 * np.abs(ru).max()
    est = fth[ok] / (ru[ok] * np.broadcast_to(dv[None], (L, E))[ok])
    err_high = np.abs(est - np.broadcast_to(exact[None], (L, E))[ok]).max()
    c0, c1 = np.asarray(m.cellsOnEdge).T
    err_two = np.abs(0.5 * (_f32_as_f64(th[0, c0]) + _f32_as_f64(th[0, c1])) - exact).max()
    if plant == 0:
        assert err_high < 0.5 * err_two
    else:
        assert err_high > 2.0 * err_two, "the slot check must catch the planted mistake"


@pytest.mark.parametrize("plant", [0, oo.PLANT_BLEND_SIGN], ids=["as-built", "planted-blend-sign"])
def test_third_order_blend_dissipates_theta_variance(get_case, case, developed, plant):
    """sum A*theta'*(tend_rt(coef) - tend_rt(0)) < 0 for a random theta: the
    upwind blend (horizontal and vertical) is a hyperdiffusion."""
    L, C, E = _sizes(case)
    case0 = get_case("x1.642-moist", nlev=NLEV,
                     config=dataclasses.replace(case.config, config_coef_3rd_order=0.0))
    th = (THETA0 + np.random.default_rng(5).normal(0.0, 1.0, (L, C))).astype(np.float32)
    a = _at_stage2(case, developed, plant, theta=th).export_internal_f64("tend_rt").reshape(L, C)
    b = _at_stage2(case0, developed, plant, theta=th).export_internal_f64("tend_rt").reshape(L, C)
    area = 1.0 / case.inv_area_cell.astype(np.float64)
    dz = 1.0 / _f32_as_f64(case.grid.rdzw)
    thp = _f32_as_f64(th) - THETA0
    wgt = area[None] * dz[:, None]
    s = (wgt * thp * (a - b)).sum()
    scale = (wgt * np.abs(thp * (a - b))).sum()
    if plant == 0:
        assert s < -0.1 * scale
    else:
        assert s > 0.1 * scale, "the dissipation sign must catch the planted mistake"


# ------------------------------------------------------------ w fluxes

W0 = 0.25


@pytest.mark.parametrize("plant", [0, oo.PLANT_W_INTERP], ids=["as-built", "planted-fzm-fzp-swap"])
def test_uniform_w_edge_flux_is_interface_ru_w0_dv(case, developed, plant):
    L, C, E = _sizes(case)
    o = _at_stage2(case, developed, plant, w=np.full((L + 1, C), W0, np.float32))
    fw = o.export_internal_f64("fw").reshape(L + 1, E)
    ru = o.export_internal_f64("c_ru").reshape(L, E)
    fzm, fzp = _f32_as_f64(case.grid.fzm), _f32_as_f64(case.grid.fzp)
    rui = fzm[1:L, None] * ru[1:] + fzp[1:L, None] * ru[:-1]
    want = rui * W0 * _f32_as_f64(case.mesh.dvEdge)[None]
    err = np.abs(fw[1:L] - want).max() / np.abs(want).max()
    if plant == 0:
        assert err < 1e-6
    else:
        assert err > 1e-3, "the closed form must catch the planted mistake"


def test_uniform_w_tendency_is_omega_of_w0_times_mass_convergence(case, developed):
    """tend_rw = omega(w0 * Dw + euler_rw, tend_ru) with Dw the w-grid mass
    convergence: -(1/A) sum sign*dv*(fzm ru_k + fzp ru_{k-1}) - rdzu*(m_{k+1} - m_k),
    m_k = (rw_k + rw_{k-1})/2 at the layer between interfaces k-1 and k
    (0 above the top layer)."""
    L, C, E = _sizes(case)
    o = _at_stage2(case, developed, w=np.full((L + 1, C), W0, np.float32))
    ru = o.export_internal_f64("c_ru").reshape(L, E)
    rw = o.export_internal_f64("c_rw").reshape(L + 1, C)
    fzm, fzp, rdzu = (_f32_as_f64(getattr(case.grid, n)) for n in ("fzm", "fzp", "rdzu"))
    dv = _f32_as_f64(case.mesh.dvEdge)
    ia = case.inv_area_cell.astype(np.float64)
    eoc, n, coe = case.mesh.edgesOnCell, case.mesh.nEdgesOnCell, case.mesh.cellsOnEdge
    rui = np.zeros((L + 1, E))
    rui[1:L] = fzm[1:L, None] * ru[1:] + fzp[1:L, None] * ru[:-1]
    div = np.zeros((L + 1, C))
    for j in range(case.mesh.max_edges):
        c = np.flatnonzero(n > j)
        e = eoc[c, j]
        sg = np.where(coe[e, 0] == c, 1.0, -1.0)
        div[:, c] += sg[None] * dv[e][None] * rui[:, e]
    mlay = np.zeros((L + 1, C))
    mlay[1:L] = 0.5 * (rw[1:L] + rw[0:L - 1])
    dw = np.zeros((L + 1, C))
    for k in range(1, L):
        dw[k] = -div[k] * ia - rdzu[k] * (mlay[k + 1] - mlay[k])
    pre = W0 * dw + o.export_internal_f64("euler_rw").reshape(L + 1, C)
    want = o.omega_conversion(pre.astype(np.float32), o.export_field("TEND_RU")).astype(np.float64)
    got = o.export_internal_f64("tend_rw").reshape(L + 1, C)
    scale = np.abs(want).max()
    assert np.abs(got[1:L] - want[1:L]).max() < 2e-6 * scale


# ------------------------------------------------------------ conservation

@pytest.mark.parametrize("mixing", ["off", "2d_smagorinsky"])
def test_rho_theta_conserved_by_a_step_fp64(get_case, mixing):
    """sum areaCell * dzw * rtheta_m (flux form throughout: edge fluxes,
    vertical fluxes with closed lids, the acoustic implicit update) holds to
    rounding over a whole step."""
    base = get_case("x1.642-moist", nlev=NLEV)
    c = get_case("x1.642-moist", nlev=NLEV, config=dataclasses.replace(base.config, config_horiz_mixing=mixing))
    o = Oracle(c, 64)
    o.set_state(initial_state_arrays(c), rebuild=True)
    L, C, _ = _sizes(c)
    wgt = (1.0 / c.inv_area_cell.astype(np.float64))[None] * (1.0 / _f32_as_f64(c.grid.rdzw))[:, None]
    t0 = (wgt * o.export_field_f64("RTHETA_M").reshape(L, C)).sum()
    assert o.step(1) == 0
    t1 = (wgt * o.export_field_f64("RTHETA_M").reshape(L, C)).sum()
    assert abs(t1 / t0 - 1) < 1e-13


# ------------------------------------------------------------ upper damping

def test_rayleigh_u_damping_rate_ramps_to_one_over_timescale(get_case, case, developed):
    """Rayleigh damping of u over the top n levels (native/upper_damping.rs:18-31):
    with everything else equal, the tendency changes by -rate_k*rho_edge*u
    with rate 1/timescale at the top level falling linearly to
    1/(n*timescale) at the lowest damped level, and not at all below."""
    n, days = 4, 5.0
    cfg = dataclasses.replace(case.config, config_rayleigh_damp_u=True, config_number_rayleigh_damp_u_levels=n,
                              config_rayleigh_damp_u_timescale_days=days)
    damped = get_case("x1.642-moist", nlev=NLEV, config=cfg)
    L, C, E = _sizes(case)
    out = []
    for c in (case, damped):
        o = Oracle(c, 64)
        o.set_state(developed, rebuild=False)
        o.begin_step()
        o.begin_subcycle(1)
        o.dyn_tend(1)
        out.append(o)
    d = out[1].export_internal_f64("tend_ru").reshape(L, E) - out[0].export_internal_f64("tend_ru").reshape(L, E)
    rho = out[0].export_internal_f64("c_rho").reshape(L, C)
    c0, c1 = np.asarray(case.mesh.cellsOnEdge).T
    ru_e = 0.5 * (rho[:, c0] + rho[:, c1]) * out[0].export_internal_f64("u").reshape(L, E)
    tau = days * 86400.0
    assert np.array_equal(d[:L - n], np.zeros_like(d[:L - n]))
    ok = np.abs(ru_e) > 1e-6 * np.abs(ru_e).max()
    rate = -d[ok.nonzero()] / ru_e[ok.nonzero()]
    levels = np.nonzero(ok)[0]
    for k in range(L - n, L):
        rk = rate[levels == k]
        want = (k - (L - n) + 1) / (n * tau)
        assert np.allclose(rk, want, rtol=1e-6, atol=0), k
    assert abs(rate[levels == L - 1].mean() * tau - 1.0) < 1e-6


def test_cam_floor_lifts_kdiff_to_the_floor_profile(get_case):
    """The CAM floor (native/upper_damping.rs:5-16): over the top n levels
    kdiff >= base*(1-(L-1-k)/n), base = 4*2.0833*len*coef, equal to it where
    the Smagorinsky value is smaller, unchanged below the layer."""
    base_case = get_case("x1.642-moist", nlev=NLEV)
    mix = dataclasses.replace(base_case.config, config_horiz_mixing="2d_smagorinsky")
    n, coef = 4, 0.2
    plain = get_case("x1.642-moist", nlev=NLEV, config=mix)
    cam = get_case("x1.642-moist", nlev=NLEV, config=dataclasses.replace(mix, config_mpas_cam_coef=coef,
                                                                        config_number_cam_damping_levels=n))
    L, C, _ = _sizes(plain)
    kd = []
    for c in (plain, cam):
        o = Oracle(c, 64)
        o.set_state(initial_state_arrays(c), rebuild=True)
        o.begin_step()
        o.begin_subcycle(1)
        kd.append(o.export_internal_f64("kdiff").reshape(L, C))
    length = float(np.float32(plain.len_disp))
    base = 4.0 * 2.0833 * length * coef
    assert np.array_equal(kd[0][:L - n], kd[1][:L - n])
    for k in range(L - n, L):
        floor = base * (1.0 - (L - 1 - k) / n)
        assert np.allclose(kd[1][k], np.maximum(kd[0][k], floor), rtol=1e-6), k
        assert (kd[1][k] >= floor * (1 - 1e-6)).all()
    assert abs(kd[1][L - 1].min() / base - 1.0) < 1e-6  # the top level sits on the floor somewhere


def test_clamp_applies_to_passive_tracers_without_moisture(get_case):
    """The end-of-step q < 0 -> 0 clamp covers every species, also when no
    species loads the air (n_moist = 0) and transport is unlimited
    (cuda_driver.py:5306-5311)."""
    base = get_case("x1.642-moist", nlev=NLEV)
    c = dataclasses.replace(base, n_moist=0, config=dataclasses.replace(base.config, config_monotonic=False))
    st = initial_state_arrays(c)
    q = st["scalars"].copy()
    q[1] = (q[1] - np.float32(2e-6)).astype(np.float32)  # negative patches the unlimited transport keeps
    st["scalars"] = q
    o = Oracle(c, 32)
    o.set_state(st, rebuild=True)
    o.set_post_transport_clamp(False)
    assert o.step(1) == 0
    unclamped = o.get_state()["scalars"]
    assert (unclamped < 0).any()
    o2 = Oracle(c, 32)
    o2.set_state(st, rebuild=True)
    assert o2.step(1) == 0
    clamped = o2.get_state()["scalars"]
    assert (clamped >= 0).all()
    assert np.array_equal(clamped, np.where(unclamped < 0, np.float32(0), unclamped))
