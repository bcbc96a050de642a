"""HIP dycore (libhexdyn.so through its C ABI) against the CPU oracle.

Strict numerics: every phase and every whole step bit for bit equal to the
binary32 restatement.  Fast numerics (FMA contraction): within relative L2
1e-5 of the FP64 restatement on u, w, theta_m and rho_zz after one step.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from hexdyn.case import initial_state_arrays  # noqa: E402
from hexdyn.config import DycoreConfig  # noqa: E402
from hexdyn.setup import rk_schedule  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402


def _driver(case):
    from hexdyn.driver import HexDycoreDriver

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return HexDycoreDriver.from_host(case)


def _bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32).ravel()
    b = np.ascontiguousarray(b, dtype=np.float32).ravel()
    return a.size == b.size and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _diff(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    ne = np.flatnonzero(a.astype(np.float32).view(np.uint32) != b.astype(np.float32).view(np.uint32))
    return f"{ne.size} words differ, max |d| {np.abs(a - b).max():.3e}"


STATE = ["RHO_ZZ", "RTHETA_M", "RU", "RW", "THETA_M", "EXNER", "RHO_P", "RTHETA_P", "PRESSURE_P", "U", "W"]


def _phase_walk(case, physics=None):
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    d = _driver(case)
    if physics is not None:
        o.set_physics(**physics)
        d.set_physics_tendencies(**physics)
    bad = []

    def cmp(tag, fields):
        for f in fields:
            a = o.export_field(f)
            b = d.export_field(f).cpu().numpy()
            if not _bits_equal(a, b):
                bad.append(f"{tag}:{f}: {_diff(a, b)}")

    state = STATE + (["SCALARS"] if case.n_scalars else [])
    mixing = case.config.config_horiz_mixing != "off"
    cmp("init", state)
    o.begin_step(); d.phase("begin_step")
    cmp("begin_step", ["V", "PV_EDGE", "KE", "DIVERGENCE", "VORTICITY"] + (["QTOT", "CQW", "CQU"] if case.n_moist else []))
    cfg = case.config
    sched = rk_schedule(cfg.config_dt, acoustic_substeps=cfg.config_number_of_sub_steps,
                        dynamics_splits=cfg.config_dynamics_split_steps)
    for sub in range(1, cfg.config_dynamics_split_steps + 1):
        o.begin_subcycle(sub); d.phase("begin_subcycle", sub)
        cmp(f"s{sub}", ["A_TRI", "ALPHA_TRI", "GAMMA_TRI", "COFWZ", "COFWT", "COFTZ", "EULER_RU", "EULER_RW",
                        "TEND_RHO"] + (["KDIFF", "EULER_THETA"] if mixing else []))
        for st in (1, 2, 3):
            o.dyn_tend(st); d.phase("dyn_tend", st)
            cmp(f"s{sub}t{st}", ["TEND_RU", "TEND_RT", "TEND_RW"])
            for ss in range(1, sched[st - 1].acoustic_steps + 1):
                o.acoustic_substep(st, ss); d.phase("acoustic_substep", st, ss)
                cmp(f"s{sub}t{st}a{ss}", ["RU_P", "RU_AVG", "RW_P", "WW_AVG", "RHO_PP", "RTHETA_PP"])
                o.divergence_damping(st); d.phase("divergence_damping", st)
                cmp(f"s{sub}t{st}d{ss}", ["RU_P"])
            o.recover(st); d.phase("recover", st)
            cmp(f"s{sub}t{st}r", STATE + ["FLUX_U", "FLUX_W"])
            o.solve_diagnostics(st); d.phase("solve_diagnostics", st)
            cmp(f"s{sub}t{st}g", ["V", "PV_EDGE", "KE", "DIVERGENCE", "VORTICITY"])
        o.end_subcycle(sub); d.phase("end_subcycle", sub)
    if case.n_scalars:
        for st in (1, 2, 3):
            o.scalars(st); d.phase("scalars", st)
            cmp(f"q{st}", ["SCALARS_STAGE"])
    o.end_step(); d.phase("end_step")
    cmp("end", state)
    return bad


@pytest.mark.parametrize("name", ["x1.642-moist", "x1.2562-dry", "r642-moist"])
def test_phase_by_phase_bitwise(get_case, name):
    # r642: ragged mesh with heptagons, which runs the maxEdges-7 kernels
    print(f"[gpu] phase walk {name}", flush=True)
    bad = _phase_walk(get_case(name))
    assert not bad, "\n".join(bad[:20])


def _mixing(get_case, name, **kw):
    import dataclasses

    base = get_case(name)
    return get_case(name, config=dataclasses.replace(base.config, config_horiz_mixing="2d_smagorinsky", **kw))


@pytest.mark.parametrize("name,visc4", [("x1.642-moist", 0.05), ("x1.2562-dry", 0.05), ("x1.2562-dry", 0.0),
                                        ("r642-moist", 0.05)])
def test_phase_by_phase_bitwise_with_mixing(get_case, name, visc4):
    print(f"[gpu] phase walk with 2d_smagorinsky mixing {name} visc4={visc4}", flush=True)
    bad = _phase_walk(_mixing(get_case, name, config_visc4_2dsmag=visc4))
    assert not bad, "\n".join(bad[:20])


def test_ragged_mesh_three_steps_bitwise(get_case):
    """Pentagons, hexagons and heptagons (maxEdges 7): three whole steps on
    the replayed graph bit for bit equal to the oracle."""
    print("[gpu] r642-moist (maxEdges 7) 3 steps", flush=True)
    case = get_case("r642-moist")
    assert case.mesh.max_edges == 7
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(3) == 0
    d = _driver(case)
    d.step_device(3)
    ref, got = o.get_state(), d.get_state()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"


@pytest.mark.parametrize("nlev", [10, 50, 63])
def test_level_counts_two_steps_bitwise(get_case, nlev):
    """A short column, one whose rows (56 floats apart) end in padding
    lanes, and the deepest one a 64-lane row holds (L + 1 = 64, the
    interface row on the last lane), with mixing: 2 steps bit for bit."""
    print(f"[gpu] x1.642-moist at {nlev} levels, mixing on, 2 steps", flush=True)
    import dataclasses

    base = get_case("x1.642-moist", nlev=nlev)
    case = get_case("x1.642-moist", nlev=nlev,
                    config=dataclasses.replace(base.config, config_horiz_mixing="2d_smagorinsky"))
    assert case.n_levels == nlev
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(2) == 0
    d = _driver(case)
    d.step_device(2)
    ref, got = o.get_state(), d.get_state()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"


def test_three_steps_bitwise_moist(get_case):
    print("[gpu] x1.2562-moist 3 steps", flush=True)
    case = get_case("x1.2562-moist")
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(3) == 0
    d = _driver(case)
    d.step_device(3)
    ref, got = o.get_state(), d.get_state()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"


@pytest.mark.parametrize("mixing", ["off", "2d_smagorinsky"])
def test_headline_mesh_step_bitwise_and_conservative(get_case, mixing):
    print(f"[gpu] x1.40962-moist 1 step, mixing {mixing}", flush=True)
    case = get_case("x1.40962-moist") if mixing == "off" else _mixing(get_case, "x1.40962-moist")
    d = _driver(case)
    s0 = d.get_state()
    d.step_device(1)
    got = d.get_state()
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(1) == 0
    ref = o.get_state()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"
    # size-independent properties of the GPU result
    area = 1.0 / case.inv_area_cell.astype(np.float64)[None, :]
    dz = (1.0 / case.grid.rdzw.astype(np.float64))[:, None]
    m0 = (area * s0["rho_zz"].astype(np.float64) * dz).sum()
    m1 = (area * got["rho_zz"].astype(np.float64) * dz).sum()
    assert abs(m1 / m0 - 1) < 1e-6
    assert got["scalars"].min() >= 0.0
    for k, v in got.items():
        assert np.isfinite(v).all(), k


@pytest.mark.parametrize("name", ["x1.10242-moist", "x1.163842-qv"])
def test_other_mesh_sizes_step_bitwise(get_case, name, capsys, monkeypatch):
    """The ensemble mesh (x1.10242, T=6, dt 1440 s) and the scale-up mesh
    (x1.163842, water vapour only, dt 360 s), with 2-D Smagorinsky mixing:
    one large step bit for bit equal to the oracle, then 4 more steps with
    dry and tracer masses held in FP64 integrals.  Building the x1.163842
    case takes a minute or two of host time; its stages are printed."""
    print(f"[gpu] {name} 1 step bitwise + 4 steps of budgets, mixing on", flush=True)
    monkeypatch.setenv("HEXDYN_VERBOSE", "1")
    with capsys.disabled():
        case = _mixing(get_case, name)
    d = _driver(case)
    a = d.integrals()
    d.step_device(1)
    got = d.get_state()
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(1) == 0
    ref = o.get_state()
    del o
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"
    d.step_device(4)
    b = d.integrals()
    assert abs(b["dry_mass"] / a["dry_mass"] - 1) < 1e-9
    for s in range(case.n_scalars):
        assert abs(b[f"tracer_mass_{s}"] / a[f"tracer_mass_{s}"] - 1) < 1e-6, s
    st = d.get_state()
    assert np.isfinite(st["rho_zz"]).all() and (st["scalars"] >= 0).all()


def test_fast_numerics_within_tolerance_of_fp64(get_case):
    """Tolerance: relative L2 <= 1e-5 against FP64 on u, theta_m and rho_zz.
    For w, a small residual of cancelling terms, binary32 evaluated in the
    reference's own order already sits near 1e-4 (the strict path, equal to
    the binary32 restatement bit for bit); fast numerics must be no worse than
    1.5 times that floor."""
    print("[gpu] fast numerics vs FP64", flush=True)
    base = get_case("x1.2562-moist")
    import dataclasses

    fast = dataclasses.replace(base, config=dataclasses.replace(base.config, numerics="fast"))
    d = _driver(fast)
    d.step_device(1)
    got = d.get_state()
    o = Oracle(base, 64)
    o.set_state(initial_state_arrays(base), rebuild=True)
    assert o.step(1) == 0
    ref = o.get_state()
    o32 = Oracle(base, 32)
    o32.set_state(initial_state_arrays(base), rebuild=True)
    assert o32.step(1) == 0
    strict = o32.get_state()

    def rel(x, k):
        return np.linalg.norm(x[k].astype(np.float64) - ref[k]) / np.linalg.norm(ref[k].astype(np.float64))

    for k in ("u", "w", "theta_m", "rho_zz"):
        print(f"  {k}: fast rel L2 vs FP64 = {rel(got, k):.3e}, strict binary32 = {rel(strict, k):.3e}")
    for k in ("u", "theta_m", "rho_zz"):
        assert rel(got, k) <= 1e-5, k
    assert rel(got, "w") <= 1.5 * rel(strict, "w")
    assert not _bits_equal(got["u"], strict["u"])  # the fast path really is a different instruction stream


def test_validation_flag_refuses_and_keeps_state(get_case):
    print("[gpu] validation flag", flush=True)
    case = get_case("x1.642-dry")
    d = _driver(case)
    before = d.get_state()
    # the carried theta_m feeds the theta flux; a NaN there must trip the flag
    bad = before["theta_m"].copy()
    bad[3, 7] = np.nan
    t = torch.from_numpy(bad).cuda()
    from hexdyn import capi
    from hexdyn.lib import lib
    import ctypes as ct

    rc = lib().hexdyn_import_field(d.handle, capi.FIELD_ID["THETA_M"], ct.c_void_p(t.data_ptr()),
                                   ct.c_void_p(d.stream_ptr()))
    assert rc == 0
    with pytest.raises(FloatingPointError):
        d.step_device(1)
    after = d.get_state()
    for k in ("rho_zz", "rtheta_m", "ru", "rw", "exner", "u", "w"):
        assert _bits_equal(after[k], before[k]), k
    assert np.isnan(after["theta_m"][3, 7])
    # the published state is intact: with the carried theta restored the step runs
    tt = torch.from_numpy(before["theta_m"].copy()).cuda()
    assert lib().hexdyn_import_field(d.handle, capi.FIELD_ID["THETA_M"], ct.c_void_p(tt.data_ptr()),
                                     ct.c_void_p(d.stream_ptr())) == 0
    d.step_device(1)


def test_determinism_and_restart(get_case):
    print("[gpu] determinism and restart", flush=True)
    case = get_case("x1.2562-moist")
    a = _driver(case)
    a.step_device(3)
    sa = a.get_state()
    b = _driver(case)
    b.step_device(1)
    mid = b.get_state()
    from hexdyn.driver import HexDycoreDriver

    c = HexDycoreDriver.from_host(case, state=mid, saved=mid)
    c.step_device(2)
    sc = c.get_state()
    b.step_device(2)
    sb = b.get_state()
    for k in sa:
        assert _bits_equal(sa[k], sb[k]), f"determinism {k}"
        assert _bits_equal(sa[k], sc[k]), f"restart {k}"


def test_fp64_integrals_match_host_sums_and_repeat(get_case):
    print("[gpu] FP64 integrals x1.2562-moist", flush=True)
    case = get_case("x1.2562-moist")
    d = _driver(case)
    st = d.get_state()
    a = d.integrals()
    assert a == d.integrals()  # fixed reduction order: identical bits
    m = case.mesh
    area = (1.0 / case.inv_area_cell.astype(np.float64))[None, :]
    dzw = (1.0 / case.grid.rdzw.astype(np.float64))[:, None]
    wgt = area * dzw
    rho = st["rho_zz"].astype(np.float64)
    assert abs(a["dry_mass"] / (wgt * rho).sum() - 1) < 1e-13
    assert abs(a["rho_theta"] / (wgt * st["rtheta_m"].astype(np.float64)).sum() - 1) < 1e-13
    dcdv = (m.dcEdge.astype(np.float32) * m.dvEdge.astype(np.float32)).astype(np.float64)
    u = st["u"].astype(np.float64)
    ke = np.zeros_like(rho)
    for s in range(m.max_edges):
        ok = s < m.nEdgesOnCell
        e = m.edgesOnCell[:, s]
        ke += np.where(ok[None, :], 0.25 * dcdv[e][None, :] * u[:, e] ** 2, 0.0)
    ke *= case.inv_area_cell.astype(np.float64)[None, :]
    qv = st["scalars"][0].astype(np.float64)
    temp = st["theta_m"].astype(np.float64) * st["exner"].astype(np.float64) / (1.0 + (461.6 / 287.0) * qv)
    zg = case.grid.zgrid.astype(np.float32).astype(np.float64)
    zmid = (0.5 * (zg[:-1] + zg[1:])).astype(np.float32).astype(np.float64)
    energy = (wgt * rho * ((1004.5 - 287.0) * temp + 9.80616 * zmid + ke)).sum()
    assert abs(a["total_energy"] / energy - 1) < 1e-12
    for s in range(case.n_scalars):
        ref = (wgt * rho * st["scalars"][s].astype(np.float64)).sum()
        assert abs(a[f"tracer_mass_{s}"] - ref) <= 1e-12 * abs(ref) + 1e-300
    # one step: dry and tracer mass conserved to binary32 rounding of the state
    d.step_device(1)
    b = d.integrals()
    assert abs(b["dry_mass"] / a["dry_mass"] - 1) < 1e-6
    for s in range(case.n_scalars):
        assert abs(b[f"tracer_mass_{s}"] / a[f"tracer_mass_{s}"] - 1) < 1e-5, s


def test_graph_replays_match_eager_steps(get_case):
    # two buffer layouts alternate, so steps 1-2 capture and 3-5 replay; eager
    # host work (state download) runs between replays
    print("[gpu] hipGraph replays vs eager x1.642-moist", flush=True)
    case = get_case("x1.642-moist")
    runs = {}
    for graphs in (True, False):
        d = _driver(case)
        d.use_graphs(graphs)
        states = []
        for _ in range(5):
            d.step_device(1)
            states.append(d.get_state())
        runs[graphs] = states
    for i, (g, e) in enumerate(zip(runs[True], runs[False])):
        for k in g:
            assert _bits_equal(g[k], e[k]), f"step {i + 1} {k}: {_diff(g[k], e[k])}"


def test_graph_cache_captures_once_per_layout(get_case):
    """The dynamics buffers return to their layout every 2 steps (one
    publish swap per step) and the scalar buffers every 3 (three stage swaps
    and a publish swap make a 3-cycle), so the layouts repeat every 6 steps:
    steps 1-6 capture one graph each and steps 7-12 replay them (the cache
    key is stable)."""
    print("[gpu] hipGraph cache x1.642-moist", flush=True)
    d = _driver(get_case("x1.642-moist"))
    d.step_device(6)
    assert d.graph_count() == 6
    d.step_device(6)
    assert d.graph_count() == 6


def test_restart_file_continues_bitwise(get_case, tmp_path):
    print("[gpu] restart file x1.642-moist", flush=True)
    case = get_case("x1.642-moist")
    a = _driver(case)
    a.step_device(3)
    ref = a.get_state()
    b = _driver(case)
    b.step_device(2)
    p = tmp_path / "restart.nc"
    b.write_restart(p, xtime="0000-01-01_00:24:00")
    c = _driver(case)
    meta = c.read_restart(p)
    assert meta["model_time_s"] == 2 * case.config.config_dt
    c.step_device(1)
    got = c.get_state()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"


# Configurations that take the other branches of the fused whole-step
# sequence: more than two acoustic substeps (ru_avg stored and read back), one
# substep everywhere (the average taken from ru_p), one dynamics subcycle,
# no divergence damping, unlimited transport with the final clamp, the RK2
# schedule, and mixing without its fourth-order part.
STEP_VARIANTS = [
    dict(config_number_of_sub_steps=6),
    dict(config_number_of_sub_steps=1),
    dict(config_dynamics_split_steps=1),
    dict(config_divergence_damping=False),
    dict(config_monotonic=False),
    dict(config_time_integration_order=2),
    dict(config_horiz_mixing="2d_smagorinsky", config_visc4_2dsmag=0.0),
    dict(config_rayleigh_damp_u=True, config_number_rayleigh_damp_u_levels=8,
         config_rayleigh_damp_u_timescale_days=0.5),
    dict(config_horiz_mixing="2d_smagorinsky", config_mpas_cam_coef=0.2, config_number_cam_damping_levels=4),
]


def test_partial_mass_loading_bitwise(get_case):
    """physics_mass_loading_count below the carried species count
    (cuda_driver.py:5063-5068): 4 of 6 species load the air, all 6 are
    transported; 2 steps bit for bit, and qtot is the leading-4 sum."""
    import dataclasses

    print("[gpu] x1.642-moist, mass loading by 4 of 6 species, 2 steps", flush=True)
    case = dataclasses.replace(get_case("x1.642-moist"), n_moist=4)
    d = _driver(case)
    d.phase("begin_step")
    q = np.asarray(case.state.scalars, dtype=np.float32)
    t = np.zeros(q.shape[1:], dtype=np.float32)
    for s in range(4):
        t = t + q[s]
    assert _bits_equal(d.export_field("QTOT").cpu().numpy(), t)
    d.close()
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(2) == 0
    d = _driver(case)
    d.step_device(2)
    ref, got = o.get_state(), d.get_state()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"


@pytest.mark.parametrize("variant", STEP_VARIANTS, ids=lambda v: ",".join(f"{k[7:]}={x}" for k, x in v.items()))
def test_whole_step_variants_bitwise(get_case, variant):
    import dataclasses

    print(f"[gpu] x1.642-moist 2 steps with {variant}", flush=True)
    base = get_case("x1.642-moist")
    case = get_case("x1.642-moist", config=dataclasses.replace(base.config, **variant))
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(2) == 0
    d = _driver(case)
    d.step_device(2)
    ref, got = o.get_state(), d.get_state()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"


def test_unlimited_transport_with_held_scalar_tendency_bitwise(get_case):
    """config_monotonic off: all three transport stages take the unlimited
    form, each with the held scalar tendency; 2 whole steps bit for bit."""
    import dataclasses

    print("[gpu] x1.642-moist, unlimited transport with a held scalar tendency, 2 steps", flush=True)
    base = get_case("x1.642-moist")
    case = get_case("x1.642-moist", config=dataclasses.replace(base.config, config_monotonic=False))
    q = np.asarray(case.state.scalars, dtype=np.float32)
    src = ((q * np.random.default_rng(13).uniform(-0.3, 0.3, q.shape).astype(np.float32))
           / np.float32(3600.0)).astype(np.float32)
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    o.set_physics(scalars=src)
    assert o.step(2) == 0
    d = _driver(case)
    d.set_physics_tendencies(scalars=src)
    d.step_device(2)
    ref, got = o.get_state(), d.get_state()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"


def test_post_transport_handoff_bitwise(get_case):
    print("[gpu] unclamped post-transport handoff x1.642-moist, unlimited transport, 2 steps", flush=True)
    import dataclasses

    base = get_case("x1.642-moist")
    case = get_case("x1.642-moist", config=dataclasses.replace(base.config, config_monotonic=False))
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    o.set_post_transport_clamp(False)
    assert o.step(2) == 0
    d = _driver(case)
    d.set_post_transport_clamp(False)
    d.step_device(2)
    ref, got = o.get_state(), d.get_state()
    assert (ref["scalars"] < 0).any()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"


def test_day_long_run_x1_2562_dry_bitwise(get_case):
    """The CPU-runnable workload (x1.2562 dry, dt 1800 s) over 24 h: 48 steps,
    replayed as graphs on the GPU, bit for bit against the oracle."""
    print("[gpu] x1.2562-dry 24 h (48 steps)", flush=True)
    case = get_case("x1.2562-dry")
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(48) == 0
    d = _driver(case)
    d.step_device(48)
    ref, got = o.get_state(), d.get_state()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"


def test_day_long_headline_run_conserves_mass(get_case):
    """x1.40962 moist over 24 h (120 steps of 720 s): the validation flag never
    trips, dry mass and every tracer mass (FCT) hold to the rounding of the
    binary32 state, and no mixing ratio goes negative.  Integrals in FP64."""
    print("[gpu] x1.40962-moist 24 h (120 steps), FP64 budgets", flush=True)
    case = get_case("x1.40962-moist")
    d = _driver(case)
    a = d.integrals()
    d.step_device(120)
    b = d.integrals()
    # measured on MI355X: dry mass -2.1e-11, water vapour -2.6e-10, condensates ~ -6e-8
    assert abs(b["dry_mass"] / a["dry_mass"] - 1) < 1e-9
    for s in range(case.n_scalars):
        assert abs(b[f"tracer_mass_{s}"] / a[f"tracer_mass_{s}"] - 1) < 1e-6, s
    st = d.get_state()
    assert np.isfinite(st["rho_zz"]).all() and (st["scalars"] >= 0).all()


@pytest.mark.parametrize("public", [False, True])
def test_case_read_from_an_mpas_init_file_steps_bitwise(get_case, tmp_path, public):
    """A model-dump file (rho_zz, theta_m, ru, rw) and an init_atmosphere-style
    file (rho, theta, u, w: converted, w -> rw with the terrain terms)."""
    print(f"[gpu] case from an MPAS init file x1.642-moist (public fields: {public}), 2 steps", flush=True)
    from hexdyn import mpas_io

    case = get_case("x1.642-moist")
    p = tmp_path / "x1.642.init.nc"
    w = None
    if public:
        o0 = Oracle(case, 32)
        o0.set_state(initial_state_arrays(case), rebuild=True)
        w = o0.get_state()["w"]
    mpas_io.write_mpas_init(case, p, public_w=w)
    back = mpas_io.case_from_mpas(p, config=case.config, cold_start=public)
    o = Oracle(back, 32)
    o.set_state(initial_state_arrays(back), rebuild=True)
    assert o.step(2) == 0
    d = _driver(back)
    d.step_device(2)
    ref, got = o.get_state(), d.get_state()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"


def test_phase_by_phase_bitwise_with_physics_tendencies(get_case):
    print("[gpu] phase walk with held physics tendencies x1.2562-moist + mixing", flush=True)
    case = _mixing(get_case, "x1.2562-moist")
    L, C, E = case.n_levels, case.mesh.n_cells, case.mesh.n_edges
    rng = np.random.default_rng(11)
    physics = {"rho": (1e-9 * rng.standard_normal((L, C))).astype(np.float32),
               "rho_u": (1e-6 * rng.standard_normal((L, E))).astype(np.float32),
               "rho_theta": (1e-6 * rng.standard_normal((L, C))).astype(np.float32)}
    # a held scalar tendency that pushes some mixing ratios negative before the limiter
    q = np.asarray(case.state.scalars, dtype=np.float32)
    physics["scalars"] = ((q * rng.uniform(-0.5, 0.5, q.shape).astype(np.float32)) / np.float32(3600.0)).astype(
        np.float32)
    bad = _phase_walk(case, physics)
    assert not bad, "\n".join(bad[:20])


def test_phase_by_phase_bitwise_nonzero_reference_wind(get_case):
    """A non-zero u_init/v_init profile: the perturbation-Coriolis term
    (neighbour angle, target f) is live in every stage's momentum tendency."""
    print("[gpu] phase walk x1.642-moist with reference wind (12, -7) m/s", flush=True)
    case = get_case("x1.642-moist", reference_wind=(12.0, -7.0))
    assert np.any(case.u_init) and np.any(case.v_init)
    bad = _phase_walk(case)
    assert not bad, "\n".join(bad[:20])


def test_phase_by_phase_bitwise_upper_damping(get_case):
    """Rayleigh damping of u (8 levels, half a day) and the CAM kdiff floor
    (4 levels) with 2-D Smagorinsky mixing, phase by phase."""
    import dataclasses

    print("[gpu] phase walk x1.642-moist with Rayleigh u damping and the CAM floor", flush=True)
    base = get_case("x1.642-moist")
    cfg = dataclasses.replace(base.config, config_horiz_mixing="2d_smagorinsky", config_mpas_cam_coef=0.2,
                              config_number_cam_damping_levels=4, config_rayleigh_damp_u=True,
                              config_number_rayleigh_damp_u_levels=8, config_rayleigh_damp_u_timescale_days=0.5)
    bad = _phase_walk(get_case("x1.642-moist", config=cfg))
    assert not bad, "\n".join(bad[:20])


def test_variable_mesh_density_mixing_bitwise(get_case):
    """A variable meshDensity in [1, 4]: del2/del4 scaled per edge; 2 steps
    with mixing bit for bit."""
    import copy
    import dataclasses
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
    from reference_walk import mixing_density

    print("[gpu] x1.642-moist variable mesh density, mixing, 2 steps", flush=True)
    base = _mixing(get_case, "x1.642-moist")
    mesh = copy.deepcopy(base.mesh)
    mesh.arrays["meshDensity"] = mixing_density(mesh)
    case = dataclasses.replace(base, mesh=mesh)
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(2) == 0
    d = _driver(case)
    d.step_device(2)
    ref, got = o.get_state(), d.get_state()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"


def test_passive_tracers_clamped_without_moisture_bitwise(get_case):
    """n_moist = 0 with six carried species, unlimited transport and negative
    patches: the end-of-step clamp of every species, 2 steps bit for bit."""
    import dataclasses

    print("[gpu] x1.642 passive tracers, unlimited transport, 2 steps", flush=True)
    base = get_case("x1.642-moist")
    case = dataclasses.replace(base, n_moist=0, config=dataclasses.replace(base.config, config_monotonic=False))
    st = initial_state_arrays(case)
    q = st["scalars"].copy()
    q[1] = (q[1] - np.float32(2e-6)).astype(np.float32)
    st["scalars"] = q
    o = Oracle(case, 32)
    o.set_state(st, rebuild=True)
    assert o.step(2) == 0
    from hexdyn.driver import HexDycoreDriver

    d = HexDycoreDriver.from_host(case, state=st)
    d.step_device(2)
    ref, got = o.get_state(), d.get_state()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"
    assert (got["scalars"] >= 0).all()


def test_cells_on_cell_out_of_edge_order_bitwise(get_case):
    """cellsOnCell rotated one slot against edgesOnCell in every cell (a
    table order MPAS files do not use, so slot s no longer names the cell
    across edge s): the FCT cell pass then reads each edge's other cell on
    its own, and the limiter bounds keep the given order; 2 moist steps bit
    for bit."""
    import copy
    import dataclasses

    print("[gpu] x1.642-moist cellsOnCell rotated one slot, 2 steps", flush=True)
    base = get_case("x1.642-moist")
    assert base.config.config_monotonic
    mesh = copy.deepcopy(base.mesh)
    coc = mesh.arrays["cellsOnCell"].copy()
    n = mesh.arrays["nEdgesOnCell"]
    for c in range(coc.shape[0]):
        coc[c, : n[c]] = np.roll(coc[c, : n[c]], 1)
    mesh.arrays["cellsOnCell"] = coc
    case = dataclasses.replace(base, mesh=mesh)
    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    assert o.step(2) == 0
    d = _driver(case)
    d.step_device(2)
    ref, got = o.get_state(), d.get_state()
    for k in ref:
        assert _bits_equal(ref[k], got[k]), f"{k}: {_diff(ref[k], got[k])}"
