"""MPAS grid files and dycore restart files: round trips through the
netCDF form, index conventions, unit-sphere scaling, bitwise continuation."""
import numpy as np
import pytest

from hexdyn import mpas_io
from hexdyn.case import initial_state_arrays
from hexdyn.mesh import icosahedral_mesh


@pytest.fixture(scope="module")
def mesh():
    return icosahedral_mesh(3)


def test_grid_round_trip_preserves_tables(mesh, tmp_path):
    p = tmp_path / "x1.642.grid.nc"
    mpas_io.write_mpas_grid(mesh, p)
    back = mpas_io.mesh_from_mpas(p)
    for name in mpas_io.CONNECTIVITY + mpas_io.COUNTS:
        a, b = mesh.arrays[name], back.arrays[name]
        assert np.array_equal(np.where(a >= 0, a, -1), b), name
    for name in ("areaCell", "dcEdge", "dvEdge", "weightsOnEdge", "angleEdge", "kiteAreasOnVertex"):
        assert np.array_equal(mesh.arrays[name], back.arrays[name]), name
    assert np.array_equal(mesh.kiteForCell, back.kiteForCell)
    assert back.sphere_radius == mesh.sphere_radius


def test_grid_file_uses_one_based_indices(mesh, tmp_path):
    from scipy.io import netcdf_file

    p = tmp_path / "g.nc"
    mpas_io.write_mpas_grid(mesh, p)
    with netcdf_file(str(p), "r", mmap=False) as f:
        coe = np.array(f.variables["cellsOnEdge"].data)
        eoc = np.array(f.variables["edgesOnCell"].data)
    assert coe.min() == 1 and coe.max() == mesh.n_cells
    # pentagon rows pad with 0 in the file and -1 in memory
    assert (eoc == 0).sum() == (mesh.edgesOnCell < 0).sum() > 0


def test_unit_sphere_grid_is_scaled_to_earth(mesh, tmp_path):
    p = tmp_path / "unit.nc"
    mpas_io.write_mpas_grid(mesh, p, unit_sphere=True)
    back = mpas_io.mesh_from_mpas(p)
    assert back.sphere_radius == pytest.approx(mesh.sphere_radius)
    for name in ("dcEdge", "xCell"):
        assert np.allclose(back.arrays[name], mesh.arrays[name], rtol=1e-14, atol=1e-6)
    assert np.allclose(back.areaCell, mesh.areaCell, rtol=1e-14)


def test_hdf5_files_are_refused(tmp_path):
    p = tmp_path / "h5.nc"
    p.write_bytes(mpas_io.HDF5_MAGIC + b"\0" * 64)
    with pytest.raises(mpas_io.MpasFileError):
        mpas_io.mesh_from_mpas(p)


def test_restart_continues_bitwise_with_the_oracle(get_case, tmp_path):
    from oracle.oracle import Oracle

    case = get_case("x1.642-moist")
    run = Oracle(case, 32)
    run.set_state(initial_state_arrays(case), rebuild=True)
    assert run.step(3) == 0
    ref = run.get_state()

    first = Oracle(case, 32)
    first.set_state(initial_state_arrays(case), rebuild=True)
    assert first.step(2) == 0
    p = tmp_path / "restart.nc"
    mpas_io.write_restart(p, first.get_state(), xtime="2026-01-01_00:48:00",
                          model_time_s=first.model_time())
    state, meta = mpas_io.read_restart(p)
    assert meta["xtime"] == "2026-01-01_00:48:00"
    assert meta["model_time_s"] == pytest.approx(2 * case.config.config_dt)
    second = Oracle(case, 32)
    second.set_state(state, rebuild=False)
    assert second.step(1) == 0
    got = second.get_state()
    for k in ref:
        assert np.array_equal(ref[k].view(np.uint32), got[k].view(np.uint32)), k


def _same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f" and a.dtype == b.dtype:
        a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
        return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))
    return np.array_equal(a, b)


def test_init_file_round_trip_rebuilds_the_case_bitwise(get_case, tmp_path):
    case = get_case("x1.642-moist")
    p = tmp_path / "x1.642.init.nc"
    mpas_io.write_mpas_init(case, p)
    back = mpas_io.case_from_mpas(p, config=case.config)
    g, h = case.grid, back.grid
    for name in ("zgrid", "zz", "zxu", "rdzw", "rdzu", "fzm", "fzp", "zb_cell", "zb3_cell", "zb_edge", "zb3_edge",
                 "dzu"):
        assert _same_bits(getattr(g, name), getattr(h, name)), name
    assert (g.cf1, g.cf2, g.cf3) == (h.cf1, h.cf2, h.cf3)
    for name in ("rho_base", "rtheta_base", "exner_base", "theta_base"):
        assert _same_bits(getattr(case.reference, name), getattr(back.reference, name)), name
    for name in ("rho_zz", "rtheta_m", "ru", "rw", "scalars"):
        assert _same_bits(getattr(case.state, name), getattr(back.state, name)), name
    for name in ("adv_coefs", "adv_coefs_3rd", "n_adv", "adv_cells", "inv_area_cell", "inv_dc_edge", "etp", "ewm",
                 "dss", "u_init", "cos_angle_edge"):
        assert _same_bits(getattr(case, name), getattr(back, name)), name
    assert (back.n_scalars, back.n_moist) == (case.n_scalars, case.n_moist)


def test_init_file_zb3_is_scaled_once_in_binary32(get_case, tmp_path):
    case = get_case("x1.642-moist")
    p = tmp_path / "i.nc"
    mpas_io.write_mpas_init(case, p)
    from scipy.io import netcdf_file

    with netcdf_file(str(p), "r", mmap=False) as f:
        zb3 = np.array(f.variables["zb3"].data)          # (nEdges, TWO, nVertLevelsP1), unscaled
        coe = np.array(f.variables["cellsOnEdge"].data) - 1
    m = case.mesh
    c, s = 7, 2
    e = m.edgesOnCell[c, s]
    side = 0 if coe[e, 0] == c else 1
    want = np.float32(case.config.config_coef_3rd_order) * zb3[e, side, :].astype(np.float32)
    assert _same_bits(case.grid.zb3_cell[:, c, s], want)


def test_init_file_steps_like_the_case_on_the_oracle(get_case, tmp_path):
    from oracle.oracle import Oracle

    case = get_case("x1.642-moist")
    p = tmp_path / "x1.642.init.nc"
    mpas_io.write_mpas_init(case, p)
    back = mpas_io.case_from_mpas(p, config=case.config)
    a, b = Oracle(case, 32), Oracle(back, 32)
    a.set_state(initial_state_arrays(case), rebuild=True)
    b.set_state(initial_state_arrays(back), rebuild=True)
    assert a.step(1) == 0 and b.step(1) == 0
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert _same_bits(sa[k], sb[k]), k


def test_init_file_without_rw_or_metrics_is_refused(get_case, tmp_path):
    from scipy.io import netcdf_file

    case = get_case("x1.642-moist")
    p = tmp_path / "i.nc"
    mpas_io.write_mpas_grid(case.mesh, p)
    with pytest.raises(mpas_io.MpasFileError, match="vertical"):
        mpas_io.case_from_mpas(p, config=case.config)


@pytest.mark.parametrize("unit_sphere", [False, True])
def test_file_deriv_two_builds_the_advection_stencils(get_case, tmp_path, unit_sphere):
    """A file that carries deriv_two (nEdges, TWO, FIFTEEN) feeds it to the
    advection-coefficient compression, as the reference builds its stencils
    from the static file's deriv_two; a file without it gets this package's
    fit.  The same values give the same stencils bit for bit; perturbed
    values give different ones, so the file's copy is the one used."""
    import dataclasses

    from hexdyn import setup as hsetup
    from hexdyn.mesh import Mesh

    case = get_case("x1.642-moist")
    d2 = np.zeros((case.mesh.n_edges, 2, 15))
    d2[:, :, :case.mesh.max_edges + 1] = hsetup.deriv_two(case.mesh)
    for scale, same in ((1.0, True), (1.0 + 2.0 ** -10, False)):
        mesh = Mesh(arrays={**case.mesh.arrays, "deriv_two": d2 * scale}, sphere_radius=case.mesh.sphere_radius)
        p = tmp_path / f"x1.642.{scale}.init.nc"
        mpas_io.write_mpas_init(dataclasses.replace(case, mesh=mesh), p)
        if unit_sphere:
            g = tmp_path / f"x1.642.{scale}.grid.nc"
            mpas_io.write_mpas_grid(mesh, g, unit_sphere=True)
            back = mpas_io.case_from_mpas(p, grid_path=g, config=case.config)
        else:
            back = mpas_io.case_from_mpas(p, config=case.config)
        assert "deriv_two" in back.mesh.arrays
        for name in ("adv_coefs", "adv_coefs_3rd"):
            assert _same_bits(getattr(case, name), getattr(back, name)) == same, (scale, name)


def _recovered_w(case):
    from oracle.oracle import Oracle
    from hexdyn.case import initial_state_arrays

    o = Oracle(case, 32)
    o.set_state(initial_state_arrays(case), rebuild=True)
    return o.get_state()["w"]


def test_w_to_rw_inverts_the_recovery_of_w(get_case):
    """rw -> w (the oracle's recovery, terrain terms included) -> rw again
    returns the case's rw to binary32 rounding; a flipped terrain term does
    not (the terrain slopes of the random 0-4 km orography matter)."""
    case = get_case("x1.642-moist")
    w = _recovered_w(case)
    st = case.state
    back = mpas_io.w_to_rw(case.mesh, case.grid, st.rho_zz, st.ru, w)
    scale = np.abs(st.rw).max()
    err = np.abs(back.astype(np.float64) - st.rw).max()
    assert err <= 1e-5 * scale, (err, scale)
    assert np.all(back[0] == 0) and np.all(back[-1] == 0)
    # planted mistake: the terrain correction added instead of subtracted
    import dataclasses

    flipped = dataclasses.replace(case.grid, zb_cell=-case.grid.zb_cell, zb3_cell=-case.grid.zb3_cell)
    wrong = mpas_io.w_to_rw(case.mesh, flipped, st.rho_zz, st.ru, w)
    assert np.abs(wrong.astype(np.float64) - st.rw).max() > 100 * err


def test_public_init_file_converts_rho_theta_u_w(get_case, tmp_path):
    """An init_atmosphere-style file (physical rho, dry theta, u, w; no
    rho_zz, theta_m, ru, rw) reads back to the case's prognostic state to
    rounding, and the case read from it steps on the oracle."""
    from oracle.oracle import Oracle

    case = get_case("x1.642-moist")
    p = tmp_path / "public.init.nc"
    mpas_io.write_mpas_init(case, p, public_w=_recovered_w(case))
    from scipy.io import netcdf_file

    with netcdf_file(str(p), "r", mmap=False) as f:
        names = set(f.variables)
    assert {"rho", "theta", "u", "w"} <= names and not names & {"rho_zz", "theta_m", "rtheta_m", "ru", "rw"}
    back = mpas_io.case_from_mpas(p, config=case.config)
    a, b = case.state, back.state

    def rel(x, y):
        return float(np.abs(x.astype(np.float64) - y).max() / np.abs(y).max())

    assert rel(b.rho_zz, a.rho_zz) <= 3e-7
    assert rel(b.rtheta_m, a.rtheta_m) <= 1e-6
    assert rel(b.ru, a.ru) <= 1e-6
    assert rel(b.rw, a.rw) <= 1e-5
    o = Oracle(back, 32)
    from hexdyn.case import initial_state_arrays

    o.set_state(initial_state_arrays(back), rebuild=True)
    assert o.step(1) == 0


def test_cold_start_prefers_public_fields(get_case, tmp_path):
    """With cold_start the public rho/theta/u/w rebuild the coupled state even
    where the file also holds rho_zz/theta_m/ru/rw (here scaled copies)."""
    from scipy.io import netcdf_file

    case = get_case("x1.642-moist")
    p = tmp_path / "both.init.nc"
    mpas_io.write_mpas_init(case, p, public_w=_recovered_w(case))
    with netcdf_file(str(p), "a") as f:
        L, C = case.state.rho_zz.shape
        for name, a, dims in (("rho_zz", case.state.rho_zz * 2, ("Time", "nCells", "nVertLevels")),
                              ("theta_m", case.state.rtheta_m / case.state.rho_zz * 2,
                               ("Time", "nCells", "nVertLevels"))):
            v = f.createVariable(name, "f4", dims)
            v[0] = np.asarray(a, np.float32).T
    warm = mpas_io.case_from_mpas(p, config=case.config)
    cold = mpas_io.case_from_mpas(p, config=case.config, cold_start=True)
    assert np.allclose(warm.state.rho_zz, 2 * case.state.rho_zz, rtol=1e-6)
    assert np.allclose(cold.state.rho_zz, case.state.rho_zz, rtol=1e-6)
