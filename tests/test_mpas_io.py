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
