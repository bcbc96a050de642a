"""libhexdyn.so loads and exports every entry point include/hexdyn.h declares."""
import re
from pathlib import Path

import pytest

from hexdyn import lib as hlib

HEADER = Path(__file__).resolve().parent.parent / "include" / "hexdyn.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int64_t|int|void|float|double)\s+\*?(hexdyn_\w+)\s*\(", text, re.M)))


def test_library_exports_header_symbols():
    L = hlib.lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(hlib.EXPORTED)


def test_abi_version():
    assert hlib.lib().hexdyn_abi_version() == hlib.ABI_VERSION == 4


def test_refuses_bad_config_without_gpu_calls(get_case):
    import ctypes as ct
    from hexdyn import capi

    case = get_case("x1.642-dry")
    m, v, c, keep = capi.build_structs(case)
    c.config_dynamics_split_steps = 2
    h = ct.c_void_p()
    assert hlib.lib().hexdyn_create(ct.byref(m), ct.byref(v), ct.byref(c), ct.byref(h)) == capi.HEXDYN_E_REFUSED
    assert not h.value


_KEEP = []


def _bad_levels(m, v, c, keep):
    v.n_levels = 64


def _bad_max_edges(m, v, c, keep):
    m.max_edges = 8


def _empty_mesh(m, v, c, keep):
    m.n_cells = 0


def _moist_over_scalars(m, v, c, keep):
    c.n_moist = c.n_scalars + 1


def _cam_floor_without_levels(m, v, c, keep):
    c.config_horiz_mixing = 1
    c.config_mpas_cam_coef = 0.1
    c.config_number_cam_damping_levels = 0


def _rayleigh_too_deep(m, v, c, keep):
    c.config_rayleigh_damp_u = 1
    c.config_number_rayleigh_damp_u_levels = v.n_levels + 1


def _rayleigh_zero_timescale(m, v, c, keep):
    c.config_rayleigh_damp_u = 1
    c.config_rayleigh_damp_u_timescale_days = 0.0


def _rows_past_32bit_offsets(m, v, c, keep):
    # scalars x edges rows of 256 B reach 4 GiB: 32-bit row offsets would wrap
    c.n_scalars = (1 << 24) // m.n_edges + 1
    c.n_moist = 0


def _cell_index_out_of_range(m, v, c, keep):
    import ctypes as ct
    import numpy as np

    coe = np.ctypeslib.as_array(ct.cast(m.cells_on_edge, ct.POINTER(ct.c_int32)), shape=(2 * m.n_edges,))
    bad = coe.copy()
    bad[7] = m.n_cells  # one past the last cell
    _KEEP.append(bad)  # alive until the create call
    m.cells_on_edge = bad.ctypes.data_as(ct.POINTER(ct.c_int32))


@pytest.mark.parametrize("mutate", [_bad_levels, _bad_max_edges, _empty_mesh, _moist_over_scalars,
                                    _cam_floor_without_levels, _rayleigh_too_deep, _rayleigh_zero_timescale,
                                    _rows_past_32bit_offsets, _cell_index_out_of_range],
                         ids=lambda f: f.__name__.strip("_"))
def test_refuses_inputs_outside_the_contract(get_case, mutate):
    """Sizes and tables outside what the kernels assume are refused at
    create (HEXDYN_E_REFUSED), before any device call: 64 levels (L + 1 must
    fit a 64-lane row), maxEdges 8, an empty mesh, more mass-loading species
    than carried ones, a CAM floor without levels, a Rayleigh layer deeper
    than the column or with no timescale, arrays whose 32-bit row offsets
    would wrap (4 GiB), a connectivity index out of range.  The reason stays
    readable after the refused handle is gone."""
    import ctypes as ct
    from hexdyn import capi

    case = get_case("x1.642-moist")
    m, v, c, keep = capi.build_structs(case)
    mutate(m, v, c, keep)
    h = ct.c_void_p()
    L = hlib.lib()
    assert L.hexdyn_create(ct.byref(m), ct.byref(v), ct.byref(c), ct.byref(h)) == capi.HEXDYN_E_REFUSED
    assert not h.value
    buf = ct.create_string_buffer(512)
    assert L.hexdyn_last_create_error(buf, len(buf)) == 0
    assert buf.value, "the refusal reason must be readable"


def test_refusal_reason_names_the_knob(get_case):
    import ctypes as ct
    from hexdyn import capi

    case = get_case("x1.642-moist")
    m, v, c, keep = capi.build_structs(case)
    _rows_past_32bit_offsets(m, v, c, keep)
    h = ct.c_void_p()
    L = hlib.lib()
    assert L.hexdyn_create(ct.byref(m), ct.byref(v), ct.byref(c), ct.byref(h)) == capi.HEXDYN_E_REFUSED
    buf = ct.create_string_buffer(512)
    L.hexdyn_last_create_error(buf, len(buf))
    assert b"32-bit row offsets" in buf.value


def test_driver_maps_refusal_to_configuration_refusal(get_case, monkeypatch):
    """from_host raises ConfigurationRefusal with the library's reason for a
    refusal (no GPU needed: the refusal comes before any device call)."""
    from dataclasses import replace

    from hexdyn import driver as hdriver
    from hexdyn.config import ConfigurationRefusal

    class _FakeCuda:
        @staticmethod
        def is_available():
            return True

    class _FakeTorch:
        cuda = _FakeCuda()

    monkeypatch.setattr(hdriver, "_torch", lambda: _FakeTorch())
    case = get_case("x1.642-moist")
    bad = replace(case, n_scalars=(1 << 24) // case.mesh.n_edges + 1, n_moist=0)
    with pytest.raises(ConfigurationRefusal, match="32-bit row offsets"):
        hdriver.HexDycoreDriver.from_host(bad)


def _struct_members(name):
    text = HEADER.read_text()
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    members = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        names = decl.split(None, 1)[1] if not decl.startswith("const") else decl.split(None, 2)[2]
        for part in names.split(","):
            members.append(part.strip().lstrip("*").strip())
    return members


def test_ctypes_mirrors_match_header():
    from hexdyn import capi

    for cname, py in (("hexdyn_mesh", capi.HexdynMesh), ("hexdyn_vertical", capi.HexdynVertical),
                      ("hexdyn_config", capi.HexdynConfig), ("hexdyn_state_view", capi.HexdynStateView)):
        assert _struct_members(cname) == [f[0] for f in py._fields_], cname
    enum = re.search(r"HEXDYN_F_RHO_ZZ = 0,(.*?)HEXDYN_F_COUNT", HEADER.read_text(), re.S).group(1)
    assert ["RHO_ZZ"] + re.findall(r"HEXDYN_F_(\w+)", enum) == capi.FIELDS
