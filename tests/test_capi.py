"""libhexdyn.so loads and exports every entry point include/hexdyn.h declares."""
import re
from pathlib import Path

from hexdyn import lib as hlib

HEADER = Path(__file__).resolve().parent.parent / "include" / "hexdyn.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|float|double)\s+\*?(hexdyn_\w+)\s*\(", text, re.M)))


def test_library_exports_header_symbols():
    L = hlib.lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(hlib.EXPORTED)


def test_abi_version():
    assert hlib.lib().hexdyn_abi_version() == hlib.ABI_VERSION == 3


def test_refuses_bad_config_without_gpu_calls(get_case):
    import ctypes as ct
    from hexdyn import capi

    case = get_case("x1.642-dry")
    m, v, c, keep = capi.build_structs(case)
    c.config_dynamics_split_steps = 2
    h = ct.c_void_p()
    assert hlib.lib().hexdyn_create(ct.byref(m), ct.byref(v), ct.byref(c), ct.byref(h)) == capi.HEXDYN_E_REFUSED
    assert not h.value


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
