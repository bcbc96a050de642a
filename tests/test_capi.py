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
    assert hlib.lib().hexdyn_abi_version() == 1


def test_refuses_bad_config_without_gpu_calls(get_case):
    import ctypes as ct
    from hexdyn import capi

    case = get_case("x1.642-dry")
    m, v, c, keep = capi.build_structs(case)
    c.config_dynamics_split_steps = 2
    h = ct.c_void_p()
    assert hlib.lib().hexdyn_create(ct.byref(m), ct.byref(v), ct.byref(c), ct.byref(h)) == capi.HEXDYN_E_REFUSED
    assert not h.value
