import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "gpuwm-hex_amd"))
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libhexdyn.so")


_CASES = {}


@pytest.fixture(scope="session")
def get_case():
    from hexdyn.case import build_case

    def get(name, config=None, **kw):
        key = (name, repr(config), tuple(sorted(kw.items())))
        if key not in _CASES:
            _CASES[key] = build_case(name, config=config, **kw)
        return _CASES[key]

    return get
