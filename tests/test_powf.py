"""The exner power function (binary64 ln/exp method rounded once) against
glibc powf and against correctly rounded binary64 pow."""
import ctypes
import ctypes.util

import numpy as np

from oracle.oracle import powf as oracle_powf
from hexdyn import lib as hlib

libm = ctypes.CDLL(ctypes.util.find_library("m"))
libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
libm.powf.restype = ctypes.c_float


def _ulps(a, b):
    ia = np.asarray(a, dtype=np.float32).view(np.int32).astype(np.int64)
    ib = np.asarray(b, dtype=np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def test_powf_against_glibc_and_binary64():
    rng = np.random.default_rng(7)
    # exner arguments: zz*R/p0*rtheta over the model's range, and the fixed exponent R/cv
    xs = np.concatenate([rng.uniform(0.005, 1.3, 20000), rng.uniform(1e-6, 50.0, 5000)]).astype(np.float32)
    y = np.float32(287.0 / 717.5)
    L = hlib.lib()
    ours = np.array([L.hexdyn_powf(float(x), float(y)) for x in xs], dtype=np.float32)
    via_oracle = np.array([oracle_powf(float(x), float(y)) for x in xs], dtype=np.float32)
    glibc = np.array([libm.powf(float(x), float(y)) for x in xs], dtype=np.float32)
    cr = np.power(xs.astype(np.float64), np.float64(y)).astype(np.float32)
    assert np.array_equal(ours.view(np.uint32), via_oracle.view(np.uint32))
    assert _ulps(ours, glibc).max() <= 1
    assert (_ulps(ours, glibc) == 0).mean() > 0.999
    assert (_ulps(ours, cr) == 0).mean() > 0.9999


def test_powf_special_values():
    L = hlib.lib()
    assert L.hexdyn_powf(1.0, 0.4) == 1.0
    assert L.hexdyn_powf(0.0, 0.4) == 0.0
    assert np.isnan(L.hexdyn_powf(-1.0, 0.4))
    assert L.hexdyn_powf(2.0, 0.5) == np.float32(np.sqrt(2.0))
    assert L.hexdyn_powf(4.0, 0.5) == 2.0
    y = np.float32(0.4)
    for v in (1e-40, 1e-42, 2e-38):  # subnormal and near-subnormal arguments
        x = np.float32(v)
        cr = np.float32(np.float64(x) ** np.float64(y))
        assert np.float32(L.hexdyn_powf(float(x), float(y))) == cr
