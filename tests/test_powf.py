"""The exner power function.

The product evaluates x^(R/cv) by a binary64 ln/exp method rounded once
(correctly rounded except within ~1e-16 of a tie) and moves the result to the
other neighbour where a bitmap, built from the host C library's powf, marks
the argument: the GNU powf bits the reference's v8.4.1 exner uses
(cuda_driver.py:1466-1472).  The golden vectors hold the reference's own GNU
powf transcription (native_powf.py) at every argument in [2^-12, 4) where it
differs from the binary64 method.
"""
import ctypes
import ctypes.util
from pathlib import Path

import numpy as np

from hexdyn import lib as hlib
from oracle.oracle import powf as oracle_powf
from oracle.oracle import powf_many

libm = ctypes.CDLL(ctypes.util.find_library("m"))
libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
libm.powf.restype = ctypes.c_float

RCV = np.float32(287.0) / (np.float32(1004.5) - np.float32(287.0))
GOLDEN = Path(__file__).resolve().parent / "golden" / "reference_powf.npz"


def _ulps(a, b):
    ia = np.asarray(a, dtype=np.float32).view(np.int32).astype(np.int64)
    ib = np.asarray(b, dtype=np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def test_binary64_method_is_correctly_rounded_almost_everywhere():
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.uniform(0.005, 1.3, 200000), rng.uniform(1e-6, 50.0, 50000)]).astype(np.float32)
    ours = powf_many(xs, RCV, "hexdyn")
    cr = np.power(xs.astype(np.float64), np.float64(RCV)).astype(np.float32)
    assert (_ulps(ours, cr) == 0).mean() > 0.99999
    assert _ulps(ours, cr).max() <= 1


def test_reference_gnu_powf_is_the_host_powf():
    """Every argument where the reference's GNU powf leaves the correctly
    rounded value: the host C library gives the same bits there."""
    z = np.load(GOLDEN, allow_pickle=False)
    assert np.float32(z["exponent"][0]) == RCV
    assert int(z["host_libm_differs"][0]) == 0
    x = z["x_bits"].view(np.float32)
    assert x.size > 70000
    host = powf_many(x, RCV, "libm")
    assert np.array_equal(host.view(np.uint32), z["reference"].view(np.uint32))
    # and the binary64 method really differs there (by one ulp)
    assert (_ulps(powf_many(x, RCV, "hexdyn"), z["reference"]) == 1).all()


def test_product_exner_power_gives_the_gnu_bits():
    """hexdyn_powf (the host copy of the device's corrected evaluation) equals
    the reference's GNU powf on every golden argument and the host powf on a
    random sample of the table's range."""
    L = hlib.lib()
    z = np.load(GOLDEN, allow_pickle=False)
    x = z["x_bits"].view(np.float32)
    pick = np.random.default_rng(3).choice(x.size, 4000, replace=False)
    got = np.array([L.hexdyn_powf(float(v), float(RCV)) for v in x[pick]], dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), z["reference"][pick].view(np.uint32))
    lo, hi = z["range_bits"]
    sample = np.random.default_rng(5).integers(lo, hi, 20000, dtype=np.uint32).view(np.float32)
    got = np.array([L.hexdyn_powf(float(v), float(RCV)) for v in sample], dtype=np.float32)
    host = powf_many(sample, RCV, "libm")
    assert np.array_equal(got.view(np.uint32), host.view(np.uint32))
    via_oracle = np.array([oracle_powf(float(v), float(RCV)) for v in sample], dtype=np.float32)
    assert np.array_equal(via_oracle.view(np.uint32), host.view(np.uint32))


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
