"""Accuracy of the kernel's Float32 erf (gaussmle.jl_amd/csrc/gmle_erf.hpp).

The same header is compiled for the host and evaluated on a dense grid against
scipy's Float64 erf, with exp(-u^2) supplied both correctly rounded and the way
the kernel forms it (exp2 of the Float32 product -u*u*log2(e)).  The reference
takes erf from OpenLibm's erff (about 1 ulp); the kernel's must stay within a
few ulp of the true value everywhere.
"""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest
from scipy.special import erf

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def kerf(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    out = tmp_path_factory.mktemp("erf") / "liberf.so"
    subprocess.run([hipcc, "-O2", "-fPIC", "-shared", "-I", str(ROOT / "gaussmle.jl_amd" / "csrc"),
                    str(ROOT / "tests" / "erf_harness.cpp"), "-o", str(out)], check=True, capture_output=True)
    lib = C.CDLL(str(out))
    P = C.POINTER(C.c_float)
    lib.erf_with_gauss_batch.argtypes = [P, P, P, C.c_long]

    def f(u, g):
        u = np.ascontiguousarray(u, np.float32)
        g = np.ascontiguousarray(g, np.float32)
        o = np.empty_like(u)
        lib.erf_with_gauss_batch(u.ctypes.data_as(P), g.ctypes.data_as(P), o.ctypes.data_as(P), u.size)
        return o
    return f


def _ulp(x):
    return np.spacing(np.abs(x).astype(np.float32)).astype(np.float64)


def test_erf_accuracy_on_dense_grid(kerf):
    u = np.linspace(-5, 5, 2_000_001).astype(np.float32)
    u = np.concatenate([u, np.float32([0.875, -0.875, np.nextafter(np.float32(0.875), 0), 3.92, 1e-30, -1e-20])])
    exact = erf(u.astype(np.float64))
    g_exact = np.exp(-(u.astype(np.float64) ** 2)).astype(np.float32)
    a = (u * u).astype(np.float32)
    g_kernel = np.exp2(-(a * np.float32(1.4426950408889634)).astype(np.float64)).astype(np.float32)
    for g in (g_exact, g_kernel):
        v = kerf(u, g).astype(np.float64)
        err = np.abs(v - exact)
        ulps = err / _ulp(exact.astype(np.float32))
        print(f"[erf] max abs {err.max():.2e} max ulp {ulps.max():.2f} mean ulp {ulps.mean():.3f}")
        assert err.max() < 1.5e-7
        assert ulps.max() <= 3.0
        np.testing.assert_array_equal(np.sign(v), np.sign(exact))   # odd, sign-exact


def test_erf_special_values(kerf):
    u = np.float32([0.0, -0.0, np.inf, -np.inf, np.nan, 10.0, -10.0])
    g = np.exp(-(u.astype(np.float64) ** 2)).astype(np.float32)
    v = kerf(u, g)
    assert v[0] == 0.0 and v[1] == 0.0
    assert v[2] == 1.0 and v[3] == -1.0 and np.isnan(v[4])
    assert v[5] == 1.0 and v[6] == -1.0


def test_erf_monotone_pixel_integrals(kerf):
    """Phi = (erf(b) - erf(a)) / 2 stays >= 0 across adjacent edges, which the
    kernel's Poisson fast-path bound relies on (up to rounding)."""
    u = np.linspace(-4.5, 4.5, 900_001).astype(np.float32)
    g = np.exp(-(u.astype(np.float64) ** 2)).astype(np.float32)
    v = kerf(u, g)
    assert np.all(np.diff(v) >= -1.2e-7)
