"""Accuracy of the kernel's Float32 erf (gaussmle.jl_amd/csrc/gmle_erf.hpp).

The same header is compiled for the host and evaluated on a dense grid against
scipy's Float64 erf, with exp(-u^2) supplied both correctly rounded and the way
the kernel forms it (exp2 of the Float32 product -u*u*log2(e)).  The kernel
evaluates erf as 1 - exp(-u^2) erfcx(u), one piece, so its error is absolute
(~2e-7, a few ulp of 1) rather than relative near u = 0.  What the fit
consumes are pixel integrals (erf(b) - erf(a)) / 2 over one pixel; those are
checked against scipy directly, absolutely and relative to their size.
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
    u = np.concatenate([u, np.float32([0.875, -0.875, 3.92, -3.92, 1.96, 1e-30, -1e-20])])
    exact = erf(u.astype(np.float64))
    g_exact = np.exp(-(u.astype(np.float64) ** 2)).astype(np.float32)
    a = (u * u).astype(np.float32)
    g_kernel = np.exp2(-(a * np.float32(1.4426950408889634)).astype(np.float64)).astype(np.float32)
    for g in (g_exact, g_kernel):
        v = kerf(u, g).astype(np.float64)
        err = np.abs(v - exact)
        big = np.abs(u) > 0.5
        ulps = err[big] / _ulp(exact[big].astype(np.float32))
        print(f"[erf] max abs {err.max():.2e}; |u| > 0.5: max ulp {ulps.max():.2f}", flush=True)
        assert err.max() < 3.0e-7
        assert ulps.max() <= 5.0
        np.testing.assert_array_equal(v[u > 0] >= 0, True)                 # sign follows u
        np.testing.assert_array_equal(kerf(-u, g), -kerf(u, g))            # odd, exactly


@pytest.mark.parametrize("sigma", [0.8, 1.3, 2.5])
def test_pixel_integrals_against_scipy(kerf, sigma):
    """Phi of a pixel at every offset of the centre, as the kernel forms it."""
    s = np.float32(np.sqrt(0.5 / sigma ** 2))
    mu = np.linspace(-6, 6, 400_001)
    lo = ((np.float32(-0.5) - mu.astype(np.float32)) * s).astype(np.float32)
    hi = ((np.float32(0.5) - mu.astype(np.float32)) * s).astype(np.float32)
    g = lambda x: np.exp2(-((x * x).astype(np.float32) * np.float32(1.4426950408889634)).astype(np.float64))  # noqa
    phi = 0.5 * (kerf(hi, g(hi).astype(np.float32)).astype(np.float64) - kerf(lo, g(lo).astype(np.float32)))
    want = 0.5 * (erf(hi.astype(np.float64)) - erf(lo.astype(np.float64)))
    err = np.abs(phi - want)
    central = want > 0.05
    rel = err[central] / want[central]
    print(f"[erf] sigma {sigma}: Phi max abs {err.max():.2e}, central max rel {rel.max():.2e}", flush=True)
    assert err.max() < 2.5e-7
    assert rel.max() < 2e-6


def test_erf_special_values(kerf):
    u = np.float32([0.0, -0.0, np.inf, -np.inf, np.nan, 10.0, -10.0, 3.92, -3.92])
    g = np.exp(-(u.astype(np.float64) ** 2)).astype(np.float32)
    v = kerf(u, g)
    assert abs(v[0]) < 3e-7 and abs(v[1]) < 3e-7 and np.signbit(v[1])
    assert v[2] == 1.0 and v[3] == -1.0 and np.isnan(v[4])
    assert v[5] == 1.0 and v[6] == -1.0 and v[7] == 1.0 and v[8] == -1.0


def test_erf_monotone_pixel_integrals(kerf):
    """Phi = (erf(b) - erf(a)) / 2 stays >= 0 across adjacent edges, which the
    kernel's Poisson fast-path bound relies on (up to rounding)."""
    u = np.linspace(-4.5, 4.5, 900_001).astype(np.float32)
    g = np.exp(-(u.astype(np.float64) ** 2)).astype(np.float32)
    v = kerf(u, g)
    assert np.all(np.diff(v) >= -3.0e-7)
