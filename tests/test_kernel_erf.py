"""Accuracy of the kernel's Float32 Gaussian pixel integrals
(gaussmle.jl_amd/csrc/gmle_erf.hpp).

The same header is compiled for the host and evaluated on dense grids against
scipy's Float64 erf / erfc, with exp(-u^2) supplied both correctly rounded and
the way the kernel forms it (exp2 of -v*v with v = sqrt(log2 e) u in Float32).
The kernel carries each pixel edge as sign(u) erfc(|u|) = sign(u) exp(-u^2)
R(t), R a polynomial fit of erfcx in t = 1 - 2c / (|v| + c), and forms pixel
integrals
(erf(b) - erf(a)) / 2 from those complements, so that tail pixels keep their
relative accuracy; the complement, the integrals (absolutely and relative to
their size, down to the far tails) and the erf the same pieces give are each
checked.
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
    lib.signed_erfc_batch.argtypes = [P, P, P, C.c_long]
    lib.pixel_integral_batch.argtypes = [P, P, P, P, P, C.c_long]

    def f(u, g):
        u = np.ascontiguousarray(u, np.float32)
        g = np.ascontiguousarray(g, np.float32)
        o = np.empty_like(u)
        lib.erf_with_gauss_batch(u.ctypes.data_as(P), g.ctypes.data_as(P), o.ctypes.data_as(P), u.size)
        return o

    def serfc(u, g):
        u = np.ascontiguousarray(u, np.float32)
        g = np.ascontiguousarray(g, np.float32)
        o = np.empty_like(u)
        lib.signed_erfc_batch(u.ctypes.data_as(P), g.ctypes.data_as(P), o.ctypes.data_as(P), u.size)
        return o
    f.serfc = serfc

    def phi(lo, glo, hi, ghi):
        a = [np.ascontiguousarray(x, np.float32) for x in (lo, glo, hi, ghi)]
        o = np.empty_like(a[0])
        lib.pixel_integral_batch(*[x.ctypes.data_as(P) for x in a], o.ctypes.data_as(P), o.size)
        return o
    f.phi = phi
    return f


K = np.float32(1.2011224031448364)   # sqrt(log2 e), kErfK


def _g_kernel(x):
    """exp(-u^2) as the kernel forms it: exp2(-v*v) with v = k u in Float32."""
    v = (np.asarray(x, np.float32) * K).astype(np.float32)
    return np.exp2(-((v * v).astype(np.float32)).astype(np.float64)).astype(np.float32)


def _ulp(x):
    return np.spacing(np.abs(x).astype(np.float32)).astype(np.float64)


def test_erf_accuracy_on_dense_grid(kerf):
    from scipy.special import erfc
    u = np.linspace(-6, 6, 2_400_001).astype(np.float32)
    u = np.concatenate([u, np.float32([0.875, -0.875, 3.92, -3.92, 5.6, -5.6, 1e-30, -1e-20])])
    exact = erf(u.astype(np.float64))
    g_exact = np.exp(-(u.astype(np.float64) ** 2)).astype(np.float32)
    for g in (g_exact, _g_kernel(u)):
        # the complement the kernels carry: relative accuracy out to |u| = 5.6
        se = kerf.serfc(u, g).astype(np.float64)
        ec = np.sign(u) * erfc(np.abs(u).astype(np.float64))
        inside = np.abs(u) <= 5.6
        rel = np.abs(se[inside] / ec[inside] - 1)
        # R's own error, plus, with the kernel's Gaussian, the Float32
        # rounding of v^2 (relative 2^-24 of ~u^2, times ln 2 in exp2)
        tol = 1.0e-6 + (2.0e-7 * u[inside].astype(np.float64) ** 2 if g is not g_exact else 0.0)
        # the erf the same pieces give (1 - erfc, absolute accuracy)
        v = kerf(u, g).astype(np.float64)
        err = np.abs(v - exact)
        print(f"[erf] erfc max rel {rel.max():.2e}, max rel / tolerance {np.max(rel / tol):.2f} (|u| <= 5.6); "
              f"erf max abs {err.max():.2e}", flush=True)
        assert np.all(rel <= tol)
        assert err.max() < 4.0e-7
        np.testing.assert_array_equal(v[u > 0] >= 0, True)                 # sign follows u
        np.testing.assert_array_equal(kerf(-u, g), -kerf(u, g))            # odd, exactly
        np.testing.assert_array_equal(kerf.serfc(-u, g), -kerf.serfc(u, g))


@pytest.mark.parametrize("sigma", [0.8, 1.3, 2.5])
def test_pixel_integrals_against_scipy(kerf, sigma):
    """Phi of a pixel at every offset of the centre, as the kernel forms it:
    relative accuracy wherever the pixel's edges lie within |u| <= 5.6, and
    an absolute error far below the tail's own size beyond."""
    from scipy.special import erfc
    s = np.float32(np.sqrt(0.5 / sigma ** 2))
    mu = np.linspace(-9, 9, 600_001)
    lo = ((np.float32(-0.5) - mu.astype(np.float32)) * s).astype(np.float32)
    hi = ((np.float32(0.5) - mu.astype(np.float32)) * s).astype(np.float32)
    phi = kerf.phi(lo, _g_kernel(lo), hi, _g_kernel(hi)).astype(np.float64)
    l64, h64 = lo.astype(np.float64), hi.astype(np.float64)
    # the exact integral without cancellation: complements on the same side
    want = np.where(l64 >= 0, 0.5 * (erfc(l64) - erfc(h64)),
                    np.where(h64 <= 0, 0.5 * (erfc(-h64) - erfc(-l64)), 0.5 * (erf(h64) - erf(l64))))
    err = np.abs(phi - want)
    inside = np.maximum(np.abs(l64), np.abs(h64)) <= 5.6
    rel = err[inside] / want[inside]
    # R's error, amplified where the pixel's two complements are close, plus
    # the Float32 rounding of the Gaussian's argument (see the test above)
    umax = np.maximum(np.abs(l64), np.abs(h64))[inside]
    tol = 2.0e-6 + 4.0e-7 * umax ** 2
    beyond = ~inside
    eb = err[beyond].max() if beyond.any() else 0.0
    print(f"[erf] sigma {sigma}: Phi max abs {err.max():.2e}; within |u| <= 5.6 max rel {rel.max():.2e} "
          f"(down to Phi = {want[inside].min():.1e}); beyond: max abs {eb:.2e}", flush=True)
    # |error of Phi| <= (|error of se_lo| + |error of se_hi|) / 2 <= R's
    # relative error (3.7e-7 with a correctly rounded reciprocal, as here)
    # times erfc <= 1, plus the Gaussian's rounding
    assert err.max() < 4.0e-7
    assert np.all(rel <= tol), float(np.max(rel / tol))
    # beyond |u| = 5.6 the integrals are below 1e-15 and stay far below that
    # in absolute error
    assert eb < 1e-15
    assert np.all(phi >= 0)


def test_tail_integrals_keep_relative_accuracy(kerf):
    """The reason for the complement form: a tail pixel's integral, here
    1e-2 to 1e-12 of the total (edges up to |u| = 5.6), is accurate
    relatively, where the Float32 difference erf(b) - erf(a) of two values
    near 1 is off by up to order one."""
    from scipy.special import erfc
    lo = np.linspace(1.5, 5.1, 10_001).astype(np.float32)
    hi = (lo + np.float32(0.5)).astype(np.float32)
    phi = kerf.phi(lo, _g_kernel(lo), hi, _g_kernel(hi)).astype(np.float64)
    want = 0.5 * (erfc(lo.astype(np.float64)) - erfc(hi.astype(np.float64)))
    rel = np.abs(phi / want - 1)
    naive = 0.5 * (kerf(hi, _g_kernel(hi)).astype(np.float64) - kerf(lo, _g_kernel(lo)))
    rel_naive = np.abs(naive / want - 1)
    print(f"[erf] tails: complement form max rel {rel.max():.2e}, erf difference max rel {rel_naive.max():.2e}",
          flush=True)
    assert np.all(rel <= 2.0e-6 + 4.0e-7 * hi.astype(np.float64) ** 2)
    assert rel_naive.max() > 0.5


def test_erf_special_values(kerf):
    u = np.float32([0.0, -0.0, np.inf, -np.inf, np.nan, 10.0, -10.0, 4.5, -4.5])
    g = np.exp(-(u.astype(np.float64) ** 2)).astype(np.float32)
    v = kerf(u, g)
    assert abs(v[0]) < 4e-7 and abs(v[1]) < 4e-7 and np.signbit(v[1])
    assert v[2] == 1.0 and v[3] == -1.0 and np.isnan(v[4])
    assert v[5] == 1.0 and v[6] == -1.0 and v[7] == 1.0 and v[8] == -1.0
    se = kerf.serfc(u, g)
    assert abs(se[0] - 1) < 4e-7 and abs(se[1] + 1) < 4e-7      # erfc(0) = 1, signed
    assert se[2] == 0 and se[3] == 0 and np.isnan(se[4]) and 0 <= se[5] < 1e-40
    # the pixel straddling the centre, and integrals over infinite edges
    one = kerf.phi(np.float32([-np.inf]), np.float32([0]), np.float32([np.inf]), np.float32([0]))
    assert one[0] == 1.0


def test_erf_monotone_pixel_integrals(kerf):
    """Phi = (erf(b) - erf(a)) / 2 stays >= 0 across adjacent edges, which the
    kernel's Poisson fast-path bound relies on (up to rounding)."""
    u = np.linspace(-4.5, 4.5, 900_001).astype(np.float32)
    g = np.exp(-(u.astype(np.float64) ** 2)).astype(np.float32)
    v = kerf(u, g)
    assert np.all(np.diff(v) >= -3.0e-7)
