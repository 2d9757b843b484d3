"""ctypes wrapper of oracle/librefcpu.so, the CPU restatement of GaussMLE.jl's fit.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or as the timed CPU baseline.
The product path (gaussmle.jl_amd) never imports it.

Precision codes: 0 = Float32 (the reference's CPU path), 1 = Float64
everywhere, 2 = Float32 iterations + Float64 final pass.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "librefcpu.so"

FP32, FP64, MIXED = 0, 1, 2

_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)

_lib = None


def build(force: bool = False) -> Path:
    """Compile the oracle with its Makefile."""
    if force or not LIB.exists() or LIB.stat().st_mtime < max(
            (HERE / "refcpu.cpp").stat().st_mtime, (HERE / "refcpu.hpp").stat().st_mtime,
            (HERE / "refboxer.cpp").stat().st_mtime, (HERE / "refsim.cpp").stat().st_mtime):
        subprocess.run(["make", "-s", "-C", str(HERE), "-B" if force else "all"], check=True)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        lib.refcpu_n_params.restype = C.c_int
        lib.refcpu_n_params.argtypes = [C.c_int]
        lib.refcpu_integral_gaussian_1d.restype = C.c_double
        lib.refcpu_integral_gaussian_1d.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        lib.refcpu_pixel_derivs.restype = C.c_double
        lib.refcpu_pixel_derivs.argtypes = [C.c_int, C.c_int, _F, C.c_int, C.c_int, _D, _D, _D]
        lib.refcpu_pixel_terms.restype = None
        lib.refcpu_pixel_terms.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _D, _D, _D]
        lib.refcpu_invert_fisher.restype = C.c_int
        lib.refcpu_invert_fisher.argtypes = [C.c_int, C.c_int, _D, _D]
        lib.refcpu_cholesky_inverse.restype = C.c_int
        lib.refcpu_cholesky_inverse.argtypes = [C.c_int, C.c_int, _D, _D]
        lib.refcpu_lu_inverse.restype = C.c_int
        lib.refcpu_lu_inverse.argtypes = [C.c_int, C.c_int, _D, _D]
        lib.refcpu_initialize.restype = None
        lib.refcpu_initialize.argtypes = [C.c_int, C.c_int, _F, C.c_int, _F, _D]
        lib.refcpu_fit.restype = C.c_int
        lib.refcpu_fit.argtypes = [C.c_int, C.c_int, _F, C.c_int, C.c_int64, _F, C.c_int, _F, C.c_int, C.c_int,
                                   _I, _I, _F, _F, _F, C.c_int, _D, _D, _D, _I, C.c_int]
        lib.refcpu_poisson_branches.restype = C.c_int
        lib.refcpu_poisson_branches.argtypes = [C.c_int, _F, C.c_int, C.c_int64, _F, _F, _F, _F, C.c_int, _I,
                                                C.c_int]
        lib.refcpu_final_pass.restype = C.c_int
        lib.refcpu_final_pass.argtypes = [C.c_int, C.c_int, _F, C.c_int, C.c_int64, _F, C.c_int, _F, C.c_int,
                                          C.c_int, _I, _I, _D, _D, _D, C.c_int]
        _lib = lib
    return _lib


def _f(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(_F)


def _psf(model_id: int, psf8) -> np.ndarray:
    p = np.zeros(8, dtype=np.float32)
    if psf8 is not None:
        a = np.asarray(psf8, dtype=np.float32).ravel()
        p[:len(a)] = a
    return p


def integral_gaussian_1d(i: int, pos: float, sigma: float, prec: int = FP64) -> float:
    return load().refcpu_integral_gaussian_1d(prec, i, pos, sigma)


def pixel_derivs(model_id: int, psf8, i: int, j: int, theta, prec: int = FP64):
    """(model value, dudt[P], d2udt2_diag[P]) at pixel (row i, column j), 1-based."""
    lib = load()
    P = lib.refcpu_n_params(model_id)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    du = np.zeros(P)
    d2 = np.zeros(P)
    psf = _psf(model_id, psf8)
    m = lib.refcpu_pixel_derivs(prec, model_id, psf.ctypes.data_as(_F), i, j, th.ctypes.data_as(_D),
                                du.ctypes.data_as(_D), d2.ctypes.data_as(_D))
    return m, du, d2


def pixel_terms(camera: int, data: float, model: float, var: float = 0.0, prec: int = FP64):
    """(cf, df, llr term) of one pixel."""
    cf, df, ll = C.c_double(), C.c_double(), C.c_double()
    load().refcpu_pixel_terms(prec, camera, data, model, var, C.byref(cf), C.byref(df), C.byref(ll))
    return cf.value, df.value, ll.value


def invert_fisher(F, prec: int = FP64):
    """(crlb[P], path) with path 0 = Cholesky, 1 = LU, 2 = singular."""
    F = np.asfortranarray(F, dtype=np.float64)
    P = F.shape[0]
    flat = F.reshape(-1, order="F").copy()
    crlb = np.zeros(P)
    path = load().refcpu_invert_fisher(prec, P, flat.ctypes.data_as(_D), crlb.ctypes.data_as(_D))
    return crlb, path


def cholesky_inverse(A, prec: int = FP64):
    A = np.asfortranarray(A, dtype=np.float64)
    P = A.shape[0]
    flat = A.reshape(-1, order="F").copy()
    out = np.zeros(P * P)
    ok = load().refcpu_cholesky_inverse(prec, P, flat.ctypes.data_as(_D), out.ctypes.data_as(_D))
    return (out.reshape((P, P), order="F") if ok else None)


def lu_inverse(A, prec: int = FP64):
    A = np.asfortranarray(A, dtype=np.float64)
    P = A.shape[0]
    flat = A.reshape(-1, order="F").copy()
    out = np.zeros(P * P)
    ok = load().refcpu_lu_inverse(prec, P, flat.ctypes.data_as(_D), out.ctypes.data_as(_D))
    return out.reshape((P, P), order="F"), bool(ok)


def initialize(model_id: int, psf8, roi: np.ndarray, prec: int = FP32) -> np.ndarray:
    """Initial estimates for one ROI given as (box, box) indexed [row, col]."""
    lib = load()
    box = roi.shape[0]
    flat = np.asfortranarray(roi, dtype=np.float32).reshape(-1, order="F")
    th = np.zeros(lib.refcpu_n_params(model_id))
    psf = _psf(model_id, psf8)
    lib.refcpu_initialize(prec, model_id, psf.ctypes.data_as(_F), box, flat.ctypes.data_as(_F),
                          th.ctypes.data_as(_D))
    return th


def fit(model_id: int, psf8, box: int, data_flat: np.ndarray, lower, upper, max_step, iterations: int,
        prec: int = FP32, var_map: np.ndarray | None = None, x_corners=None, y_corners=None, nthreads: int = 0):
    """Fit every ROI of data_flat (n*box*box Float32 in the reference's layout).

    Returns dict(params (n, P), crlb (n, P), llr (n,), path (n,)) in Float64
    (Float32 results converted exactly).  var_map (ny, nx) indexed [row, col]
    selects the sCMOS likelihood.
    """
    lib = load()
    P = lib.refcpu_n_params(model_id)
    d = np.ascontiguousarray(data_flat, dtype=np.float32)
    n = d.size // (box * box)
    params = np.zeros(n * P)
    crlb = np.zeros(n * P)
    llr = np.zeros(n)
    path = np.zeros(n, dtype=np.int32)
    camera = 0 if var_map is None else 1
    ny = nx = 0
    vflat = None
    if camera:
        v = np.asfortranarray(var_map, dtype=np.float32)
        ny, nx = v.shape
        vflat = v.reshape(-1, order="F")
    xc = None if x_corners is None else np.ascontiguousarray(x_corners, dtype=np.int32)
    yc = None if y_corners is None else np.ascontiguousarray(y_corners, dtype=np.int32)
    psf = _psf(model_id, psf8)
    rc = lib.refcpu_fit(prec, model_id, psf.ctypes.data_as(_F), box, n, d.ctypes.data_as(_F), camera, _f(vflat),
                        ny, nx, None if xc is None else xc.ctypes.data_as(_I),
                        None if yc is None else yc.ctypes.data_as(_I), _f(lower), _f(upper), _f(max_step),
                        iterations, params.ctypes.data_as(_D), crlb.ctypes.data_as(_D), llr.ctypes.data_as(_D),
                        path.ctypes.data_as(_I), nthreads)
    if rc != 0:
        raise ValueError("refcpu_fit: bad arguments")
    return {"params": params.reshape(n, P), "crlb": crlb.reshape(n, P), "llr": llr, "path": path}


def fit_workload(w, prec: int = FP32, nthreads: int = 0, n: int | None = None):
    """Fit a gaussmle_amd.synth.Workload (optionally its first n ROIs)."""
    ww = w if n is None else w.subset(n)
    c = ww.constraints
    return fit(ww.psf_pixels.model_id, ww.psf_pixels.psf_array(), ww.box, ww.flat_data(), c.lower, c.upper,
               c.max_step, ww.iterations, prec, ww.var_map, ww.x_corners, ww.y_corners, nthreads)


def poisson_branches(w, nthreads: int = 0) -> np.ndarray:
    """Coverage counters of the Float32 fit of a Poisson synth.Workload, (n, 4):
    pixel terms below the 0.01 threshold, cf caps, df caps (summed over the
    Newton iterations), and iterations in which the GPU kernel evaluates these
    branches (its fast-path precondition fails)."""
    lib = load()
    c = w.constraints
    d = np.ascontiguousarray(w.flat_data(), dtype=np.float32)
    out = np.zeros((w.n, 4), dtype=np.int32)
    psf = _psf(w.psf_pixels.model_id, w.psf_pixels.psf_array())
    rc = lib.refcpu_poisson_branches(w.psf_pixels.model_id, psf.ctypes.data_as(_F), w.box, w.n,
                                     d.ctypes.data_as(_F), _f(c.lower), _f(c.upper), _f(c.max_step), w.iterations,
                                     out.ctypes.data_as(_I), nthreads)
    if rc != 0:
        raise ValueError("refcpu_poisson_branches: bad arguments")
    return out


def final_pass(w, theta: np.ndarray, prec: int = FP64, nthreads: int = 0):
    """LLR / Fisher / CRLB of a synth.Workload at given estimates theta (n, P)."""
    lib = load()
    P = len(w.psf_pixels)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    n = th.shape[0]
    crlb = np.zeros(n * P)
    llr = np.zeros(n)
    d = np.ascontiguousarray(w.flat_data()[: n * w.box * w.box], dtype=np.float32)
    camera = 0 if w.var_map is None else 1
    ny = nx = 0
    vflat = None
    if camera:
        v = np.asfortranarray(w.var_map, dtype=np.float32)
        ny, nx = v.shape
        vflat = v.reshape(-1, order="F")
    xc = None if w.x_corners is None else np.ascontiguousarray(w.x_corners[:n], dtype=np.int32)
    yc = None if w.y_corners is None else np.ascontiguousarray(w.y_corners[:n], dtype=np.int32)
    psf = _psf(w.psf_pixels.model_id, w.psf_pixels.psf_array())
    rc = lib.refcpu_final_pass(prec, w.psf_pixels.model_id, psf.ctypes.data_as(_F), w.box, n, d.ctypes.data_as(_F),
                               camera, _f(vflat), ny, nx, None if xc is None else xc.ctypes.data_as(_I),
                               None if yc is None else yc.ctypes.data_as(_I), th.ctypes.data_as(_D),
                               crlb.ctypes.data_as(_D), llr.ctypes.data_as(_D), nthreads)
    if rc != 0:
        raise ValueError("refcpu_final_pass: bad arguments")
    return {"crlb": crlb.reshape(n, P), "llr": llr}
