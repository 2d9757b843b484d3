"""ctypes wrapper of the simulator restatement in oracle/refsim.cpp.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
gmle_simulate_device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import refcpu

_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_U = C.POINTER(C.c_uint32)
_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        refcpu.build()
        lib = C.CDLL(str(refcpu.LIB))
        lib.refsim_philox4x32_10.restype = None
        lib.refsim_philox4x32_10.argtypes = [_U, _U, _U]
        lib.refsim_draws.restype = None
        lib.refsim_draws.argtypes = [C.c_int, C.c_double, C.c_uint64, C.c_int64, _D]
        lib.refsim_simulate.restype = None
        lib.refsim_simulate.argtypes = [C.c_int, _F, C.c_int, C.c_int64, _F, C.c_int, C.c_int, _I, _I, _F, _F, _F,
                                        _F, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                        C.c_uint64, _F, _F]
        _lib = lib
    return _lib


def philox(counter, key) -> list:
    c = (C.c_uint32 * 4)(*counter)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    load().refsim_philox4x32_10(c, k, o)
    return list(o)


def draws(kind: str, mean: float, seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float64)
    load().refsim_draws(0 if kind == "poisson" else 1, float(mean), int(seed), int(n), out.ctypes.data_as(_D))
    return out


def _f(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(_F)


def simulate(psf_pixels, box: int, params: np.ndarray, seed: int, camera: int = 0, xc=None, yc=None, maps=None,
             scalars=None, ny: int = 0, nx: int = 0):
    """params (n, stride); maps / scalars dicts with qe, readnoise, gain,
    offset (maps column-major flat).  Returns (data, expected), each
    (n * box * box,) in the fitting layout."""
    params = np.ascontiguousarray(params, dtype=np.float32)
    n, stride = params.shape
    maps = maps or {}
    sc = {"qe": 1.0, "readnoise": 0.0, "gain": 1.0, "offset": 0.0}
    sc.update(scalars or {})
    keep = {k: np.ascontiguousarray(v, dtype=np.float32) for k, v in maps.items()}
    out = np.zeros(n * box * box, np.float32)
    exp = np.zeros(n * box * box, np.float32)
    xci = None if xc is None else np.ascontiguousarray(xc, dtype=np.int32)
    yci = None if yc is None else np.ascontiguousarray(yc, dtype=np.int32)
    psf = np.ascontiguousarray(psf_pixels.psf_array(), dtype=np.float32)
    load().refsim_simulate(psf_pixels.model_id, psf.ctypes.data_as(_F), box, n, params.ctypes.data_as(_F), stride,
                           camera, None if xci is None else xci.ctypes.data_as(_I),
                           None if yci is None else yci.ctypes.data_as(_I),
                           *[_f(keep.get(k)) for k in ("qe", "readnoise", "gain", "offset")], ny, nx,
                           sc["qe"], sc["readnoise"], sc["gain"], sc["offset"], int(seed), out.ctypes.data_as(_F),
                           exp.ctypes.data_as(_F))
    return out, exp
