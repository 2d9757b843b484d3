"""ctypes wrapper of the ROI-extraction restatement in oracle/refboxer.cpp.

TEST INFRASTRUCTURE ONLY: imported by tests/ and the cpu_baseline leg of
bench.py, as the checker or as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import refcpu

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)
_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        refcpu.build()
        lib = C.CDLL(str(refcpu.LIB))
        lib.refboxer_weights.restype = C.c_int
        lib.refboxer_weights.argtypes = [C.c_float, _F, C.c_int]
        lib.refboxer_dog.restype = None
        lib.refboxer_dog.argtypes = [_F, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, _F]
        lib.refboxer_find.restype = C.c_int64
        lib.refboxer_find.argtypes = [_F, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                      C.c_float, C.c_int64, _F, _I, _I, _I, _F, _F]
        _lib = lib
    return _lib


def _stack(frames) -> np.ndarray:
    a = np.asarray(frames, dtype=np.float32)
    if a.ndim == 2:
        a = a[:, :, None]
    return np.asfortranarray(a)


def weights(sigma: float) -> np.ndarray:
    w = np.zeros(64, np.float32)
    R = load().refboxer_weights(float(sigma), w.ctypes.data_as(_F), 64)
    return w[:R + 1].copy()


def dog(frames, sigma_small: float, sigma_large: float) -> np.ndarray:
    a = _stack(frames)
    ny, nx, nf = a.shape
    out = np.zeros((ny, nx, nf), np.float32, order="F")
    load().refboxer_dog(a.ctypes.data_as(_F), ny, nx, nf, float(sigma_small), float(sigma_large),
                        out.ctypes.data_as(_F))
    return out


def find(frames, boxsize: int, overlap: float, sigma_small: float, sigma_large: float, minval: float,
         max_boxes: int | None = None) -> dict:
    """All detections: rois (box, box, n), x/y corners and frames (1-based),
    strength, and the filtered stack."""
    a = _stack(frames)
    ny, nx, nf = a.shape
    cap = max_boxes if max_boxes is not None else ny * nx * nf
    cap = max(cap, 1)
    rois = np.zeros((boxsize, boxsize, cap), np.float32, order="F")
    xc, yc, fr = (np.zeros(cap, np.int32) for _ in range(3))
    st = np.zeros(cap, np.float32)
    d = np.zeros((ny, nx, nf), np.float32, order="F")
    n = load().refboxer_find(a.ctypes.data_as(_F), ny, nx, nf, int(boxsize), float(overlap), float(sigma_small),
                             float(sigma_large), float(minval), cap, rois.ctypes.data_as(_F), xc.ctypes.data_as(_I),
                             yc.ctypes.data_as(_I), fr.ctypes.data_as(_I), st.ctypes.data_as(_F),
                             d.ctypes.data_as(_F))
    m = min(n, cap)
    return {"n": int(n), "rois": np.asfortranarray(rois[:, :, :m]), "x_corners": xc[:m], "y_corners": yc[:m],
            "frames": fr[:m], "strength": st[:m], "dog": d}
