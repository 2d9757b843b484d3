"""ctypes binding of libgaussmle_hip.so (the C ABI in include/gaussmle.h).

The product path has no CPU fallback: if the HIP library is missing, loading it
raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
from pathlib import Path

import numpy as np

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_ROOT / "lib" / "libgaussmle_hip.so"

GMLE_OK = 0
GMLE_ERR_ARG = -1
GMLE_ERR_BOUNDS = -2
GMLE_ERR_DATA = -3
GMLE_ERR_HIP = -4
GMLE_ERR_NOMEM = -5
GMLE_ERR_NODEVICE = -6
GMLE_ERR_UNSUPPORTED = -7

MODEL_IDS = {"GaussianXYNB": 0, "GaussianXYNBS": 1, "GaussianXYNBSXSY": 2, "AstigmaticXYZNB": 3}
CAMERA_IDS = {"ideal": 0, "scmos": 1}
PRECISION_IDS = {"fp32": 0, "mixed": 1}

# Every function the header declares, with its ctypes signature.
_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_I8 = C.POINTER(C.c_int8)
SIGNATURES = {
    "gmle_version": (C.c_char_p, []),
    "gmle_device_count": (C.c_int, []),
    "gmle_last_error": (C.c_size_t, [C.c_char_p, C.c_size_t]),
    "gmle_n_params": (C.c_int, [C.c_int]),
    "gmle_default_constraints": (C.c_int, [C.c_int, C.c_int, _F, _F, _F]),
    "gmle_max_box": (C.c_int, []),
    "gmle_lanes_supported": (C.c_int, [C.c_int, C.c_int]),
    "gmle_default_lanes": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64]),
    "gmle_fit": (
        C.c_int,
        [C.c_int, _F, C.c_int, C.c_int64, _F, C.c_int, _F, C.c_int, C.c_int, _I32, _I32,
         _F, _F, _F, C.c_int, C.c_int, C.c_int, _F, _F, _F, _D, _D],
    ),
    "gmle_fit_device": (
        C.c_int,
        [C.c_int, _F, C.c_int, C.c_int64, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
         C.c_void_p, _F, _F, _F, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
         C.c_void_p, C.c_void_p],
    ),
    "gmle_boxer_workspace_bytes": (C.c_int64, [C.c_int, C.c_int, C.c_int]),
    "gmle_find_boxes_device": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                         C.c_float, C.c_float, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gmle_find_boxes": (C.c_int, [_F, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                  C.c_float, C.c_int64, _F, _I32, _I32, _I32, _F, _I64]),
    "gmle_boxer_weights": (C.c_int, [C.c_float, _F, C.c_int]),
    "gmle_simulate_device": (C.c_int, [C.c_int, _F, C.c_int, C.c_int64, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                       C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint64, C.c_void_p,
                                       C.c_void_p]),
    "gmle_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "gmle_epilogue_device": (C.c_int, [C.c_int64, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                       C.c_void_p, C.c_void_p]),
    "gmle_chi2_pvalue": (C.c_float, [C.c_float, C.c_int]),
    "gmle_shape_index": (C.c_int, [C.c_int, C.c_int]),
    "gmle_one_launch_lanes": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "gmle_batch_workspace_bytes": (C.c_size_t, [C.c_int64]),
    "gmle_fit_batch": (
        C.c_int,
        [C.c_int64, _I8, _I8, _F, _F, _F, _F, _F, C.c_int, _F, C.c_int, C.c_int, _I32, _I32, C.c_int, C.c_int,
         _I32, C.c_int, _F, _F, _F, _D, _D],
    ),
    "gmle_fit_batch_device": (
        C.c_int,
        [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _F, _F, _F, _F, C.c_int, C.c_void_p, C.c_int,
         C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, _I32, C.c_int, _I64, C.c_void_p, C.c_void_p,
         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p],
    ),
    "gmle_fit_batch_device_noalloc": (
        C.c_int,
        [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _F, _F, _F, _F, C.c_int, C.c_void_p, C.c_int,
         C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, _I32, C.c_int, _I64, C.c_void_p, C.c_void_p,
         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    ),
}



class Request(C.Structure):
    """struct gmle_request (include/gaussmle.h)."""
    _fields_ = [("n_rois", C.c_int64), ("models", C.c_void_p), ("boxes", C.c_void_p), ("x_corners", C.c_void_p),
                ("y_corners", C.c_void_p), ("adu", C.c_void_p), ("params", C.c_void_p), ("crlb", C.c_void_p),
                ("llr", C.c_void_p)]


_RP = C.POINTER(Request)
_DP = C.POINTER(C.c_double)
SIGNATURES.update({
    "gmle_host_alloc": (C.c_void_p, [C.c_size_t]),
    "gmle_host_free": (C.c_int, [C.c_void_p]),
    "gmle_service_create": (C.c_int, [C.c_int, C.c_size_t, C.c_int, C.c_int, C.c_int, _F, _F, _F, _F, _I32,
                                      C.POINTER(C.c_void_p)]),
    "gmle_service_calibrate": (C.c_int, [C.c_void_p, _F, _F, _F, C.c_int, C.c_int]),
    "gmle_service_submit": (C.c_int64, [C.c_void_p, _RP]),
    "gmle_service_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_double]),
    "gmle_service_replay": (C.c_int, [C.c_void_p, C.c_int64, _RP, _DP, _DP, _DP]),
    "gmle_service_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                    C.POINTER(C.c_int64)]),
    "gmle_service_stats": (C.c_int, [C.c_void_p, _DP, C.c_int]),
    "gmle_service_destroy": (C.c_int, [C.c_void_p]),
})

NUM_SHAPES = 56   # GMLE_NUM_SHAPES: 4 models x box sizes 3..16
BATCH_STRIDE = 6  # params / crlb values per ROI slot of a mixed batch


class GaussMLEError(RuntimeError):
    """Raised when the HIP library returns an error code."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class ArgumentError(GaussMLEError, ValueError):
    """The reference's ArgumentError (invalid input rejected before launch)."""


_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load (once) and return the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    # torch (when installed) bundles its own HIP / HSA runtime under the same
    # sonames as the system ROCm, and one process can only drive the GPU
    # through one of them: load torch's first so this library binds to it.
    if "torch" not in sys.modules and os.environ.get("GMLE_NO_TORCH_PRELOAD") != "1":
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    if not p.exists():
        raise RuntimeError(
            f"HIP library not found at {p}; build it with `python gaussmle.jl_amd/build.py` "
            "(there is no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def last_error() -> str:
    lib = load()
    buf = C.create_string_buffer(1024)
    lib.gmle_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int) -> None:
    if rc == GMLE_OK:
        return
    msg = last_error()
    if rc in (GMLE_ERR_ARG, GMLE_ERR_BOUNDS, GMLE_ERR_DATA):
        raise ArgumentError(rc, msg)
    raise GaussMLEError(rc, msg)


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I32)


def i8ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int8 and a.flags.c_contiguous
    return a.ctypes.data_as(_I8)


def i64ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_I64)


def device_count() -> int:
    return int(load().gmle_device_count())
