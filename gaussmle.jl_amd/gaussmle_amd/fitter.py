"""GaussMLEFitter / fit: the host-side API of the drop-in.

Mirrors GaussMLE.jl's src/interface.jl (GaussMLEFitter :115-145, fit(Array)
:177-291, fit(Matrix) :294-301, fit(ROIBatch{IdealCamera}) :304-386,
fit(ROIBatch{SCMOSCamera}) :389-473) and src/devices.jl.  Every fit runs on
the GPU through the C ABI of libgaussmle_hip.so; there is no CPU fallback (the
reference's own CPU path, device = :cpu, is the Julia package itself).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy import stats

from . import _lib
from .models import (AstigmaticXYZNB, GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, IdealCamera,
                     ParameterConstraints, PSFModel, ROIBatch, SCMOSCamera, default_constraints, to_electrons,
                     to_pixel_units, variance_map)
from .results import BasicSMLD, LocalizationResult, to_smld

F32 = np.float32


# ---------------------------------------------------------------------------
# Devices (src/devices.jl)
# ---------------------------------------------------------------------------
class ComputeDevice:
    pass


class GPU(ComputeDevice):
    def __repr__(self):
        return "GPU()"


class CPU(ComputeDevice):
    def __repr__(self):
        return "CPU()"


def auto_device() -> ComputeDevice:
    """GPU when the HIP library sees a device (src/devices.jl:11-19)."""
    if _lib.device_count() > 0:
        return GPU()
    raise RuntimeError("no HIP GPU available; this library has no CPU path")


def select_device(device=None) -> ComputeDevice:
    """Resolve :cpu / :gpu / :auto / None (src/interface.jl:123-137).

    Unlike the reference, a GPU request without a GPU raises instead of falling
    back to a CPU path, and CPU requests are refused: the CPU path is served by
    GaussMLE.jl itself.
    """
    if device in (None, "auto", ":auto"):
        return auto_device()
    if isinstance(device, GPU) or device in ("gpu", ":gpu"):
        if _lib.device_count() <= 0:
            raise RuntimeError("GPU requested but no HIP device is available")
        return GPU()
    if isinstance(device, CPU) or device in ("cpu", ":cpu"):
        raise ValueError("device=:cpu is GaussMLE.jl's own CPU path; this library only provides the GPU path")
    raise ValueError(f"unknown device {device!r}")


# ---------------------------------------------------------------------------
# Fitter
# ---------------------------------------------------------------------------
@dataclass
class GaussMLEFitter:
    """Configuration of a fit (src/interface.jl:73-79).

    precision: "fp32" reproduces the reference's all-Float32 arithmetic;
    "mixed" keeps Float32 Newton iterations but evaluates the final
    log-likelihood, Fisher matrix and its inversion (CRLB) in Float64.
    lanes: lanes per ROI of the kernel (0 = the library's choice).
    """

    device: ComputeDevice
    psf_model: PSFModel
    iterations: int
    constraints: ParameterConstraints
    batch_size: int
    precision: str = "fp32"
    lanes: int = 0

    def __init__(self, psf_model: PSFModel | None = None, device=None, iterations: int = 20,
                 constraints: ParameterConstraints | None = None, batch_size: int = 10_000,
                 precision: str = "fp32", lanes: int = 0):
        self.psf_model = GaussianXYNB(0.13) if psf_model is None else psf_model
        self.device = select_device(device)
        self.iterations = int(iterations)
        self.constraints = default_constraints(self.psf_model, 11) if constraints is None else constraints
        if len(self.constraints) != len(self.psf_model):
            raise ValueError("constraints length does not match the PSF model")
        self.batch_size = int(batch_size)
        if precision not in _lib.PRECISION_IDS:
            raise ValueError(f"precision must be one of {sorted(_lib.PRECISION_IDS)}")
        self.precision = precision
        self.lanes = int(lanes)


def validate_fit_input(data: np.ndarray) -> None:
    """src/interface.jl:10-40 (NaN / Inf and shape checks; negatives allowed)."""
    if data.size == 0:
        raise _lib.ArgumentError(_lib.GMLE_ERR_ARG, "Input data array is empty")
    if data.ndim != 3 or data.shape[0] != data.shape[1]:
        raise _lib.ArgumentError(_lib.GMLE_ERR_ARG, f"ROIs must be square, got size {data.shape[0]}x{data.shape[1]}")
    if data.shape[0] < 3:
        raise _lib.ArgumentError(_lib.GMLE_ERR_ARG,
                                 f"ROI size must be at least 3x3, got {data.shape[0]}x{data.shape[0]}")
    if not np.all(np.isfinite(data)):
        raise _lib.ArgumentError(_lib.GMLE_ERR_DATA, "Input data contains NaN or Inf values")


@dataclass
class FitArrays:
    """Raw outputs of one fit: params / crlb (P, n), llr (n) in Float32, and the
    Float64 CRLB / LLR in mixed precision."""

    params: np.ndarray
    crlb: np.ndarray
    llr: np.ndarray
    crlb64: np.ndarray | None = None
    llr64: np.ndarray | None = None


def fit_arrays(fitter: GaussMLEFitter, data: np.ndarray, psf_pixels: PSFModel, var_map: np.ndarray | None = None,
               x_corners: np.ndarray | None = None, y_corners: np.ndarray | None = None) -> FitArrays:
    """One call of the C ABI's gmle_fit on host arrays in the reference's layout.

    data: (box, box, n) indexed [row, col, roi] (Julia layout).  var_map:
    (ny, nx) indexed [row, col] for sCMOS, else None.
    """
    lib = _lib.load()
    d = np.asfortranarray(data, dtype=F32)
    box, _, n = d.shape
    flat = d.reshape(-1, order="F")
    P = len(psf_pixels)
    params = np.empty(n * P, dtype=F32)
    crlb = np.empty(n * P, dtype=F32)
    llr = np.empty(n, dtype=F32)
    mixed = fitter.precision == "mixed"
    crlb64 = np.empty(n * P, dtype=np.float64) if mixed else None
    llr64 = np.empty(n, dtype=np.float64) if mixed else None
    camera = 1 if var_map is not None else 0
    vflat = None
    ny = nx = 0
    if camera:
        v = np.asfortranarray(var_map, dtype=F32)
        ny, nx = v.shape
        vflat = v.reshape(-1, order="F")
        x_corners = np.ascontiguousarray(x_corners, dtype=np.int32)
        y_corners = np.ascontiguousarray(y_corners, dtype=np.int32)
    c = fitter.constraints
    rc = lib.gmle_fit(psf_pixels.model_id, _lib.fptr(psf_pixels.psf_array()), box, n, _lib.fptr(flat), camera,
                      _lib.fptr(vflat), ny, nx, _lib.iptr(x_corners if camera else None),
                      _lib.iptr(y_corners if camera else None), _lib.fptr(c.lower), _lib.fptr(c.upper),
                      _lib.fptr(c.max_step), fitter.iterations, _lib.PRECISION_IDS[fitter.precision], fitter.lanes,
                      _lib.fptr(params), _lib.fptr(crlb), _lib.fptr(llr), _lib.dptr(crlb64), _lib.dptr(llr64))
    _lib.check(rc)
    shape = (P, n)
    return FitArrays(params.reshape(shape, order="F"), crlb.reshape(shape, order="F"), llr,
                     None if crlb64 is None else crlb64.reshape(shape, order="F"), llr64)


def pvalues(llr: np.ndarray, box: int, n_params: int) -> np.ndarray:
    """Goodness of fit: chi2 = -2 LLR with box^2 - P degrees of freedom,
    p = 1 - Float32(cdf) (src/interface.jl:265-275)."""
    chi2 = F32(-2.0) * np.asarray(llr, dtype=F32)
    cdf = stats.chi2.cdf(chi2.astype(np.float64), box * box - n_params).astype(F32)
    return (F32(1.0) - cdf).astype(F32)


def _dummy_corners(n: int, box: int):
    """fit(Array) corners: x = 1 + (k-1) box, y = 1 (src/interface.jl:203-205)."""
    xc = (1 + np.arange(n, dtype=np.int64) * box).astype(np.int32)
    yc = np.ones(n, dtype=np.int32)
    return xc, yc


def fit(fitter: GaussMLEFitter, data, variance_map: np.ndarray | None = None):
    """fit(fitter, data) -> BasicSMLD, dispatching like the reference:

    * ndarray (box, box, n): src/interface.jl:177-291 (pixel size 0.1 um,
      dummy corners; variance_map switches to the sCMOS likelihood);
    * ndarray (box, box): a single ROI, returns its emitter (:294-301);
    * ROIBatch with IdealCamera (:304-386) or SCMOSCamera (:389-473).
    """
    if isinstance(data, ROIBatch):
        return _fit_batch(fitter, data)
    arr = np.asarray(data)
    if arr.ndim == 2:
        smld = fit(fitter, arr.reshape(arr.shape[0], arr.shape[1], 1), variance_map=variance_map)
        return smld.emitters[0]
    validate_fit_input(arr)
    box, _, n = arr.shape
    pixel_size = 0.1
    psf_pixels = to_pixel_units(fitter.psf_model, pixel_size)
    xc, yc = _dummy_corners(n, box)
    if variance_map is None:
        out = fit_arrays(fitter, arr, psf_pixels)
    else:
        out = fit_arrays(fitter, arr, psf_pixels, np.asarray(variance_map, dtype=F32), xc, yc)
    pv = pvalues(out.llr, box, len(fitter.psf_model))
    camera = IdealCamera.square(1024, 1024, pixel_size, dtype=np.float32)   # 0.1f0 (src/interface.jl:191,284)
    batch = ROIBatch(np.asarray(arr, dtype=F32), xc, yc, np.ones(n, dtype=np.int32), camera)
    loc = LocalizationResult.create(out, pv, batch, fitter.psf_model)
    return to_smld(loc, batch)


def _fit_batch(fitter: GaussMLEFitter, batch: ROIBatch) -> BasicSMLD:
    validate_fit_input(batch.data)
    box = batch.data.shape[0]
    psf_pixels = to_pixel_units(fitter.psf_model, batch.camera.pixel_size)
    if isinstance(batch.camera, SCMOSCamera):
        data = to_electrons(batch.data, batch.camera, batch.x_corners, batch.y_corners)
        out = fit_arrays(fitter, data, psf_pixels, variance_map(batch.camera), batch.x_corners, batch.y_corners)
    else:
        out = fit_arrays(fitter, batch.data, psf_pixels)
    pv = pvalues(out.llr, box, len(fitter.psf_model))
    loc = LocalizationResult.create(out, pv, batch, fitter.psf_model)
    return to_smld(loc, batch)


__all__ = [
    "GaussMLEFitter", "fit", "fit_arrays", "FitArrays", "pvalues", "validate_fit_input", "CPU", "GPU",
    "auto_device", "select_device", "GaussianXYNB", "GaussianXYNBS", "GaussianXYNBSXSY", "AstigmaticXYZNB",
]
