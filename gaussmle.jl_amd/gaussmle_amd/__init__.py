"""gaussmle_amd -- MI355X (gfx950) drop-in for GaussMLE.jl's per-ROI MLE fit.

The fit runs in libgaussmle_hip.so (hand-written HIP kernels behind a C ABI,
include/gaussmle.h); this package is the host-side mirror of GaussMLE.jl's
GaussMLEFitter / fit API (src/interface.jl) and its result conversion
(src/roi_batch.jl).
"""
from . import _lib, synth
from .boxer import find_boxes, getboxes
from ._lib import ArgumentError, GaussMLEError
from .fitter import (CPU, GPU, FitArrays, GaussMLEFitter, auto_device, fit, fit_arrays, pvalues, select_device,
                     validate_fit_input)
from .models import (AstigmaticXYZNB, GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, IdealCamera,
                     ParameterConstraints, PSFModel, ROIBatch, SCMOSCamera, default_constraints, to_electrons,
                     to_pixel_units, variance_map)
from .results import (BasicSMLD, Emitter2DFitGaussMLE, Emitter2DFitSigma, Emitter2DFitSigmaXY, Emitter3DFitGaussMLE,
                      LocalizationResult, to_emitter, to_smld)

__version__ = "0.1.0"

__all__ = [
    "GaussMLEFitter", "fit", "fit_arrays", "FitArrays", "pvalues", "validate_fit_input", "CPU", "GPU",
    "auto_device", "select_device", "GaussianXYNB", "GaussianXYNBS", "GaussianXYNBSXSY", "AstigmaticXYZNB",
    "PSFModel", "ParameterConstraints", "default_constraints", "to_pixel_units", "IdealCamera", "SCMOSCamera",
    "ROIBatch", "to_electrons", "variance_map", "BasicSMLD", "LocalizationResult", "to_emitter", "to_smld",
    "Emitter2DFitGaussMLE", "Emitter2DFitSigma", "Emitter2DFitSigmaXY", "Emitter3DFitGaussMLE", "ArgumentError",
    "GaussMLEError", "synth", "getboxes", "find_boxes",
]
