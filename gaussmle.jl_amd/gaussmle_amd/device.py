"""Device-resident fits: gmle_fit_device on torch tensors already in HBM.

torch only supplies device memory and the stream; the computation is the HIP
library's.  Used by the benchmark and the GPU parity tests.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .models import PSFModel


def fit_device(psf_pixels: PSFModel, box: int, d_data, constraints, iterations: int, precision: str = "fp32",
               lanes: int = 0, d_var=None, ny: int = 0, nx: int = 0, d_xc=None, d_yc=None, out=None,
               stream=None):
    """Enqueue one fit on torch's current stream (or `stream`).

    d_data: float32 CUDA tensor of n*box*box values in the reference's layout.
    Returns dict(params=(n, P) tensor, crlb=(n, P), llr=(n,), crlb64, llr64).
    """
    import torch

    lib = _lib.load()
    n = d_data.numel() // (box * box)
    P = len(psf_pixels)
    dev = d_data.device
    if out is None:
        out = {
            "params": torch.empty((n, P), dtype=torch.float32, device=dev),
            "crlb": torch.empty((n, P), dtype=torch.float32, device=dev),
            "llr": torch.empty((n,), dtype=torch.float32, device=dev),
        }
        if precision == "mixed":
            out["crlb64"] = torch.empty((n, P), dtype=torch.float64, device=dev)
            out["llr64"] = torch.empty((n,), dtype=torch.float64, device=dev)
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    camera = 0 if d_var is None else 1
    c = constraints

    def ptr(t):
        return None if t is None else t.data_ptr()

    rc = lib.gmle_fit_device(
        psf_pixels.model_id, _lib.fptr(psf_pixels.psf_array()), box, n, ptr(d_data), camera, ptr(d_var), ny, nx,
        ptr(d_xc), ptr(d_yc), _lib.fptr(c.lower), _lib.fptr(c.upper), _lib.fptr(c.max_step), iterations,
        _lib.PRECISION_IDS[precision], lanes, ptr(out["params"]), ptr(out["crlb"]), ptr(out["llr"]),
        ptr(out.get("crlb64")), ptr(out.get("llr64")), s.cuda_stream)
    _lib.check(rc)
    return out


def upload_workload(w, device="cuda:0"):
    """Copy a synth.Workload's arrays into device tensors."""
    import torch

    d = {"data": torch.from_numpy(np.ascontiguousarray(w.flat_data())).to(device)}
    if w.var_map is not None:
        d["var"] = torch.from_numpy(np.asfortranarray(w.var_map).reshape(-1, order="F").copy()).to(device)
        d["xc"] = torch.from_numpy(w.x_corners).to(device)
        d["yc"] = torch.from_numpy(w.y_corners).to(device)
    return d


EPILOGUE_FIELDS = ("x_um", "y_um", "crlb_x_um", "crlb_y_um", "pvalue", "width_x_um", "width_y_um",
                   "crlb_width_x_um", "crlb_width_y_um", "chi2")


def epilogue_device(d_params, d_crlb, d_llr, pixel_size_x: float, pixel_size_y: float | None = None, model: int = 0,
                    box: int = 11, d_models=None, d_boxes=None, d_xc=None, d_yc=None, stream=None, out=None):
    """gmle_epilogue_device on torch tensors: (n, 10) Float32 per-emitter
    results (EPILOGUE_FIELDS) in microns with p-values, enqueued on torch's
    current stream.  d_params / d_crlb are (n, stride)."""
    import torch

    lib = _lib.load()
    n = d_llr.numel()
    stride = d_params.shape[1]
    if out is None:
        out = torch.empty((n, len(EPILOGUE_FIELDS)), dtype=torch.float32, device=d_llr.device)
    s = stream if stream is not None else torch.cuda.current_stream(d_llr.device)

    def ptr(t):
        return None if t is None else t.data_ptr()

    rc = lib.gmle_epilogue_device(n, model, box, ptr(d_models), ptr(d_boxes), stride, ptr(d_params), ptr(d_crlb),
                                  ptr(d_llr), ptr(d_xc), ptr(d_yc), float(pixel_size_x),
                                  float(pixel_size_x if pixel_size_y is None else pixel_size_y), ptr(out),
                                  s.cuda_stream)
    _lib.check(rc)
    return out
