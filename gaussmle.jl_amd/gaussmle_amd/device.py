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
