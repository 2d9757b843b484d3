"""ROI extraction on the device: the step before the fit.

GaussMLE.jl's pipeline gets its ROIBatch from SMLMBoxer.getboxes
(examples/full_pipeline_boxer.jl:70-80, examples/full_pipeline.jl:73-81):
difference-of-Gaussians filter, local maxima above `minval`, boxes that
overlap by at most `overlap` pixels, cut from the raw frames.  `getboxes`
here takes the same arguments and returns the ROIBatch that `fit` takes; the
work runs in the HIP library (gmle_find_boxes / gmle_find_boxes_device), whose
conventions are stated in include/gaussmle.h.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib
from .models import ROIBatch


def _stack(imagestack) -> np.ndarray:
    a = np.asarray(imagestack, dtype=np.float32)
    if a.ndim == 2:
        a = a[:, :, None]
    if a.ndim != 3:
        raise ValueError(f"imagestack must be (rows, cols[, frames]), got shape {a.shape}")
    return np.asfortranarray(a)


def suppression_radius(boxsize: int, overlap: float) -> int:
    """Detections closer than this + 1 pixels (Chebyshev) suppress each other."""
    return max(1, math.ceil(np.float32(boxsize) - np.float32(overlap)) - 1)


def max_detections(ny: int, nx: int, nf: int, boxsize: int, overlap: float) -> int:
    """Upper bound on detections: one per (r + 1) x (r + 1) cell of a frame."""
    c = suppression_radius(boxsize, overlap) + 1
    return nf * -(-ny // c) * -(-nx // c)


def find_boxes(imagestack, boxsize: int = 7, overlap: float = 2.0, sigma_small: float = 1.0,
               sigma_large: float = 2.0, minval: float = 0.0, max_boxes: int | None = None) -> dict:
    """Detections of a (rows, cols, frames) stack: rois (box, box, n),
    1-based x (column) / y (row) corners and frame indices, strength (the
    filtered value at each detection) and n (all detections found)."""
    a = _stack(imagestack)
    ny, nx, nf = a.shape
    cap = max_detections(ny, nx, nf, boxsize, overlap) if max_boxes is None else int(max_boxes)
    lib = _lib.load()
    rois = np.zeros((boxsize, boxsize, max(cap, 1)), np.float32, order="F")
    xc, yc, fr = (np.zeros(max(cap, 1), np.int32) for _ in range(3))
    st = np.zeros(max(cap, 1), np.float32)
    n = np.zeros(1, np.int64)
    rc = lib.gmle_find_boxes(_lib.fptr(a.reshape(-1, order="F")), ny, nx, nf, int(boxsize), float(overlap),
                             float(sigma_small), float(sigma_large), float(minval), cap,
                             _lib.fptr(rois.reshape(-1, order="F")), _lib.iptr(xc), _lib.iptr(yc), _lib.iptr(fr),
                             _lib.fptr(st), _lib.i64ptr(n))
    _lib.check(rc)
    m = min(int(n[0]), cap)
    return {"n": int(n[0]), "rois": np.asfortranarray(rois[:, :, :m]), "x_corners": xc[:m].copy(),
            "y_corners": yc[:m].copy(), "frames": fr[:m].copy(), "strength": st[:m].copy()}


def getboxes(imagestack, camera, boxsize: int = 7, overlap: float = 2.0, sigma_small: float = 1.0,
             sigma_large: float = 2.0, minval: float = 0.0, use_gpu: bool = True) -> ROIBatch:
    """SMLMBoxer.getboxes' call (examples/full_pipeline.jl:73-81) on the GPU,
    returning the ROIBatch `fit` takes.  There is no CPU path."""
    if not use_gpu:
        raise ValueError("getboxes runs on the GPU only (use_gpu=False is not served by this package)")
    r = find_boxes(imagestack, boxsize, overlap, sigma_small, sigma_large, minval)
    return ROIBatch(r["rois"], r["x_corners"], r["y_corners"], r["frames"], camera)


def find_boxes_device(d_frames, ny: int, nx: int, n_frames: int, boxsize: int = 7, overlap: float = 2.0,
                      sigma_small: float = 1.0, sigma_large: float = 2.0, minval: float = 0.0,
                      max_boxes: int | None = None, out: dict | None = None, stream=None) -> dict:
    """gmle_find_boxes_device on a float32 CUDA tensor holding the stack in
    column-major order (ny * nx * n_frames values), enqueued on torch's current
    stream.  Returns device tensors (rois, x/y corners, frames, strength,
    count, workspace); nothing is synchronized."""
    import torch

    lib = _lib.load()
    dev = d_frames.device
    cap = max_detections(ny, nx, n_frames, boxsize, overlap) if max_boxes is None else int(max_boxes)
    if out is None:
        ws = lib.gmle_boxer_workspace_bytes(ny, nx, n_frames)
        out = {
            "rois": torch.empty((max(cap, 1), boxsize * boxsize), dtype=torch.float32, device=dev),
            "x_corners": torch.empty(max(cap, 1), dtype=torch.int32, device=dev),
            "y_corners": torch.empty(max(cap, 1), dtype=torch.int32, device=dev),
            "frames": torch.empty(max(cap, 1), dtype=torch.int32, device=dev),
            "strength": torch.empty(max(cap, 1), dtype=torch.float32, device=dev),
            "count": torch.zeros(1, dtype=torch.int64, device=dev),
            "workspace": torch.empty(max(ws, 1), dtype=torch.uint8, device=dev),
        }
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    rc = lib.gmle_find_boxes_device(d_frames.data_ptr(), ny, nx, n_frames, int(boxsize), float(overlap),
                                    float(sigma_small), float(sigma_large), float(minval), cap,
                                    out["rois"].data_ptr(), out["x_corners"].data_ptr(), out["y_corners"].data_ptr(),
                                    out["frames"].data_ptr(), out["strength"].data_ptr(), out["count"].data_ptr(),
                                    out["workspace"].data_ptr(), s.cuda_stream)
    _lib.check(rc)
    return out
