"""Synthetic ROI batches drawn on the GPU (GaussMLE.jl's generate_roi_batch).

generate_roi_batch(camera, psf_model; n_rois, roi_size, true_params, corners,
frame_indices, xy_variation, corner_mode, min_spacing, seed) mirrors
src/simulator.jl:35-108: PSF-specific default parameters with +-xy_variation
on x, y (:7-10, :113-130), corners placed at random with a minimum spacing, on
a grid or in clusters (:135-223), then every pixel drawn on the device by
gmle_simulate_device (Poisson for an ideal camera; Poisson, read noise and the
ADU conversion for sCMOS, :259-307).  Host-side draws (parameters, corners)
use numpy's seeded generator and the pixels use the library's counter-based
generator, so a seed reproduces a batch exactly here, though not Julia's
stream.  The returned ROIBatch is what `fit` takes.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .models import (AstigmaticXYZNB, GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, IdealCamera, PSFModel,
                     ROIBatch, SCMOSCamera, to_pixel_units)

F32 = np.float32

# src/simulator.jl:7-10
DEFAULT_PARAMS = {
    GaussianXYNB: [6.0, 6.0, 1000.0, 10.0],
    GaussianXYNBS: [6.0, 6.0, 1000.0, 10.0, 1.3],
    GaussianXYNBSXSY: [6.0, 6.0, 1000.0, 10.0, 1.3, 1.3],
    AstigmaticXYZNB: [6.0, 6.0, 0.0, 1000.0, 10.0],
}


def default_params(psf: PSFModel, n_rois: int, xy_variation: float, rng: np.random.Generator) -> np.ndarray:
    """(P, n) Float32: the model's defaults with x, y moved by up to
    +-xy_variation (src/simulator.jl:113-130)."""
    base = np.array(DEFAULT_PARAMS[type(psf)], dtype=F32)
    out = np.repeat(base[:, None], n_rois, axis=1)
    if xy_variation > 0:
        out[0] += ((rng.random(n_rois) - 0.5) * 2.0 * xy_variation).astype(F32)
        out[1] += ((rng.random(n_rois) - 0.5) * 2.0 * xy_variation).astype(F32)
    return out


def make_corners(n_rois: int, camera_size: tuple, roi_size: int, mode: str, min_spacing: int,
                 rng: np.random.Generator) -> np.ndarray:
    """(2, n) Int32 1-based corners (src/simulator.jl:135-223)."""
    nx, ny = camera_size
    max_x, max_y = nx - roi_size + 1, ny - roi_size + 1
    corners = np.zeros((2, n_rois), dtype=np.int32)
    if mode == "random":
        placed = []
        attempts, max_attempts = 0, n_rois * 100
        for i in range(n_rois):
            valid = False
            while not valid and attempts < max_attempts:
                attempts += 1
                x, y = int(rng.integers(1, max_x + 1)), int(rng.integers(1, max_y + 1))
                valid = all(not (abs(x - px) < min_spacing and abs(y - py) < min_spacing) for px, py in placed)
                if valid:
                    corners[:, i] = (x, y)
                    placed.append((x, y))
            if not valid:
                corners[:, i] = (rng.integers(1, max_x + 1), rng.integers(1, max_y + 1))
    elif mode == "grid":
        n_grid = int(np.ceil(np.sqrt(n_rois)))
        sx = max(min_spacing, (max_x - 1) // n_grid)
        sy = max(min_spacing, (max_y - 1) // n_grid)
        idx = 0
        for j in range(1, max_y + 1, sy):
            for i in range(1, max_x + 1, sx):
                if idx < n_rois:
                    corners[:, idx] = (i, j)
                    idx += 1
    elif mode == "clustered":
        n_clusters = max(1, n_rois // 10)
        per = n_rois // n_clusters
        idx = 0
        for _ in range(n_clusters):
            cx = int(rng.integers(roi_size, max_x - roi_size + 1))
            cy = int(rng.integers(roi_size, max_y - roi_size + 1))
            for _ in range(min(per, n_rois - idx)):
                ox, oy = rng.integers(-min_spacing, min_spacing + 1, 2)
                corners[:, idx] = (np.clip(cx + ox, 1, max_x), np.clip(cy + oy, 1, max_y))
                idx += 1
    else:
        raise ValueError(f"Unknown corner_mode: {mode}. Use 'random', 'grid', or 'clustered'")
    return corners


def _map_or_scalar(v, shape):
    a = np.asarray(v, dtype=F32)
    if a.ndim == 0:
        return None, float(a)
    if a.shape != shape:
        raise ValueError(f"camera map has shape {a.shape}, camera is {shape}")
    return np.asfortranarray(a).reshape(-1, order="F"), 0.0


def simulate_device(psf_pixels: PSFModel, box: int, d_params, seed: int, camera=None, d_xc=None, d_yc=None,
                    out=None, stream=None):
    """gmle_simulate_device on torch tensors: d_params (n, stride) Float32 on
    the device; returns (n * box * box,) Float32 in the fitting layout,
    enqueued on torch's current stream."""
    import torch

    lib = _lib.load()
    n, stride = d_params.shape
    dev = d_params.device
    if out is None:
        out = torch.empty(n * box * box, dtype=torch.float32, device=dev)
    scmos = isinstance(camera, SCMOSCamera)
    maps = {}
    ny = nx = 0
    scal = {"qe": 1.0, "readnoise": 0.0, "gain": 1.0, "offset": 0.0}
    if scmos:
        ny, nx = camera.shape
        for name in ("qe", "readnoise", "gain", "offset"):
            m, v = _map_or_scalar(getattr(camera, name), (ny, nx))
            scal[name] = v
            if m is not None:
                maps[name] = torch.from_numpy(m.copy()).to(dev)
    s = stream if stream is not None else torch.cuda.current_stream(dev)

    def ptr(t):
        return None if t is None else t.data_ptr()

    psf = np.ascontiguousarray(psf_pixels.psf_array(), dtype=F32)
    rc = lib.gmle_simulate_device(psf_pixels.model_id, _lib.fptr(psf), box, n, d_params.data_ptr(), stride,
                                  1 if scmos else 0, ptr(d_xc), ptr(d_yc), ptr(maps.get("qe")),
                                  ptr(maps.get("readnoise")), ptr(maps.get("gain")), ptr(maps.get("offset")), ny, nx,
                                  scal["qe"], scal["readnoise"], scal["gain"], scal["offset"], int(seed) & (2**64 - 1),
                                  out.data_ptr(), s.cuda_stream)
    _lib.check(rc)
    return out


def generate_roi_batch(camera: IdealCamera, psf_model: PSFModel, n_rois: int = 100, roi_size: int = 11,
                       true_params=None, corners=None, frame_indices=None, xy_variation: float = 1.0,
                       corner_mode: str = "random", min_spacing: int = 20, seed: int | None = None) -> ROIBatch:
    """src/simulator.jl:35-108 with the pixels drawn on the GPU."""
    import torch

    rng = np.random.default_rng(seed)
    pixel_size = camera.pixel_edges_x[1] - camera.pixel_edges_x[0]
    psf_pixels = to_pixel_units(psf_model, pixel_size)
    if true_params is None:
        params = default_params(psf_pixels, n_rois, xy_variation, rng)
    else:
        params = np.asarray(true_params, dtype=F32)
        n_rois = params.shape[1]
    ny, nx = camera.shape
    if corners is None:
        corners = make_corners(n_rois, (nx, ny), roi_size, corner_mode, min_spacing, rng)
    else:
        corners = np.asarray(corners, dtype=np.int32)
        if corners.shape != (2, n_rois):
            raise ValueError("Corners must be 2 x n_rois")
        max_x, max_y = nx - roi_size + 1, ny - roi_size + 1
        for i in range(n_rois):
            if not 1 <= corners[0, i] <= max_x:
                raise ValueError(f"ROI {i + 1}: x corner ({corners[0, i]}) must be in range [1, {max_x}] "
                                 f"for ROI size {roi_size} on camera width {nx}")
            if not 1 <= corners[1, i] <= max_y:
                raise ValueError(f"ROI {i + 1}: y corner ({corners[1, i]}) must be in range [1, {max_y}] "
                                 f"for ROI size {roi_size} on camera height {ny}")
    frames = np.ones(n_rois, np.int32) if frame_indices is None else np.asarray(frame_indices, np.int32)
    if len(frames) != n_rois:
        raise ValueError("Must have one frame index per ROI")
    dev = torch.device("cuda:0")
    d_params = torch.from_numpy(np.ascontiguousarray(params.T)).to(dev)
    d_xc = torch.from_numpy(np.ascontiguousarray(corners[0])).to(dev)
    d_yc = torch.from_numpy(np.ascontiguousarray(corners[1])).to(dev)
    s = int(rng.integers(0, 2**63)) if seed is None else int(seed)
    out = simulate_device(psf_pixels, roi_size, d_params, s, camera, d_xc, d_yc)
    data = out.cpu().numpy().reshape(n_rois, roi_size, roi_size).transpose(2, 1, 0)
    return ROIBatch(np.asfortranarray(data), corners[0].copy(), corners[1].copy(), frames, camera)
