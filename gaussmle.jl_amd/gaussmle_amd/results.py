"""Fit results -> emitters -> BasicSMLD.

Host mirror of GaussMLE.jl's src/roi_batch.jl: LocalizationResult (:9-74,
ROI -> camera coordinates :29-34), to_emitter per PSF model (:125-302) and
to_smld (:326-353).  Emitter field names follow SMLMData's emitter types.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .models import AstigmaticXYZNB, GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, PSFModel, ROIBatch

F32 = np.float32


@dataclass
class Emitter2DFitGaussMLE:
    x: float
    y: float
    photons: float
    bg: float
    σ_x: float
    σ_y: float
    σ_photons: float
    σ_bg: float
    pvalue: float
    frame: int
    dataset: int
    track_id: int
    id: int


@dataclass
class Emitter2DFitSigma:
    x: float
    y: float
    photons: float
    bg: float
    σ: float
    σ_x: float
    σ_y: float
    σ_photons: float
    σ_bg: float
    σ_σ: float
    pvalue: float
    frame: int
    dataset: int
    track_id: int
    id: int


@dataclass
class Emitter2DFitSigmaXY:
    x: float
    y: float
    photons: float
    bg: float
    σx: float
    σy: float
    σ_x: float
    σ_y: float
    σ_photons: float
    σ_bg: float
    σ_σx: float
    σ_σy: float
    pvalue: float
    frame: int
    dataset: int
    track_id: int
    id: int


@dataclass
class Emitter3DFitGaussMLE:
    x: float
    y: float
    z: float
    photons: float
    bg: float
    σ_x: float
    σ_y: float
    σ_z: float
    σ_photons: float
    σ_bg: float
    pvalue: float
    frame: int
    dataset: int
    track_id: int
    id: int


@dataclass
class LocalizationResult:
    parameters: np.ndarray       # (P, n), ROI coordinates
    uncertainties: np.ndarray    # (P, n)
    log_likelihoods: np.ndarray  # (n,)
    pvalues: np.ndarray          # (n,)
    x_camera: np.ndarray
    y_camera: np.ndarray
    frame_indices: np.ndarray
    x_roi_corners: np.ndarray
    y_roi_corners: np.ndarray
    psf_model: PSFModel
    n_fits: int

    @classmethod
    def create(cls, out, pvalues, batch: ROIBatch, psf_model: PSFModel) -> "LocalizationResult":
        """src/roi_batch.jl:37-74; camera = corner + roi - 1 (:29-34)."""
        params = out.params
        xc = batch.x_corners.astype(F32)
        yc = batch.y_corners.astype(F32)
        x_cam = (xc + params[0] - F32(1)).astype(F32)
        y_cam = (yc + params[1] - F32(1)).astype(F32)
        return cls(params, out.crlb, out.llr, pvalues, x_cam, y_cam, batch.frame_indices, batch.x_corners,
                   batch.y_corners, psf_model, params.shape[1])

    def __len__(self):
        return self.n_fits


@dataclass
class BasicSMLD:
    emitters: list
    camera: object
    n_frames: int
    n_datasets: int
    metadata: dict = field(default_factory=dict)

    def __len__(self) -> int:
        return len(self.emitters)


def to_emitter(result: LocalizationResult, idx: int, camera, dataset: int = 1, track_id: int = 0, id: int | None = None):
    """src/roi_batch.jl:121-302 (idx is 0-based here; id defaults to idx + 1)."""
    # pixel size in the edges' own type: Float32 values times a Float64 pixel
    # size are computed in Float64 and rounded to Float32 once (T(...))
    psx = camera.pixel_edges_x[1] - camera.pixel_edges_x[0]
    psy = camera.pixel_edges_y[1] - camera.pixel_edges_y[0]
    p = result.parameters[:, idx]
    u = result.uncertainties[:, idx]
    x_um = float(F32((result.x_camera[idx] - F32(1)) * psx))
    y_um = float(F32((result.y_camera[idx] - F32(1)) * psy))
    frame = int(result.frame_indices[idx])
    pv = float(result.pvalues[idx])
    eid = idx + 1 if id is None else id
    m = result.psf_model
    sx = lambda v: float(F32(v * psx))  # noqa: E731
    sy = lambda v: float(F32(v * psy))  # noqa: E731
    if isinstance(m, GaussianXYNB):
        return Emitter2DFitGaussMLE(x_um, y_um, float(p[2]), float(p[3]), sx(u[0]), sy(u[1]), float(u[2]),
                                    float(u[3]), pv, frame, dataset, track_id, eid)
    if isinstance(m, GaussianXYNBS):
        return Emitter2DFitSigma(x_um, y_um, float(p[2]), float(p[3]), sx(p[4]), sx(u[0]), sy(u[1]), float(u[2]),
                                 float(u[3]), sx(u[4]), pv, frame, dataset, track_id, eid)
    if isinstance(m, GaussianXYNBSXSY):
        return Emitter2DFitSigmaXY(x_um, y_um, float(p[2]), float(p[3]), sx(p[4]), sy(p[5]), sx(u[0]), sy(u[1]),
                                   float(u[2]), float(u[3]), sx(u[4]), sy(u[5]), pv, frame, dataset, track_id, eid)
    if isinstance(m, AstigmaticXYZNB):
        return Emitter3DFitGaussMLE(x_um, y_um, float(p[2]), float(p[3]), float(p[4]), sx(u[0]), sy(u[1]),
                                    float(u[2]), float(u[3]), float(u[4]), pv, frame, dataset, track_id, eid)
    raise TypeError(f"unknown PSF model {m!r}")


def _emitters(result: LocalizationResult, camera, dataset: int = 1, track_id: int = 0) -> list:
    """Every emitter of `result`, as to_emitter builds them one at a time: the
    same Float32 roundings, computed on whole columns first."""
    psx = camera.pixel_edges_x[1] - camera.pixel_edges_x[0]
    psy = camera.pixel_edges_y[1] - camera.pixel_edges_y[0]
    n = result.n_fits
    p, u = result.parameters, result.uncertainties

    def sx(v):
        return (v * psx).astype(F32).tolist()

    def sy(v):
        return (v * psy).astype(F32).tolist()

    def raw(v):
        return v.tolist()

    x_um = ((result.x_camera - F32(1)) * psx).astype(F32).tolist()
    y_um = ((result.y_camera - F32(1)) * psy).astype(F32).tolist()
    tail = (np.asarray(result.pvalues).tolist(), np.asarray(result.frame_indices).astype(int).tolist(),
            [dataset] * n, [track_id] * n, list(range(1, n + 1)))
    m = result.psf_model
    if isinstance(m, GaussianXYNB):
        cls, cols = Emitter2DFitGaussMLE, (x_um, y_um, raw(p[2]), raw(p[3]), sx(u[0]), sy(u[1]), raw(u[2]),
                                           raw(u[3]))
    elif isinstance(m, GaussianXYNBS):
        cls, cols = Emitter2DFitSigma, (x_um, y_um, raw(p[2]), raw(p[3]), sx(p[4]), sx(u[0]), sy(u[1]), raw(u[2]),
                                        raw(u[3]), sx(u[4]))
    elif isinstance(m, GaussianXYNBSXSY):
        cls, cols = Emitter2DFitSigmaXY, (x_um, y_um, raw(p[2]), raw(p[3]), sx(p[4]), sy(p[5]), sx(u[0]), sy(u[1]),
                                          raw(u[2]), raw(u[3]), sx(u[4]), sy(u[5]))
    elif isinstance(m, AstigmaticXYZNB):
        cls, cols = Emitter3DFitGaussMLE, (x_um, y_um, raw(p[2]), raw(p[3]), raw(p[4]), sx(u[0]), sy(u[1]),
                                           raw(u[2]), raw(u[3]), raw(u[4]))
    else:
        raise TypeError(f"unknown PSF model {m!r}")
    return [cls(*f) for f in zip(*cols, *tail)]


def to_smld(result: LocalizationResult, batch: ROIBatch, dataset: int = 1, metadata: dict | None = None) -> BasicSMLD:
    """src/roi_batch.jl:326-353."""
    emitters = _emitters(result, batch.camera, dataset=dataset)
    md = dict(metadata or {})
    md["psf_model"] = type(result.psf_model).__name__
    md["n_fits"] = result.n_fits
    n_frames = int(result.frame_indices.max()) if result.n_fits else 0
    return BasicSMLD(emitters, batch.camera, n_frames, 1, md)
