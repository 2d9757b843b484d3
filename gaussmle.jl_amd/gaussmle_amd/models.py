"""PSF models, parameter constraints and camera descriptions.

Host-side mirror of GaussMLE.jl's model types (src/psf_models.jl:13-161),
to_pixel_units (src/psf_models.jl:311-331), ParameterConstraints /
default_constraints (src/constraints.jl:6-47) and the SMLMData camera types the
fit methods dispatch on (IdealCamera / SCMOSCamera), plus the sCMOS
preprocessing of src/camera_models.jl:103-124 and src/interface.jl:42-55.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

F32 = np.float32


# ---------------------------------------------------------------------------
# PSF models (spatial parameters in microns, as in the reference)
# ---------------------------------------------------------------------------
class PSFModel:
    """Abstract PSF model; ``len(model)`` is the number of fitted parameters."""

    n_params: int = 0
    model_id: int = -1
    param_names: tuple = ()

    def __len__(self) -> int:
        return self.n_params

    def psf_array(self) -> np.ndarray:
        """The 8-float PSF argument of the C ABI (pixel units expected)."""
        return np.zeros(8, dtype=F32)


@dataclass(frozen=True)
class GaussianXYNB(PSFModel):
    """Fixed-width Gaussian, parameters [x, y, N, bg] (src/psf_models.jl:39-41)."""

    sigma: float = 0.13
    n_params = 4
    model_id = 0
    param_names = ("x", "y", "photons", "bg")

    def psf_array(self) -> np.ndarray:
        a = np.zeros(8, dtype=F32)
        a[0] = F32(self.sigma)
        return a


@dataclass(frozen=True)
class GaussianXYNBS(PSFModel):
    """Fitted width, parameters [x, y, N, bg, sigma] (src/psf_models.jl:64)."""

    n_params = 5
    model_id = 1
    param_names = ("x", "y", "photons", "bg", "sigma")


@dataclass(frozen=True)
class GaussianXYNBSXSY(PSFModel):
    """Separate x / y widths, [x, y, N, bg, sigma_x, sigma_y] (src/psf_models.jl:91)."""

    n_params = 6
    model_id = 2
    param_names = ("x", "y", "photons", "bg", "sigma_x", "sigma_y")


@dataclass(frozen=True)
class AstigmaticXYZNB(PSFModel):
    """Astigmatic 3-D PSF, parameters [x, y, z, N, bg] (src/psf_models.jl:141-154).

    sigma_x(z) = sigma_x0 sqrt(1 + u^2 + Ax u^3 + Bx u^4), u = (z - gamma)/d, and
    sigma_y likewise with u = (z + gamma)/d.
    """

    sigma_x0: float
    sigma_y0: float
    Ax: float
    Ay: float
    Bx: float
    By: float
    gamma: float
    d: float
    n_params = 5
    model_id = 3
    param_names = ("x", "y", "z", "photons", "bg")

    def psf_array(self) -> np.ndarray:
        return np.array([self.sigma_x0, self.sigma_y0, self.Ax, self.Ay, self.Bx, self.By, self.gamma, self.d],
                        dtype=F32)


def to_pixel_units(psf: PSFModel, pixel_size: float) -> PSFModel:
    """Convert lateral widths from microns to pixels (src/psf_models.jl:311-331).

    gamma and d of the astigmatic model stay in their own units (z is never
    scaled by the pixel size, src/psf_models.jl:328-329).
    """
    px = F32(pixel_size)
    if isinstance(psf, GaussianXYNB):
        return GaussianXYNB(float(F32(psf.sigma) / px))
    if isinstance(psf, AstigmaticXYZNB):
        return AstigmaticXYZNB(float(F32(psf.sigma_x0) / px), float(F32(psf.sigma_y0) / px), psf.Ax, psf.Ay,
                               psf.Bx, psf.By, psf.gamma, psf.d)
    return psf


# ---------------------------------------------------------------------------
# Constraints (src/constraints.jl)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class ParameterConstraints:
    """Box bounds and per-iteration step limits (src/constraints.jl:6-10)."""

    lower: np.ndarray
    upper: np.ndarray
    max_step: np.ndarray

    def __post_init__(self):
        for name in ("lower", "upper", "max_step"):
            object.__setattr__(self, name, np.ascontiguousarray(getattr(self, name), dtype=F32))
        n = len(self.lower)
        if len(self.upper) != n or len(self.max_step) != n:
            raise ValueError("lower, upper and max_step must have the same length")

    def __len__(self) -> int:
        return len(self.lower)


def default_constraints(psf: PSFModel, box_size: int) -> ParameterConstraints:
    """Per-model defaults (src/constraints.jl:16-47).  The fitter always passes
    box_size = 11, whatever the ROI size (src/interface.jl:141)."""
    inf = np.inf
    hi = box_size + 2
    if isinstance(psf, GaussianXYNB):
        return ParameterConstraints([-2, -2, 1, 0.01], [hi, hi, inf, inf], [1, 1, 100, 2])
    if isinstance(psf, GaussianXYNBS):
        return ParameterConstraints([-2, -2, 1, 0.01, 0.3], [hi, hi, inf, inf, 5], [1, 1, 100, 2, 0.5])
    if isinstance(psf, GaussianXYNBSXSY):
        return ParameterConstraints([-2, -2, 1, 0.01, 0.3, 0.3], [hi, hi, inf, inf, 5, 5],
                                    [1, 1, 100, 2, 0.5, 0.5])
    if isinstance(psf, AstigmaticXYZNB):
        return ParameterConstraints([-2, -2, -1, 1, 0.01], [hi, hi, 1, inf, inf], [1, 1, 0.1, 100, 2])
    raise TypeError(f"unknown PSF model {psf!r}")


# ---------------------------------------------------------------------------
# Cameras (the SMLMData types the reference's fit methods dispatch on)
# ---------------------------------------------------------------------------
@dataclass
class IdealCamera:
    """Poisson-only camera with pixel edges in microns."""

    pixel_edges_x: np.ndarray
    pixel_edges_y: np.ndarray

    @classmethod
    def square(cls, nx: int, ny: int, pixel_size: float, dtype=np.float64) -> "IdealCamera":
        """Edges 0, ps, 2 ps, ... in `dtype` (the reference's pixel size type)."""
        ps = np.asarray(pixel_size, dtype=dtype)
        return cls(np.arange(nx + 1, dtype=dtype) * ps, np.arange(ny + 1, dtype=dtype) * ps)

    @property
    def pixel_size(self) -> float:
        return float(self.pixel_edges_x[1] - self.pixel_edges_x[0])

    @property
    def shape(self) -> tuple:
        return (len(self.pixel_edges_y) - 1, len(self.pixel_edges_x) - 1)


@dataclass
class SCMOSCamera(IdealCamera):
    """sCMOS camera: offset (ADU), gain (e-/ADU), readnoise (e- rms) and qe, each
    a scalar or a (ny, nx) map indexed [row = y, col = x]."""

    offset: float | np.ndarray = 100.0
    gain: float | np.ndarray = 1.0
    readnoise: float | np.ndarray = 1.0
    qe: float | np.ndarray = 1.0

    @classmethod
    def square(cls, nx: int, ny: int, pixel_size: float, readnoise=1.0, offset=100.0, gain=1.0,
               qe=1.0) -> "SCMOSCamera":
        return cls(np.arange(nx + 1, dtype=np.float64) * pixel_size, np.arange(ny + 1, dtype=np.float64) * pixel_size,
                   offset=offset, gain=gain, readnoise=readnoise, qe=qe)


def variance_map(camera: SCMOSCamera) -> np.ndarray:
    """readnoise^2 as a full (ny, nx) Float32 map (src/interface.jl:43-55)."""
    ny, nx = camera.shape
    v = np.asarray(camera.readnoise, dtype=np.float64) ** 2
    if v.ndim == 0:
        return np.full((ny, nx), F32(v), dtype=F32)
    if v.shape != (ny, nx):
        raise ValueError(f"readnoise map has shape {v.shape}, camera is {(ny, nx)}")
    return v.astype(F32)


def to_electrons(data_adu: np.ndarray, camera: SCMOSCamera, x_corners=None, y_corners=None) -> np.ndarray:
    """(ADU - offset) * gain (src/camera_models.jl:110-113).

    data_adu has the reference's (box, box, n) layout.  Scalar offset / gain
    broadcast as in the reference; per-pixel maps are gathered at each ROI's
    camera pixels through its corners (which the reference's broadcast cannot
    do).
    """
    d = np.asarray(data_adu, dtype=F32)
    off = np.asarray(camera.offset, dtype=F32)
    gain = np.asarray(camera.gain, dtype=F32)
    if off.ndim == 0 and gain.ndim == 0:
        return (d - off) * gain
    if x_corners is None or y_corners is None:
        raise ValueError("per-pixel offset / gain maps need ROI corners")
    box = d.shape[0]
    ii = np.arange(box)
    rows = (np.asarray(y_corners, dtype=np.int64) - 1)[None, None, :] + ii[:, None, None]
    cols = (np.asarray(x_corners, dtype=np.int64) - 1)[None, None, :] + ii[None, :, None]
    off_r = off if off.ndim == 0 else off[rows, cols]
    gain_r = gain if gain.ndim == 0 else gain[rows, cols]
    return ((d - off_r) * gain_r).astype(F32)


@dataclass
class ROIBatch:
    """A stack of ROIs cut from camera frames (SMLMData.ROIBatch).

    data: (box, box, n) indexed [row, col, roi]; x/y_corners: 1-based camera
    coordinates of each ROI's first pixel; frame_indices: 1-based frames.
    """

    data: np.ndarray
    x_corners: np.ndarray
    y_corners: np.ndarray
    frame_indices: np.ndarray
    camera: IdealCamera

    def __post_init__(self):
        self.data = np.asarray(self.data, dtype=F32)
        n = self.data.shape[2]
        self.x_corners = np.ascontiguousarray(self.x_corners, dtype=np.int32)
        self.y_corners = np.ascontiguousarray(self.y_corners, dtype=np.int32)
        if self.frame_indices is None:
            self.frame_indices = np.ones(n, dtype=np.int32)
        self.frame_indices = np.ascontiguousarray(self.frame_indices, dtype=np.int32)
        if not (len(self.x_corners) == len(self.y_corners) == len(self.frame_indices) == n):
            raise ValueError("one corner pair and frame index per ROI")

    def __len__(self) -> int:
        return self.data.shape[2]

    @property
    def roi_size(self) -> int:
        """Side of the square ROIs (SMLMData.ROIBatch.roi_size)."""
        return self.data.shape[0]
