"""Seeded synthetic ROI stacks for parity tests and benchmarks.

Follows the simulator semantics of GaussMLE.jl (src/simulator.jl:259-345):
each pixel's expected count is bg + N * Phi_x(j) * Phi_y(i) with the
pixel-integrated Gaussian Phi (src/gausslib/baselibrary.jl:8-14), drawn as
Poisson; the sCMOS path adds Gaussian read noise and converts to ADU with
per-pixel gain and offset.  Camera maps are indexed [row = y, col = x] at
(y_corner + i - 1, x_corner + j - 1), the kernel's convention
(src/unified_kernel.jl:277-282); the reference simulator indexes them
transposed (src/simulator.jl:292-304), which is not followed.

Expected values are computed in Float64 with numpy's PCG64 stream, so the same
seed gives the same stack on every machine.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.special import erf

from .models import (AstigmaticXYZNB, GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, ParameterConstraints,
                     PSFModel, default_constraints)

F32 = np.float32


def integrated_gaussian(n: int, mu: np.ndarray, sigma: np.ndarray) -> np.ndarray:
    """Phi(k; mu, sigma) for pixels k = 1..n (1-based), shape (len(mu), n), Float64."""
    k = np.arange(1, n + 1, dtype=np.float64)[None, :]
    mu = np.asarray(mu, dtype=np.float64)[:, None]
    s = np.sqrt(0.5 / np.asarray(sigma, dtype=np.float64) ** 2)[:, None]
    return 0.5 * (erf((k - mu + 0.5) * s) - erf((k - mu - 0.5) * s))


def astig_sigmas(z, psf: AstigmaticXYZNB):
    z = np.asarray(z, dtype=np.float64)
    ux = (z - psf.gamma) / psf.d
    uy = (z + psf.gamma) / psf.d
    ax = 1 + ux ** 2 + psf.Ax * ux ** 3 + psf.Bx * ux ** 4
    ay = 1 + uy ** 2 + psf.Ay * uy ** 3 + psf.By * uy ** 4
    return psf.sigma_x0 * np.sqrt(ax), psf.sigma_y0 * np.sqrt(ay)


def expected_counts(psf_pixels: PSFModel, theta: np.ndarray, box: int) -> np.ndarray:
    """Expected counts, (box, box, n) in the reference's [row, col, roi] layout
    (Fortran order, so ROI k is contiguous).  theta is (P, n) in pixel units."""
    th = np.asarray(theta, dtype=np.float64)
    if isinstance(psf_pixels, GaussianXYNB):
        sx = sy = np.full(th.shape[1], psf_pixels.sigma)
        x, y, N, bg = th[0], th[1], th[2], th[3]
    elif isinstance(psf_pixels, GaussianXYNBS):
        x, y, N, bg = th[0], th[1], th[2], th[3]
        sx = sy = th[4]
    elif isinstance(psf_pixels, GaussianXYNBSXSY):
        x, y, N, bg = th[0], th[1], th[2], th[3]
        sx, sy = th[4], th[5]
    else:
        x, y, N, bg = th[0], th[1], th[3], th[4]
        sx, sy = astig_sigmas(th[2], psf_pixels)
    px = integrated_gaussian(box, x, sx)   # (n, cols)
    py = integrated_gaussian(box, y, sy)   # (n, rows)
    mu = bg[:, None, None] + N[:, None, None] * px[:, :, None] * py[:, None, :]   # (n, col, row)
    return np.transpose(mu, (2, 1, 0))     # (row, col, n) view, Fortran-contiguous


@dataclass
class Workload:
    """One synthetic fitting problem in the reference's array layout."""

    name: str
    psf_pixels: PSFModel
    box: int
    data: np.ndarray            # (box, box, n) Float32, Fortran order; photons / electrons
    truth: np.ndarray           # (P, n) Float64
    constraints: ParameterConstraints
    iterations: int
    var_map: np.ndarray | None = None      # (ny, nx) Float32 (e-^2), sCMOS only
    x_corners: np.ndarray | None = None
    y_corners: np.ndarray | None = None

    @property
    def n(self) -> int:
        return self.data.shape[2]

    @property
    def camera(self) -> int:
        return 0 if self.var_map is None else 1

    def flat_data(self) -> np.ndarray:
        return self.data.reshape(-1, order="F")

    def subset(self, k: int) -> "Workload":
        d = np.asfortranarray(self.data[:, :, :k])
        return Workload(self.name, self.psf_pixels, self.box, d, self.truth[:, :k], self.constraints,
                        self.iterations, self.var_map, None if self.x_corners is None else self.x_corners[:k].copy(),
                        None if self.y_corners is None else self.y_corners[:k].copy())


def _poisson_stack(rng: np.random.Generator, mu: np.ndarray) -> np.ndarray:
    out = np.empty(mu.shape, dtype=F32, order="F")
    out[...] = rng.poisson(mu)
    return out


def config0(n: int = 10_000, seed: int = 0, box: int = 11, iterations: int = 20) -> Workload:
    """11x11 GaussianXYNB (sigma 1.3 px), Poisson, 20 iterations.  x, y within
    1.5 px of the ROI centre (n+1)/2, N ~ U(500, 5000), bg ~ U(1, 20)."""
    rng = np.random.default_rng(seed)
    psf = GaussianXYNB(1.3)
    c = (box + 1) / 2
    theta = np.stack([rng.uniform(c - 1.5, c + 1.5, n), rng.uniform(c - 1.5, c + 1.5, n),
                      rng.uniform(500, 5000, n), rng.uniform(1, 20, n)])
    data = _poisson_stack(rng, expected_counts(psf, theta, box))
    return Workload("xynb_ideal_11", psf, box, data, theta, default_constraints(psf, 11), iterations)


def scmos_maps(rng: np.random.Generator, ny: int, nx: int):
    """Offset 100 ADU, gain ~ U(0.4, 0.6) e-/ADU, read variance ~ LogNormal(ln 2, 0.5) e-^2."""
    offset = np.full((ny, nx), 100.0, dtype=F32)
    gain = rng.uniform(0.4, 0.6, (ny, nx)).astype(F32)
    var = rng.lognormal(np.log(2.0), 0.5, (ny, nx)).astype(F32)
    return offset, gain, var


def scmos_stack(rng, mu, var, gain, offset, xc, yc):
    """Electrons -> ADU -> electrons, as the camera and src/camera_models.jl:110-113 do."""
    box = mu.shape[0]
    ii = np.arange(box)
    nx = var.shape[1]
    # flat C-order index of map element [row = yc + i - 1, col = xc + j - 1]
    flat = (((yc.astype(np.int64) - 1)[None, None, :] + ii[:, None, None]) * nx
            + (xc.astype(np.int64) - 1)[None, None, :] + ii[None, :, None])
    v = np.take(var.ravel(), flat)
    g = np.take(gain.ravel(), flat)
    o = np.take(offset.ravel(), flat)
    e = rng.poisson(mu) + rng.standard_normal(mu.shape) * np.sqrt(v.astype(np.float64))
    adu = (e / g.astype(np.float64) + o.astype(np.float64)).astype(F32)
    out = np.empty(mu.shape, dtype=F32, order="F")
    out[...] = (adu - o) * g
    return out


def config1(n: int = 1_000_000, seed: int = 1, box: int = 13, cam: int = 256, iterations: int = 20) -> Workload:
    """13x13 GaussianXYNBS, sigma ~ U(1.0, 1.6), sCMOS on a 256x256 sensor with
    random ROI corners."""
    rng = np.random.default_rng(seed)
    psf = GaussianXYNBS()
    c = (box + 1) / 2
    theta = np.stack([rng.uniform(c - 1.5, c + 1.5, n), rng.uniform(c - 1.5, c + 1.5, n),
                      rng.uniform(500, 5000, n), rng.uniform(1, 20, n), rng.uniform(1.0, 1.6, n)])
    offset, gain, var = scmos_maps(rng, cam, cam)
    xc = rng.integers(1, cam - box + 2, n).astype(np.int32)
    yc = rng.integers(1, cam - box + 2, n).astype(np.int32)
    mu = expected_counts(psf, theta, box)
    data = scmos_stack(rng, mu, var, gain, offset, xc, yc)
    return Workload("xynbs_scmos_13", psf, box, data, theta, default_constraints(psf, 11), iterations, var, xc, yc)


ASTIG_CONFIG_PSF = AstigmaticXYZNB(1.3, 1.3, 0.05, 0.05, 0.3, 0.3, 50.0, 100.0)


def astig_constraints() -> ParameterConstraints:
    """Defaults (box 11) with z in [-600, 600] and a z step of 25 (= d/4); the
    reference's z defaults ([-1, 1], step 0.1) assume microns."""
    c = default_constraints(ASTIG_CONFIG_PSF, 11)
    lo, up, ms = c.lower.copy(), c.upper.copy(), c.max_step.copy()
    lo[2], up[2], ms[2] = -600, 600, 25
    return ParameterConstraints(lo, up, ms)


def config2(n: int = 4_000_000, seed: int = 2, box: int = 15, iterations: int = 30) -> Workload:
    """15x15 AstigmaticXYZNB (sx0 = sy0 = 1.3, A = 0.05, B = 0.3, gamma = 50,
    d = 100), z ~ U(-300, 300), N ~ U(1000, 10000), bg ~ U(2, 30), Poisson."""
    rng = np.random.default_rng(seed)
    psf = ASTIG_CONFIG_PSF
    c = (box + 1) / 2
    theta = np.stack([rng.uniform(c - 1.5, c + 1.5, n), rng.uniform(c - 1.5, c + 1.5, n),
                      rng.uniform(-300, 300, n), rng.uniform(1000, 10000, n), rng.uniform(2, 30, n)])
    data = _poisson_stack(rng, expected_counts(psf, theta, box))
    return Workload("astig_ideal_15", psf, box, data, theta, astig_constraints(), iterations)


def generic(psf_pixels: PSFModel, n: int, box: int, seed: int, iterations: int = 20, scmos: bool = False,
            cam: int = 64, constraints: ParameterConstraints | None = None, n_range=(500, 5000),
            bg_range=(1, 20), sigma_range=(1.0, 1.6), z_range=(-300, 300)) -> Workload:
    """Any model / box / camera with the config-0 style parameter draws."""
    rng = np.random.default_rng(seed)
    c = (box + 1) / 2
    x = rng.uniform(c - 1.5, c + 1.5, n)
    y = rng.uniform(c - 1.5, c + 1.5, n)
    N = rng.uniform(*n_range, n)
    bg = rng.uniform(*bg_range, n)
    if isinstance(psf_pixels, GaussianXYNB):
        theta = np.stack([x, y, N, bg])
    elif isinstance(psf_pixels, GaussianXYNBS):
        theta = np.stack([x, y, N, bg, rng.uniform(*sigma_range, n)])
    elif isinstance(psf_pixels, GaussianXYNBSXSY):
        theta = np.stack([x, y, N, bg, rng.uniform(*sigma_range, n), rng.uniform(*sigma_range, n)])
    else:
        theta = np.stack([x, y, rng.uniform(*z_range, n), N, bg])
    mu = expected_counts(psf_pixels, theta, box)
    if constraints is None:
        constraints = astig_constraints() if isinstance(psf_pixels, AstigmaticXYZNB) else default_constraints(
            psf_pixels, 11)
    name = f"{type(psf_pixels).__name__}_{'scmos' if scmos else 'ideal'}_{box}"
    if not scmos:
        return Workload(name, psf_pixels, box, _poisson_stack(rng, mu), theta, constraints, iterations)
    offset, gain, var = scmos_maps(rng, cam, cam)
    xc = rng.integers(1, cam - box + 2, n).astype(np.int32)
    yc = rng.integers(1, cam - box + 2, n).astype(np.int32)
    data = scmos_stack(rng, mu, var, gain, offset, xc, yc)
    return Workload(name, psf_pixels, box, data, theta, constraints, iterations, var, xc, yc)


def low_background(psf_pixels: PSFModel, n: int, box: int, seed: int, iterations: int = 20,
                   bg_range=(0.002, 0.04), log_n_range=(200.0, 1.0e5), hot_fraction: float = 0.25,
                   hot_range=(200.0, 2000.0), bg_lower: float = 1e-4,
                   sigma_range=(1.0, 1.6)) -> Workload:
    """Poisson ROIs on which the reference's likelihood branches bind
    (src/camera_models.jl:23-32): background U(bg_range) photons/px (below the
    0.01 model threshold in the tails), N log-uniform in log_n_range, and in a
    hot_fraction of the ROIs one hot pixel away from the spot (+U(hot_range)
    counts, as a cosmic-ray hit), where data/model exceeds the 1e4 caps.  The
    background's lower bound is bg_lower (the default constraints' 0.01 would
    clamp it above the threshold)."""
    rng = np.random.default_rng(seed)
    c = (box + 1) / 2
    x = rng.uniform(c - 1.5, c + 1.5, n)
    y = rng.uniform(c - 1.5, c + 1.5, n)
    N = np.exp(rng.uniform(np.log(log_n_range[0]), np.log(log_n_range[1]), n))
    bg = rng.uniform(*bg_range, n)
    if isinstance(psf_pixels, GaussianXYNB):
        theta = np.stack([x, y, N, bg])
    elif isinstance(psf_pixels, GaussianXYNBS):
        theta = np.stack([x, y, N, bg, rng.uniform(*sigma_range, n)])
    elif isinstance(psf_pixels, GaussianXYNBSXSY):
        theta = np.stack([x, y, N, bg, rng.uniform(*sigma_range, n), rng.uniform(*sigma_range, n)])
    else:
        theta = np.stack([x, y, rng.uniform(-300, 300, n), N, bg])
    data = _poisson_stack(rng, expected_counts(psf_pixels, theta, box))
    hot = np.flatnonzero(rng.random(n) < hot_fraction)
    # an interior pixel (the border ring sets the initial background) at least
    # 3 px from the spot centre, where the model is near the background
    for k in hot:
        while True:
            i, j = rng.integers(1, box - 1, 2)
            if max(abs(i + 1 - theta[1, k]), abs(j + 1 - theta[0, k])) >= 3:
                break
        data[i, j, k] += F32(np.floor(rng.uniform(*hot_range)))
    if isinstance(psf_pixels, AstigmaticXYZNB):
        cons = astig_constraints()
    else:
        cons = default_constraints(psf_pixels, 11)
    ibg = 4 if isinstance(psf_pixels, AstigmaticXYZNB) else 3
    lower = np.array(cons.lower, dtype=F32)
    lower[ibg] = F32(bg_lower)
    cons = ParameterConstraints(lower, np.array(cons.upper, dtype=F32), np.array(cons.max_step, dtype=F32))
    name = f"{type(psf_pixels).__name__}_lowbg_{box}"
    return Workload(name, psf_pixels, box, data, theta, cons, iterations)


def spot_frames(ny: int, nx: int, nf: int, spots_per_frame: int, seed: int = 5, bg: float = 10.0,
                photons=(500.0, 3000.0), sigma: float = 1.3, edge: int = 4, window: int = 8):
    """Camera frames (ny, nx, nf) Float32, column-major, in photons: a flat
    background plus pixel-integrated Gaussian spots (each added within +-window
    pixels of its centre), with Gaussian noise of variance equal to the mean
    (the Poisson limit).  Returns (frames, truth) with truth rows (frame, y, x,
    photons), 0-based pixel coordinates."""
    rng = np.random.default_rng(seed)
    frames = np.empty((ny, nx, nf), dtype=F32, order="F")
    truth = []
    k = np.arange(-window, window + 2, dtype=np.float64) - 0.5
    for f in range(nf):
        mu = np.full((ny, nx), bg, dtype=np.float64)
        ys = rng.uniform(edge, ny - 1 - edge, spots_per_frame)
        xs = rng.uniform(edge, nx - 1 - edge, spots_per_frame)
        Ns = rng.uniform(*photons, spots_per_frame)
        for y, x, N in zip(ys, xs, Ns):
            iy, ix = int(round(y)), int(round(x))
            py = np.diff(0.5 * erf((iy + k - y) / (np.sqrt(2.0) * sigma)))
            px = np.diff(0.5 * erf((ix + k - x) / (np.sqrt(2.0) * sigma)))
            r0, r1 = max(iy - window, 0), min(iy + window + 1, ny)
            c0, c1 = max(ix - window, 0), min(ix + window + 1, nx)
            mu[r0:r1, c0:c1] += N * np.outer(py[r0 - iy + window:r1 - iy + window],
                                             px[c0 - ix + window:c1 - ix + window])
            truth.append((f, y, x, N))
        noise = rng.standard_normal((ny, nx), dtype=np.float32)
        frames[:, :, f] = (mu + np.sqrt(mu) * noise).astype(F32)
    return frames, np.array(truth)
