"""The reference's local performance benchmark (test/local_performance_benchmark.jl,
run by test/runtests.jl:60-95), restated on the GPU.

Every PSF model x camera: 10,000 ROIs of 11 x 11 drawn by generate_roi_batch
(x, y = 6 + 0.3 randn, N = 1000 + 100 randn, bg = 10 + randn; sigma models
1.3 + 0.1 randn; astigmatic z = 0.2 randn um with sigma0 0.13 um, no
aberration terms, gamma 0.25 um, d 0.40 um; :129-170), fitted with
fit(fitter, ROIBatch) at 20 iterations after a warm-up fit (:172-208).  The
ideal camera is 0.1 um pixels; the sCMOS one adds 5 e- rms read noise,
offset 100 ADU, gain 0.5 e-/ADU and qe 0.82 (:106-124).  The check is
runtests.jl's: every parameter's std(fitted - true) / mean(CRLB) within
[0.8, 1.2], with fitted values compared to the drawn truth as the
benchmark does (for sCMOS the fit measures detected electrons, N qe, so
its N and bg errors also carry the spread of the truth times 1 - qe).

Differences, on purpose: numpy's seeded draws; ROIs on a grid of a
2048 x 2048 sensor instead of random placement on 512 x 512, so that every
ROI lies on the sensor without overlap.  fits/s of the whole fit() call
(device fit plus the host conversion to emitters) is printed, not checked.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import gaussmle_amd as g
from gaussmle_amd.models import (AstigmaticXYZNB, GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, IdealCamera,
                                 SCMOSCamera)

BOX = 11
N_ROIS = 10_000
PIXEL = 0.1
PSFS = {
    "GaussianXYNB": lambda: GaussianXYNB(0.13),
    "GaussianXYNBS": GaussianXYNBS,
    "GaussianXYNBSXSY": GaussianXYNBSXSY,
    "AstigmaticXYZNB": lambda: AstigmaticXYZNB(0.13, 0.13, 0.0, 0.0, 0.0, 0.0, 0.25, 0.40),
}


def camera_of(kind):
    if kind == "ideal":
        return IdealCamera.square(2048, 2048, PIXEL)
    return SCMOSCamera.square(2048, 2048, PIXEL, readnoise=5.0, offset=100.0, gain=0.5, qe=0.82)


def grid_corners(n, sensor=2048):
    per_row = (sensor - BOX + 1) // BOX
    i = np.arange(n)
    return np.stack([1 + (i % per_row) * BOX, 1 + (i // per_row) * BOX]).astype(np.int32)


def truth_of(psf, n, seed):
    """local_performance_benchmark.jl:129-158."""
    rng = np.random.default_rng(seed)
    c = (BOX + 1) / 2
    base = [c + 0.3 * rng.standard_normal(n), c + 0.3 * rng.standard_normal(n),
            1000.0 + 100.0 * rng.standard_normal(n), 10.0 + 1.0 * rng.standard_normal(n)]
    if isinstance(psf, GaussianXYNBS):
        base.append(1.3 + 0.1 * rng.standard_normal(n))
    elif isinstance(psf, GaussianXYNBSXSY):
        base += [1.3 + 0.1 * rng.standard_normal(n), 1.3 + 0.1 * rng.standard_normal(n)]
    elif isinstance(psf, AstigmaticXYZNB):
        base = base[:2] + [0.2 * rng.standard_normal(n)] + base[2:]
    return np.stack(base).astype(np.float32)


def columns(psf, smld, batch):
    """Fitted values and CRLBs per parameter in the benchmark's units
    (:210-279): positions and widths in pixels, z in um."""
    es = smld.emitters
    x = np.array([e.x / PIXEL for e in es]) - batch.x_corners + 2
    y = np.array([e.y / PIXEL for e in es]) - batch.y_corners + 2
    cols = [(x, [e.σ_x / PIXEL for e in es]), (y, [e.σ_y / PIXEL for e in es])]
    if isinstance(psf, AstigmaticXYZNB):
        cols.append(([e.z for e in es], [e.σ_z for e in es]))
    cols += [([e.photons for e in es], [e.σ_photons for e in es]), ([e.bg for e in es], [e.σ_bg for e in es])]
    if isinstance(psf, GaussianXYNBS):
        cols.append(([e.σ / PIXEL for e in es], [e.σ_σ / PIXEL for e in es]))
    elif isinstance(psf, GaussianXYNBSXSY):
        cols += [([e.σx / PIXEL for e in es], [e.σ_σx / PIXEL for e in es]),
                 ([e.σy / PIXEL for e in es], [e.σ_σy / PIXEL for e in es])]
    return cols


@pytest.mark.parametrize("camera", ["ideal", "scmos"])
@pytest.mark.parametrize("model", list(PSFS))
def test_std_over_crlb(model, camera):
    from gaussmle_amd.simulator import generate_roi_batch
    psf = PSFS[model]()
    cam = camera_of(camera)
    fitter = g.GaussMLEFitter(psf_model=psf, device="gpu", iterations=20)
    warm = generate_roi_batch(cam, psf, n_rois=100, roi_size=BOX, true_params=truth_of(psf, 100, 42),
                              corners=grid_corners(100), seed=42)
    g.fit(fitter, warm)
    truth = truth_of(psf, N_ROIS, 43)
    batch = generate_roi_batch(cam, psf, n_rois=N_ROIS, roi_size=BOX, true_params=truth,
                               corners=grid_corners(N_ROIS), seed=43)
    t0 = time.perf_counter()
    smld = g.fit(fitter, batch)
    fits_per_s = N_ROIS / (time.perf_counter() - t0)
    ratios = []
    for k, (fitted, crlb) in enumerate(columns(psf, smld, batch)):
        err = np.asarray(fitted, np.float64) - truth[k].astype(np.float64)
        ratios.append(err.std(ddof=1) / np.mean(np.asarray(crlb, np.float64)))
    print(f"[local benchmark] {model}-{camera}: fit() {fits_per_s:,.0f} fits/s, std/CRLB "
          + " ".join(f"{r:.3f}" for r in ratios), flush=True)
    assert fits_per_s > 0
    for k, r in enumerate(ratios):
        if np.isfinite(r):
            assert 0.8 <= r <= 1.2, f"{model}-{camera} parameter {k}: std/CRLB {r}"
