"""The remaining scenarios of the reference's model validation
(test/model_validation_tests.jl, box 15), restated:

  - Astigmatic 3-D (:97-135, helper test/validation_utils_roibatch.jl): the
    Huang et al. calibration in microns (sigma0 0.13, Ax / Ay +-0.05, Bx / By
    +-0.01, gamma 0.2, d 0.5), x, y = 7.5 + 0.5 randn, z ~ U(-0.25, 0.25) um,
    N = 2000 U(0.8, 1.2), bg = 1 U(0.8, 1.2); simulated with
    generate_roi_batch on an ideal 0.1 um camera and fitted through
    fit(fitter, ROIBatch).  Checks (validation_utils_roibatch.jl:130-162):
    |1 - std/CRLB| < 0.25 for x, y, z, N, bg, and |bias| below 0.03 um (z),
    100 (N) and 2 (bg); x and y have no bias test there.
  - sCMOS camera (:208-281): readnoise 5 e- rms, offset 100 ADU, gain 0.5,
    qe 0.82; x, y = 7.5 + 0.5 randn, N = 1000 + 200 randn, bg = 10 + 2 randn;
    ROIs on a grid of the sensor.  Checks: photon bias against N qe below 50,
    |1 - std/CRLB| < 0.3 for N, and a mean N CRLB above 30.
  - Spots near the ROI edges (:395-431): x, y ~ 2 + 12 U, N 1000, bg 10,
    Poisson(max(0.01, mu)) pixels, fit(fitter, data): no NaN position and no
    infinite position uncertainty.  Also run on the Float32 oracle.

Differences, on purpose: numpy's seeded draws and the library's device
simulator instead of Julia's streams; 4,000 ROIs instead of 500-1,000 for
the statistical checks; a 2048 x 2048 sensor with ROIs on a grid, so that
every ROI lies on it.
"""
import numpy as np
import pytest
from scipy.special import erf

import gaussmle_amd as g
from gaussmle_amd import synth
from gaussmle_amd.models import (AstigmaticXYZNB, GaussianXYNB, IdealCamera, SCMOSCamera, default_constraints,
                                 to_pixel_units)

F32 = np.float32
BOX = 15
PIXEL = 0.1


def grid_corners(n, sensor=2048):
    per_row = (sensor - BOX + 1) // BOX
    i = np.arange(n)
    return np.stack([1 + (i % per_row) * BOX, 1 + (i // per_row) * BOX]).astype(np.int32)


def stats(fitted, truth, crlb):
    e = np.asarray(fitted, np.float64) - np.asarray(truth, np.float64)
    c = np.asarray(crlb, np.float64)
    std = e.std(ddof=1)
    return e.mean(), std, std / c[np.isfinite(c)].mean()


@pytest.mark.gpu
def test_astigmatic_validation():
    from gaussmle_amd.simulator import generate_roi_batch
    n = 4000
    psf = AstigmaticXYZNB(0.13, 0.13, 0.05, -0.05, 0.01, -0.01, 0.2, 0.5)
    rng = np.random.default_rng(42)
    truth = np.stack([BOX / 2 + 0.5 * rng.standard_normal(n), BOX / 2 + 0.5 * rng.standard_normal(n),
                      -0.25 + 0.5 * rng.random(n), 2000.0 * (0.8 + 0.4 * rng.random(n)),
                      1.0 * (0.8 + 0.4 * rng.random(n))]).astype(F32)
    camera = IdealCamera.square(2048, 2048, PIXEL, dtype=np.float32)
    batch = generate_roi_batch(camera, psf, n_rois=n, roi_size=BOX, true_params=truth, corners=grid_corners(n),
                               seed=42)
    smld = g.fit(g.GaussMLEFitter(psf_model=psf, device="gpu", iterations=20), batch)
    es = smld.emitters
    x_roi = np.array([e.x / PIXEL + 1 for e in es]) - batch.x_corners + 1
    y_roi = np.array([e.y / PIXEL + 1 for e in es]) - batch.y_corners + 1
    checks = {
        "x": (x_roi, truth[0], [e.σ_x / PIXEL for e in es], None),
        "y": (y_roi, truth[1], [e.σ_y / PIXEL for e in es], None),
        "z": ([e.z for e in es], truth[2], [e.σ_z for e in es], 0.03),
        "photons": ([e.photons for e in es], truth[3], [e.σ_photons for e in es], 100.0),
        "background": ([e.bg for e in es], truth[4], [e.σ_bg for e in es], 2.0),
    }
    for name, (fit, tru, crlb, bias_tol) in checks.items():
        bias, std, ratio = stats(fit, tru, crlb)
        print(f"[astig validation] {name}: bias {bias:+.4f} std {std:.4f} std/CRLB {ratio:.3f}", flush=True)
        assert abs(1.0 - ratio) < 0.25, name
        if bias_tol is not None:
            assert abs(bias) < bias_tol, name


@pytest.mark.gpu
def test_scmos_camera_model():
    from gaussmle_amd.simulator import generate_roi_batch
    n = 4000
    psf = GaussianXYNB(0.13)
    camera = SCMOSCamera.square(2048, 2048, PIXEL, readnoise=5.0, offset=100.0, gain=0.5, qe=0.82)
    rng = np.random.default_rng(42)
    truth = np.stack([BOX / 2 + 0.5 * rng.standard_normal(n), BOX / 2 + 0.5 * rng.standard_normal(n),
                      1000.0 + 200.0 * rng.standard_normal(n), 10.0 + 2.0 * rng.standard_normal(n)]).astype(F32)
    batch = generate_roi_batch(camera, psf, n_rois=n, roi_size=BOX, true_params=truth, corners=grid_corners(n),
                               seed=42)
    smld = g.fit(g.GaussMLEFitter(psf_model=psf, device="gpu"), batch)
    photons = np.array([e.photons for e in smld.emitters])
    crlb = np.array([e.σ_photons for e in smld.emitters])
    bias, std, ratio = stats(photons, truth[2] * F32(0.82), crlb)
    print(f"[scmos validation] photons: bias {bias:+.3f} std {std:.3f} std/CRLB {ratio:.3f} "
          f"mean CRLB {crlb.mean():.2f}", flush=True)
    assert abs(bias) < 50.0
    assert abs(1.0 - ratio) < 0.3
    assert crlb.mean() > 30.0


def edge_spots(n=1000, seed=42):
    """model_validation_tests.jl:398-418."""
    rng = np.random.default_rng(seed)
    x = 2.0 + (BOX - 3) * rng.random(n)
    y = 2.0 + (BOX - 3) * rng.random(n)
    k = np.arange(1, BOX + 1, dtype=np.float64)
    s = np.sqrt(2.0) * 1.3

    def phi(mu):
        return 0.5 * (erf((k[:, None] - mu[None, :] + 0.5) / s) - erf((k[:, None] - mu[None, :] - 0.5) / s))

    mu = 10.0 + 1000.0 * phi(y)[:, None, :] * phi(x)[None, :, :]
    return np.asfortranarray(rng.poisson(np.maximum(0.01, mu)).astype(F32))


def test_edge_spots_oracle():
    from oracle import refcpu
    pp = to_pixel_units(GaussianXYNB(0.13), PIXEL)
    data = edge_spots()
    w = synth.Workload("edges", pp, BOX, data, np.zeros((4, data.shape[2])), default_constraints(pp, 11), 20)
    r = refcpu.fit_workload(w, refcpu.FP32)
    assert not np.isnan(r["params"][:, :2]).any()
    assert not np.isinf(r["crlb"][:, :2]).any()


@pytest.mark.gpu
def test_edge_spots():
    smld = g.fit(g.GaussMLEFitter(psf_model=GaussianXYNB(0.13), device="gpu"), edge_spots())
    es = smld.emitters
    assert not np.isnan([e.x for e in es]).any() and not np.isnan([e.y for e in es]).any()
    assert not np.isinf([e.σ_x for e in es]).any() and not np.isinf([e.σ_y for e in es]).any()
