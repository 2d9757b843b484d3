"""Host-side mirror of GaussMLE.jl's API (CPU only: no kernel runs here)."""
import math

import numpy as np
import pytest
from scipy import special, stats

import gaussmle_amd as g
from gaussmle_amd import fitter as F
from gaussmle_amd import models, results, synth


def test_psf_models_and_pixel_units():
    assert [len(m) for m in (g.GaussianXYNB(0.13), g.GaussianXYNBS(), g.GaussianXYNBSXSY(),
                             g.AstigmaticXYZNB(0.13, 0.13, 0.05, -0.05, 0.01, -0.01, 0.2, 0.5))] == [4, 5, 6, 5]
    p = g.to_pixel_units(g.GaussianXYNB(0.13), 0.1)
    assert p.sigma == pytest.approx(1.3, rel=1e-6)
    a = g.to_pixel_units(g.AstigmaticXYZNB(0.13, 0.15, 0.05, -0.05, 0.01, -0.01, 0.2, 0.5), 0.1)
    assert (a.sigma_x0, a.sigma_y0) == (pytest.approx(1.3, rel=1e-6), pytest.approx(1.5, rel=1e-6))
    assert (a.gamma, a.d) == (0.2, 0.5)   # z units are never scaled (src/psf_models.jl:328-329)
    assert g.to_pixel_units(g.GaussianXYNBS(), 0.1) == g.GaussianXYNBS()


def test_pvalues_follow_chi2():
    llr = np.array([-50.0, -60.0, -200.0, 0.0], np.float32)
    p = g.pvalues(llr, 11, 4)
    expect = (np.float32(1) - stats.chi2.cdf(-2.0 * llr.astype(np.float64), 117).astype(np.float32))
    np.testing.assert_array_equal(p, expect)
    np.testing.assert_allclose(p[:2], stats.chi2.sf(-2.0 * llr[:2], 117), rtol=1e-5)
    assert p.dtype == np.float32 and p[3] == 1.0


def test_validate_fit_input():
    with pytest.raises(g.ArgumentError, match="empty"):
        g.validate_fit_input(np.zeros((11, 11, 0), np.float32))
    with pytest.raises(g.ArgumentError, match="square"):
        g.validate_fit_input(np.zeros((11, 9, 3), np.float32))
    with pytest.raises(g.ArgumentError, match="at least 3"):
        g.validate_fit_input(np.zeros((2, 2, 3), np.float32))
    bad = np.ones((7, 7, 2), np.float32)
    bad[3, 3, 1] = np.nan
    with pytest.raises(g.ArgumentError, match="NaN"):
        g.validate_fit_input(bad)
    g.validate_fit_input(-np.ones((7, 7, 2), np.float32))   # negatives allowed (sCMOS)


def test_device_selection_never_falls_back_to_cpu():
    with pytest.raises(ValueError, match="CPU path"):
        g.select_device("cpu")
    if g._lib.device_count() == 0:
        with pytest.raises(RuntimeError):
            g.select_device("gpu")
        with pytest.raises(RuntimeError):
            g.GaussMLEFitter()


def test_to_electrons_scalar_and_maps():
    cam = models.SCMOSCamera.square(32, 32, 0.1, readnoise=1.5, offset=100.0, gain=0.5)
    adu = np.full((5, 5, 2), 120.0, np.float32)
    np.testing.assert_allclose(g.to_electrons(adu, cam), 10.0)
    np.testing.assert_allclose(g.variance_map(cam), 2.25)
    ii, jj = np.mgrid[0:32, 0:32]
    cam2 = models.SCMOSCamera.square(32, 32, 0.1, readnoise=1.0, offset=(100 + ii).astype(np.float32),
                                     gain=(1 + 0.01 * jj).astype(np.float32))
    xc = np.array([3, 20], np.int32)
    yc = np.array([10, 1], np.int32)
    e = g.to_electrons(adu, cam2, xc, yc)
    # pixel (row i, col j) of ROI k reads camera [yc+i-1, xc+j-1] (0-based here)
    for k in range(2):
        for i in range(5):
            for j in range(5):
                r, c = yc[k] - 1 + i, xc[k] - 1 + j
                assert e[i, j, k] == pytest.approx((120 - (100 + r)) * (1 + 0.01 * c), rel=1e-6)


def test_expected_counts_layout():
    th = np.array([[3.0], [8.0], [1000.0], [2.0]])
    mu = synth.expected_counts(g.GaussianXYNB(1.3), th, 11)
    assert mu.shape == (11, 11, 1)
    phi = lambda k, m: 0.5 * (special.erf((k - m + 0.5) / (math.sqrt(2) * 1.3)) -
                              special.erf((k - m - 0.5) / (math.sqrt(2) * 1.3)))
    assert mu[7, 2, 0] == pytest.approx(2 + 1000 * phi(3, 3.0) * phi(8, 8.0))   # row 8 (y), col 3 (x)
    flat = mu.reshape(-1, order="F")
    assert flat[7 + 11 * 2] == mu[7, 2, 0]   # Julia layout: (i-1) + n (j-1)


@pytest.mark.parametrize("psf", [g.GaussianXYNB(0.13), g.GaussianXYNBS(), g.GaussianXYNBSXSY(),
                                 g.AstigmaticXYZNB(0.13, 0.13, 0.05, -0.05, 0.01, -0.01, 0.2, 0.5)])
def test_emitter_conversion(psf):
    """src/roi_batch.jl: camera = corner + roi - 1; microns = (camera - 1) * pixel."""
    P = len(psf)
    n = 3
    params = np.arange(1, P * n + 1, dtype=np.float32).reshape((P, n), order="F")
    crlb = params / 10
    out = F.FitArrays(params, crlb, np.array([-50, -60, -70], np.float32))
    cam = models.IdealCamera.square(256, 256, 0.1)
    batch = models.ROIBatch(np.zeros((7, 7, n), np.float32), [5, 10, 20], [7, 1, 3], [1, 2, 2], cam)
    pv = g.pvalues(out.llr, 7, P)
    loc = results.LocalizationResult.create(out, pv, batch, psf)
    smld = results.to_smld(loc, batch)
    assert smld.n_frames == 2 and len(smld.emitters) == n
    e = smld.emitters[1]
    px = np.float32(0.1)
    assert e.x == pytest.approx(float((np.float32(10) + params[0, 1] - 1 - 1) * px))
    assert e.y == pytest.approx(float((np.float32(1) + params[1, 1] - 1 - 1) * px))
    assert e.σ_x == pytest.approx(float(crlb[0, 1] * px))
    assert e.frame == 2 and e.id == 2 and e.pvalue == pytest.approx(float(pv[1]))
    if isinstance(psf, g.AstigmaticXYZNB):
        assert e.z == params[2, 1] and e.photons == params[3, 1] and e.σ_z == pytest.approx(float(crlb[2, 1]))
    else:
        assert e.photons == params[2, 1] and e.bg == params[3, 1]


def test_fitter_defaults_mirror_reference():
    # src/interface.jl:115-145 (the device probe needs a GPU, so only the pieces)
    c = g.default_constraints(g.GaussianXYNB(0.13), 11)
    np.testing.assert_array_equal(c.lower, np.float32([-2, -2, 1, 0.01]))
    np.testing.assert_array_equal(c.upper, np.float32([13, 13, np.inf, np.inf]))
    np.testing.assert_array_equal(c.max_step, np.float32([1, 1, 100, 2]))


@pytest.mark.parametrize("pixel_dtype", [np.float64, np.float32])
@pytest.mark.parametrize("model", ["xynb", "xynbs", "xynbsxsy", "astig"])
def test_to_smld_matches_per_emitter_conversion(model, pixel_dtype):
    """to_smld converts whole columns at once; every emitter must equal the
    one to_emitter builds alone (src/roi_batch.jl:121-353)."""
    from types import SimpleNamespace
    from gaussmle_amd.models import (AstigmaticXYZNB, GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, IdealCamera,
                                     ROIBatch)
    from gaussmle_amd.results import LocalizationResult, to_emitter, to_smld
    psf = {"xynb": GaussianXYNB(0.13), "xynbs": GaussianXYNBS(), "xynbsxsy": GaussianXYNBSXSY(),
           "astig": AstigmaticXYZNB(0.13, 0.13, 0.05, -0.05, 0.01, -0.01, 0.2, 0.5)}[model]
    n, P = 500, len(psf)
    rng = np.random.default_rng(3)
    out = SimpleNamespace(params=(rng.random((P, n)) * 10).astype(np.float32),
                          crlb=rng.random((P, n)).astype(np.float32), llr=rng.random(n).astype(np.float32))
    cam = IdealCamera.square(2048, 2048, 0.1, dtype=pixel_dtype)
    batch = ROIBatch(np.zeros((11, 11, n), np.float32), rng.integers(1, 2000, n), rng.integers(1, 2000, n),
                     rng.integers(1, 50, n), cam)
    loc = LocalizationResult.create(out, rng.random(n), batch, psf)
    smld = to_smld(loc, batch)
    assert smld.emitters == [to_emitter(loc, i, cam) for i in range(n)]
    assert len(smld) == n
