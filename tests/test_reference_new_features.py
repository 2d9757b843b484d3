"""The reference's simulator and ROIBatch tests (test/test_all_new_features.jl),
restated: generate_roi_batch on the device, then fit(fitter, ROIBatch).

  - Basic simulator (:7-18): a batch of 10 default ROIs is an ROIBatch of
    length 10, ROI size 11, data (11, 11, 10), on the given camera.
  - sCMOS support (:20-40): readnoise 5 e- rms, offset 100, gain 0.5,
    qe 0.82; the batch keeps the camera.
  - Fitting with ROIBatch (:42-67): 20 default ROIs; mean photons within
    50 % of 1000 and every position uncertainty in (0.002, 0.02) um.
  - Conversion to BasicSMLD (:69-94): two ROIs at corners (10, 20) and
    (30, 40), no x / y variation; finite, non-negative positions in um.
  - Different PSF models (:96-121): XYNB, XYNBS, XYNBSXSY, 5 ROIs each;
    finite position and photon uncertainties.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import gaussmle_amd as g
from gaussmle_amd.models import (GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, IdealCamera, ROIBatch,
                                 SCMOSCamera)


def gen(*a, **k):
    from gaussmle_amd.simulator import generate_roi_batch
    return generate_roi_batch(*a, **k)


def test_basic_simulator_with_ideal_camera():
    camera = IdealCamera.square(512, 512, 0.1)
    batch = gen(camera, GaussianXYNB(0.13), n_rois=10, seed=42)
    assert isinstance(batch, ROIBatch)
    assert len(batch) == 10
    assert batch.roi_size == 11
    assert batch.camera is camera
    assert batch.data.shape == (11, 11, 10)


def test_scmos_camera_support():
    scmos = SCMOSCamera.square(256, 256, 0.1, readnoise=5.0, offset=100.0, gain=0.5, qe=0.82)
    assert scmos.readnoise == 5.0 and scmos.gain == 0.5
    batch = gen(scmos, GaussianXYNB(0.13), n_rois=5, seed=42)
    assert batch.camera is scmos
    assert np.isfinite(batch.data).all()


def test_fitting_with_roibatch():
    camera = IdealCamera.square(512, 512, 0.1)
    psf = GaussianXYNB(0.13)
    batch = gen(camera, psf, n_rois=20, seed=42)
    smld = g.fit(g.GaussMLEFitter(psf_model=psf, device="gpu", iterations=20), batch)
    assert isinstance(smld, g.BasicSMLD)
    assert len(smld.emitters) == 20
    photons = np.array([e.photons for e in smld.emitters])
    sx = np.array([e.σ_x for e in smld.emitters])
    assert abs(photons.mean() - 1000.0) <= 0.5 * max(photons.mean(), 1000.0)
    assert np.all((0.002 < sx) & (sx < 0.02))


def test_conversion_to_basic_smld():
    camera = IdealCamera.square(256, 256, 0.1)
    psf = GaussianXYNB(0.13)
    corners = np.array([[10, 30], [20, 40]], dtype=np.int32)   # Int32[10 20; 30 40]'
    batch = gen(camera, psf, n_rois=2, corners=corners, xy_variation=0.0, seed=42)
    smld = g.fit(g.GaussMLEFitter(psf_model=psf, device="gpu", iterations=20), batch)
    assert isinstance(smld, g.BasicSMLD)
    assert len(smld) == 2
    assert isinstance(smld.camera, IdealCamera)
    assert all(np.isfinite(e.x) and e.x >= 0 for e in smld.emitters)
    assert all(np.isfinite(e.y) and e.y >= 0 for e in smld.emitters)


@pytest.mark.parametrize("psf", [GaussianXYNB(0.13), GaussianXYNBS(), GaussianXYNBSXSY()],
                         ids=["xynb", "xynbs", "xynbsxsy"])
def test_different_psf_models(psf):
    camera = IdealCamera.square(256, 256, 0.1)
    batch = gen(camera, psf, n_rois=5, seed=42)
    smld = g.fit(g.GaussMLEFitter(psf_model=psf, device="gpu", iterations=20), batch)
    assert len(smld.emitters) == 5
    sx = np.array([e.σ_x for e in smld.emitters])
    sn = np.array([e.σ_photons for e in smld.emitters])
    assert not np.isinf(sx).any() and not np.isnan(sx).any()
    assert not np.isinf(sn).any() and not np.isnan(sn).any()
