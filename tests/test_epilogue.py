"""The result epilogue (gmle_epilogue_device, gmle_chi2_pvalue).

GaussMLE.jl's host epilogue: p = 1 - Float32(cdf(Chisq(box^2 - P), -2 LLR))
(src/interface.jl:265-275), camera coordinates x_corner + x_roi - 1
(src/roi_batch.jl:29-34) and microns (x_camera - 1) * pixel_size with
uncertainties and widths scaled by the pixel size (src/roi_batch.jl:125-302).
Distributions.jl's Chisq CDF cannot run here; the p-value is pinned against
scipy.stats.chi2 in Float64, then rounded the way the reference rounds.
"""
import numpy as np
import pytest
from scipy import stats

from gaussmle_amd import _lib

F32 = np.float32


def reference_pvalue(chi2, df):
    return (F32(1.0) - np.asarray(stats.chi2.cdf(np.float64(chi2), df), dtype=F32)).astype(F32)


@pytest.mark.parametrize("df", [5, 117, 164, 165, 219, 220, 252])
def test_chi2_pvalue_matches_scipy(hiplib, df):
    chi2 = np.concatenate([np.linspace(0, 4 * df, 4001), [0.5 * df, df - 1e-3, df + 1e-3, 1e-6, 3 * df + 0.25]])
    chi2 = chi2.astype(F32)
    got = np.array([hiplib.gmle_chi2_pvalue(float(c), df) for c in chi2], dtype=F32)
    want = reference_pvalue(chi2, df)
    ulp = np.spacing(np.maximum(np.abs(want), F32(1e-30))).astype(np.float64)
    err = np.abs(got.astype(np.float64) - want) / ulp
    assert err.max() <= 1.0, (df, float(chi2[np.argmax(err)]), got[np.argmax(err)], want[np.argmax(err)])
    assert (err == 0).mean() > 0.99


def test_chi2_pvalue_special_values(hiplib):
    assert hiplib.gmle_chi2_pvalue(0.0, 10) == 1.0
    assert hiplib.gmle_chi2_pvalue(float("inf"), 10) == 0.0
    assert np.isnan(hiplib.gmle_chi2_pvalue(float("nan"), 10))
    assert np.isnan(hiplib.gmle_chi2_pvalue(1.0, 0))
    # chi^2 below zero (an LLR rounded just above zero): CDF 0, p = 1, as
    # Distributions.jl's gamma CDF clamps x at 0 and scipy gives
    for c in (-1e-3, -1.0, -1e30, float("-inf")):
        assert hiplib.gmle_chi2_pvalue(c, 10) == 1.0
        assert reference_pvalue(F32(c), 10) == 1.0


def mirror(model, box, params, crlb, llr, xc, yc, psx, psy):
    """numpy restatement of the reference's per-emitter arithmetic."""
    n = len(llr)
    P = [4, 5, 6, 5][model]
    out = np.full((n, 10), np.nan, dtype=F32)
    xcam = (xc.astype(F32) + params[:, 0]) - F32(1)
    ycam = (yc.astype(F32) + params[:, 1]) - F32(1)
    out[:, 0] = ((xcam - F32(1)).astype(np.float64) * psx).astype(F32)
    out[:, 1] = ((ycam - F32(1)).astype(np.float64) * psy).astype(F32)
    out[:, 2] = (crlb[:, 0].astype(np.float64) * psx).astype(F32)
    out[:, 3] = (crlb[:, 1].astype(np.float64) * psy).astype(F32)
    chi2 = (F32(-2) * llr).astype(F32)
    out[:, 4] = reference_pvalue(chi2, box * box - P)
    if model == 1:
        out[:, 5] = out[:, 6] = (params[:, 4].astype(np.float64) * psx).astype(F32)
        out[:, 7] = out[:, 8] = (crlb[:, 4].astype(np.float64) * psx).astype(F32)
    elif model == 2:
        out[:, 5] = (params[:, 4].astype(np.float64) * psx).astype(F32)
        out[:, 6] = (params[:, 5].astype(np.float64) * psy).astype(F32)
        out[:, 7] = (crlb[:, 4].astype(np.float64) * psx).astype(F32)
        out[:, 8] = (crlb[:, 5].astype(np.float64) * psy).astype(F32)
    out[:, 9] = chi2
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("model,box", [(0, 11), (1, 13), (2, 9), (3, 15)])
def test_epilogue_device_matches_reference_arithmetic(model, box):
    import torch
    from gaussmle_amd.device import epilogue_device
    rng = np.random.default_rng(60 + model)
    n = 5000
    P = [4, 5, 6, 5][model]
    params = np.abs(rng.normal(5, 2, (n, P))).astype(F32)
    crlb = np.abs(rng.normal(0.05, 0.02, (n, P))).astype(F32)
    llr = (-rng.chisquare(box * box - P, n) / 2).astype(F32)
    # LLRs rounded just above zero (chi^2 < 0) and exactly zero: p = 1
    llr[:4] = np.array([5e-4, 0.5, 0.0, 1e-7], F32)
    xc = rng.integers(1, 2000, n).astype(np.int32)
    yc = rng.integers(1, 2000, n).astype(np.int32)
    d = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    got = epilogue_device(d(params), d(crlb), d(llr), 0.1, 0.107, model, box, d_xc=d(xc), d_yc=d(yc))
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    want = mirror(model, box, params, crlb, llr, xc, yc, 0.1, 0.107)
    for q in (0, 1, 2, 3, 5, 6, 7, 8, 9):
        np.testing.assert_array_equal(got[:, q], want[:, q], err_msg=f"field {q}")
    assert np.all(got[:4, 4] == 1.0) and np.all(want[:4, 4] == 1.0)
    ulp = np.spacing(np.maximum(want[:, 4], F32(1e-30))).astype(np.float64)
    err = np.abs(got[:, 4].astype(np.float64) - want[:, 4]) / ulp
    print(f"[epilogue] model={model} box={box} p-value max {err.max():.1f} ulp, exact {np.mean(err == 0):.4f}",
          flush=True)
    assert err.max() <= 1.0


@pytest.mark.gpu
def test_epilogue_device_mixed_batch_and_array_corners():
    """Per-ROI models / boxes with stride 6, and the fit(Array) corners."""
    import torch
    from gaussmle_amd.device import epilogue_device
    rng = np.random.default_rng(70)
    n = 3000
    models = rng.integers(0, 4, n).astype(np.int8)
    boxes = rng.choice([7, 9, 11, 13, 15], n).astype(np.int8)
    params = np.abs(rng.normal(5, 2, (n, 6))).astype(F32)
    crlb = np.abs(rng.normal(0.05, 0.02, (n, 6))).astype(F32)
    llr = (-rng.chisquare(100, n) / 2).astype(F32)
    d = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    got = epilogue_device(d(params), d(crlb), d(llr), 0.1, None, d_models=d(models), d_boxes=d(boxes))
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    for m in range(4):
        for b in (7, 9, 11, 13, 15):
            idx = np.flatnonzero((models == m) & (boxes == b))
            xc = (1 + idx * b).astype(np.int32)
            yc = np.ones(len(idx), np.int32)
            P = [4, 5, 6, 5][m]
            want = mirror(m, b, params[idx, :P], crlb[idx, :P], llr[idx], xc, yc, 0.1, 0.1)
            np.testing.assert_array_equal(got[idx][:, [0, 1, 2, 3, 5, 6, 7, 8, 9]],
                                          want[:, [0, 1, 2, 3, 5, 6, 7, 8, 9]])
            assert np.all(np.abs(got[idx, 4] - want[:, 4]) <= np.spacing(np.maximum(want[:, 4], F32(1e-30))))


def _emitter_fields(e):
    """x, y, their uncertainties, the p-value and fitted widths (with theirs)."""
    w = {"Emitter2DFitSigma": ("σ", "σ", "σ_σ", "σ_σ"), "Emitter2DFitSigmaXY": ("σx", "σy", "σ_σx", "σ_σy")}
    names = w.get(type(e).__name__)
    widths = [getattr(e, k) for k in names] if names else [np.nan] * 4
    return [e.x, e.y, e.σ_x, e.σ_y, e.pvalue, *widths]


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["array_xynb", "batch_xynbs_f64_camera", "batch_sxsy_f32_camera"])
def test_epilogue_device_matches_host_emitters(case):
    """The device epilogue gives the emitters fit() builds on the host."""
    import torch
    from gaussmle_amd import synth
    from gaussmle_amd.device import epilogue_device
    from gaussmle_amd.fitter import GaussMLEFitter, fit, fit_arrays
    from gaussmle_amd.models import (GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, IdealCamera, ROIBatch,
                                     to_pixel_units)
    rng = np.random.default_rng(80)
    if case == "array_xynb":
        w = synth.config0(n=1500, seed=81)
        fitter = GaussMLEFitter(GaussianXYNB(0.13))
        smld = fit(fitter, w.data)
        xc = yc = None
        ps = (0.1, 0.1)
        psx = psy = float(F32(0.1))
        psf_um, model = fitter.psf_model, 0
    else:
        model = 1 if case == "batch_xynbs_f64_camera" else 2
        w = synth.generic(GaussianXYNBS() if model == 1 else GaussianXYNBSXSY(), 1500, 9, seed=82)
        fitter = GaussMLEFitter(w.psf_pixels, constraints=w.constraints)
        dtype = np.float64 if model == 1 else np.float32
        cam = IdealCamera.square(512, 512, 0.107, dtype=dtype)
        xc = rng.integers(1, 500, w.n).astype(np.int32)
        yc = rng.integers(1, 500, w.n).astype(np.int32)
        smld = fit(fitter, ROIBatch(w.data, xc, yc, np.ones(w.n, np.int32), cam))
        psx = float(cam.pixel_edges_x[1] - cam.pixel_edges_x[0])
        psy = float(cam.pixel_edges_y[1] - cam.pixel_edges_y[0])
        psf_um = w.psf_pixels
        ps = (psx, psy)
    out = fit_arrays(fitter, w.data, to_pixel_units(psf_um, ps[0]))
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    got = epilogue_device(d(out.params.T), d(out.crlb.T), d(out.llr), psx, psy, model, w.box,
                          d_xc=None if xc is None else d(xc), d_yc=None if yc is None else d(yc))
    torch.cuda.synchronize()
    got = got.cpu().numpy()[:, :9]
    want = np.array([_emitter_fields(e) for e in smld.emitters], dtype=F32)
    exact = [0, 1, 2, 3, 5, 6, 7, 8]
    np.testing.assert_array_equal(got[:, exact], want[:, exact])
    assert np.all(np.abs(got[:, 4] - want[:, 4]) <= np.spacing(np.maximum(want[:, 4], F32(1e-30))))


@pytest.mark.parametrize("model,dtype", [(0, np.float32), (1, np.float64), (2, np.float64), (3, np.float32)])
def test_host_emitters_follow_pixel_size_type(model, dtype):
    """to_emitter rounds once to Float32 whatever the camera's edge type; the
    device epilogue's arithmetic (mirror) is the same."""
    from gaussmle_amd.fitter import FitArrays
    from gaussmle_amd.models import (AstigmaticXYZNB, GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, IdealCamera,
                                     ROIBatch)
    from gaussmle_amd.results import LocalizationResult, to_smld
    rng = np.random.default_rng(90 + model)
    n, box = 400, 11
    P = [4, 5, 6, 5][model]
    psf = [GaussianXYNB(1.3), GaussianXYNBS(), GaussianXYNBSXSY(),
           AstigmaticXYZNB(1.3, 1.3, 0.05, -0.05, 0.03, -0.03, 250.0, 400.0)][model]
    params = np.abs(rng.normal(5, 2, (P, n))).astype(F32)
    crlb = np.abs(rng.normal(0.05, 0.02, (P, n))).astype(F32)
    llr = (-rng.chisquare(box * box - P, n) / 2).astype(F32)
    xc = rng.integers(1, 400, n).astype(np.int32)
    yc = rng.integers(1, 400, n).astype(np.int32)
    cam = IdealCamera.square(512, 512, 0.107, dtype=dtype)
    batch = ROIBatch(np.zeros((box, box, n), F32), xc, yc, np.ones(n, np.int32), cam)
    pv = reference_pvalue(F32(-2) * llr, box * box - P)
    loc = LocalizationResult.create(FitArrays(params, crlb, llr), pv, batch, psf)
    got = np.array([_emitter_fields(e) for e in to_smld(loc, batch).emitters], dtype=F32)
    psx = float(cam.pixel_edges_x[1] - cam.pixel_edges_x[0])
    psy = float(cam.pixel_edges_y[1] - cam.pixel_edges_y[0])
    want = mirror(model, box, params.T, crlb.T, llr, xc, yc, psx, psy)[:, :9]
    np.testing.assert_array_equal(got, want)
    if dtype == np.float64:
        # a Float32 pixel size of the same nominal value rounds differently somewhere
        alt = mirror(model, box, params.T, crlb.T, llr, xc, yc, float(F32(psx)), float(F32(psy)))[:, :2]
        assert np.any(alt != want[:, :2])
