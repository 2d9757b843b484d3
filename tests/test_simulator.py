"""The on-device simulator (gmle_simulate_device, gaussmle_amd.simulator).

GaussMLE.jl's generate_roi_batch (src/simulator.jl:35-345) draws with Julia's
RNG and Distributions.jl, which cannot run here, so bit parity with it is
unpinned by construction.  What is pinned:
  * the generator: Philox4x32-10 against its published known-answer vectors
    (Salmon et al., SC'11, Random123 test vectors), in the library and in the
    oracle;
  * the samplers (oracle/refsim.cpp): Poisson and normal draws against
    scipy's distributions (chi-square goodness of fit, moments);
  * the expected counts: the oracle against synth.expected_counts (scipy erf,
    Float64), all four models;
  * the device against the oracle on the same seed, pixel for pixel;
  * end to end: fits of simulated batches meet the reference's statistical
    tolerances (std / CRLB, bias).
"""
import numpy as np
import pytest
from scipy import stats

from gaussmle_amd import simulator as gsim
from gaussmle_amd import synth
from gaussmle_amd.models import (AstigmaticXYZNB, GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, IdealCamera,
                                 SCMOSCamera)
from oracle import refsim

KAT = [  # (counter, key) -> output, Philox4x32-10 known-answer vectors
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(hiplib, ctr, key, want):
    import ctypes as C
    o = (C.c_uint32 * 4)()
    hiplib.gmle_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), o)
    assert list(o) == want
    assert refsim.philox(ctr, key) == want


@pytest.mark.parametrize("mean", [0.3, 2.5, 9.99, 10.0, 37.5, 400.0, 5000.0])
def test_oracle_poisson_sampler_follows_the_distribution(mean):
    x = refsim.draws("poisson", mean, seed=7, n=200_000)
    assert np.all(x == np.floor(x)) and x.min() >= 0
    n = len(x)
    assert abs(x.mean() - mean) < 5 * np.sqrt(mean / n)
    assert abs(x.var() / mean - 1) < 5 * np.sqrt(2.0 / n) * 1.5
    lo, hi = stats.poisson.ppf(1e-4, mean), stats.poisson.ppf(1 - 1e-4, mean)
    bounds = np.append(np.unique(np.linspace(lo, hi + 1, 60).astype(int)), np.inf)
    obs, _ = np.histogram(x, bins=bounds)                 # counts in [b_i, b_i+1)
    expct = n * (stats.poisson.cdf(bounds[1:] - 1, mean) - stats.poisson.cdf(bounds[:-1] - 1, mean))
    ok = expct > 5
    chi2 = np.sum((obs[ok] - expct[ok]) ** 2 / expct[ok])
    assert stats.chi2.sf(chi2, ok.sum() - 1) > 1e-4, (mean, chi2, ok.sum())


def test_oracle_normal_sampler_follows_the_distribution():
    x = refsim.draws("normal", 0.0, seed=8, n=200_000)
    assert abs(x.mean()) < 5 / np.sqrt(len(x)) and abs(x.var() - 1) < 5 * np.sqrt(2 / len(x))
    assert stats.kstest(x, "norm").pvalue > 1e-4


@pytest.mark.parametrize("psf,params", [
    (GaussianXYNB(1.3), [6.2, 5.7, 1500.0, 7.0]),
    (GaussianXYNBS(), [5.4, 6.3, 900.0, 3.0, 1.45]),
    (GaussianXYNBSXSY(), [6.0, 6.6, 1200.0, 12.0, 1.1, 1.7]),
    (AstigmaticXYZNB(1.3, 1.3, 0.05, -0.05, 0.3, 0.3, 50.0, 100.0), [6.1, 5.9, -40.0, 2000.0, 5.0]),
])
def test_oracle_expected_counts_match_the_model(psf, params):
    p = np.array(params, np.float32)[:, None]
    _, exp = refsim.simulate(psf, 11, p.T, seed=1)
    want = synth.expected_counts(psf, p.astype(np.float64), 11)[:, :, 0]   # (row, col)
    got = exp.reshape(11, 11).T                                           # (col, row) -> (row, col)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-4)


def test_corner_placement_modes():
    rng = np.random.default_rng(3)
    c = gsim.make_corners(40, (512, 256), 11, "random", 20, rng)
    assert c.shape == (2, 40) and c.min() >= 1 and c[0].max() <= 502 and c[1].max() <= 246
    d = np.abs(c[:, :, None] - c[:, None, :])
    close = (d[0] < 20) & (d[1] < 20)
    np.fill_diagonal(close, False)
    assert not close.any()
    g = gsim.make_corners(9, (200, 200), 11, "grid", 20, rng)
    assert len({tuple(v) for v in g.T}) == 9 and g.min() >= 1
    k = gsim.make_corners(30, (300, 300), 11, "clustered", 10, rng)
    assert k.min() >= 1 and k.max() <= 290
    with pytest.raises(ValueError, match="corner_mode"):
        gsim.make_corners(3, (64, 64), 11, "spiral", 5, rng)


def test_default_parameters_follow_the_reference():
    rng = np.random.default_rng(4)
    p = gsim.default_params(GaussianXYNBS(), 1000, 1.0, rng)
    assert p.dtype == np.float32 and p.shape == (5, 1000)
    assert np.all(np.abs(p[:2] - 6.0) <= 1.0) and np.ptp(p[0]) > 1.8
    np.testing.assert_array_equal(p[2:], np.array([[1000.0], [10.0], [1.3]], np.float32).repeat(1000, 1))
    a = gsim.default_params(AstigmaticXYZNB(1, 1, 0, 0, 0, 0, 1, 1), 5, 0.0, rng)
    np.testing.assert_array_equal(a[:, 0], np.array([6, 6, 0, 1000, 10], np.float32))


# --- the device ---------------------------------------------------------------

def _same_as_oracle(got, want, label):
    eq = got == want
    frac = eq.mean()
    print(f"[sim] {label}: identical {frac:.6f} of {got.size}", flush=True)
    # draws differ only where erff rounds the expected count differently
    assert frac > 0.999, label
    diff = np.abs(got[~eq] - want[~eq])
    assert diff.size == 0 or np.median(diff) <= 2.0


@pytest.mark.gpu
@pytest.mark.parametrize("model", [0, 1, 2, 3])
def test_device_matches_oracle_ideal(model):
    import torch
    psf = [GaussianXYNB(1.3), GaussianXYNBS(), GaussianXYNBSXSY(),
           AstigmaticXYZNB(1.3, 1.3, 0.05, -0.05, 0.3, 0.3, 50.0, 100.0)][model]
    w = synth.generic(psf, 3000, 11, seed=50 + model)
    params = np.ascontiguousarray(w.truth.T, dtype=np.float32)
    want, _ = refsim.simulate(psf, 11, params, seed=1234 + model)
    got = gsim.simulate_device(psf, 11, torch.from_numpy(params).cuda(), 1234 + model)
    torch.cuda.synchronize()
    _same_as_oracle(got.cpu().numpy(), want, f"ideal model {model}")


@pytest.mark.gpu
def test_device_matches_oracle_scmos_maps():
    import torch
    rng = np.random.default_rng(60)
    ny, nx = 96, 128
    cam = SCMOSCamera.square(nx, ny, 0.1, readnoise=rng.uniform(1, 3, (ny, nx)), offset=rng.uniform(95, 105, (ny, nx)),
                             gain=rng.uniform(0.4, 0.6, (ny, nx)), qe=0.9)
    psf = GaussianXYNB(1.3)
    n = 2000
    params = np.stack([rng.uniform(5, 7, n), rng.uniform(5, 7, n), rng.uniform(500, 3000, n),
                       rng.uniform(2, 20, n)], axis=1).astype(np.float32)
    xc = rng.integers(1, nx - 11 + 2, n).astype(np.int32)
    yc = rng.integers(1, ny - 11 + 2, n).astype(np.int32)
    maps = {k: np.asfortranarray(getattr(cam, k)).reshape(-1, order="F") for k in ("readnoise", "offset", "gain")}
    want, _ = refsim.simulate(psf, 11, params, 99, camera=1, xc=xc, yc=yc, maps=maps, scalars={"qe": 0.9}, ny=ny,
                              nx=nx)
    got = gsim.simulate_device(psf, 11, torch.from_numpy(params).cuda(), 99, cam, torch.from_numpy(xc).cuda(),
                               torch.from_numpy(yc).cuda())
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    close = np.isclose(got, want, rtol=0, atol=1e-3)
    print(f"[sim] scmos: within 1e-3 ADU {close.mean():.6f}", flush=True)
    assert close.mean() > 0.999


@pytest.mark.gpu
def test_pixel_statistics_match_the_model():
    """20,000 ROIs of one emitter: per-pixel mean and variance of the draws
    against the expected counts (Poisson: variance = mean)."""
    import torch
    psf = GaussianXYNB(1.3)
    n = 20_000
    p = np.tile(np.array([[6.3, 5.8, 2000.0, 8.0]], np.float32), (n, 1))
    got = gsim.simulate_device(psf, 11, torch.from_numpy(p).cuda(), 5).cpu().numpy().reshape(n, 121)
    _, exp = refsim.simulate(psf, 11, p[:1], seed=5)
    m, v = got.mean(0), got.var(0)
    z_mean = (m - exp) / np.sqrt(exp / n)
    assert np.abs(z_mean).max() < 5.0
    assert np.abs(v / exp - 1).max() < 6 * np.sqrt(2.0 / n) * 1.5


@pytest.mark.gpu
@pytest.mark.parametrize("camera", ["ideal", "scmos"])
def test_generate_roi_batch_then_fit_meets_reference_tolerances(camera):
    """src/simulator.jl's use in test/test_strict_validation.jl:112-138: fits of
    simulated ROIs are unbiased and their std matches the CRLB within 10 %."""
    from gaussmle_amd import GaussMLEFitter, fit_arrays, to_electrons, variance_map
    from gaussmle_amd.models import to_pixel_units
    if camera == "ideal":
        cam = IdealCamera.square(512, 512, 0.1)
    else:
        cam = SCMOSCamera.square(512, 512, 0.1, readnoise=1.5, offset=100.0, gain=0.5, qe=1.0)
    psf = GaussianXYNB(0.13)
    n = 4000
    batch = gsim.generate_roi_batch(cam, psf, n_rois=n, roi_size=11, seed=11, corner_mode="grid", min_spacing=1)
    assert batch.data.shape == (11, 11, n) and batch.data.flags.f_contiguous
    fitter = GaussMLEFitter(psf)
    psf_px = to_pixel_units(psf, 0.1)
    if camera == "ideal":
        out = fit_arrays(fitter, batch.data, psf_px)
    else:
        e = to_electrons(batch.data, cam, batch.x_corners, batch.y_corners)
        out = fit_arrays(fitter, e, psf_px, variance_map(cam), batch.x_corners, batch.y_corners)
    rng = np.random.default_rng(11)
    truth = gsim.default_params(psf_px, n, 1.0, rng)
    err = out.params[:2] - truth[:2]
    std = err.std(axis=1)
    crlb = np.sqrt(np.mean(out.crlb[:2] ** 2, axis=1))
    print(f"[sim] {camera}: bias {err.mean(axis=1)}, std/CRLB {std / crlb}", flush=True)
    assert np.all(np.abs(err.mean(axis=1)) < 3 * std / np.sqrt(n) + 1e-3)
    assert np.all(np.abs(std / crlb - 1) < 0.1)
