"""The reference's model validation (test/comprehensive_cpu_tests.jl with
test/validation_utils.jl), restated.

Each scenario draws blobs as generate_test_data does (validation_utils.jl:
68-160): x, y = box/2 + 0.5 randn, N and bg times U(0.8, 1.2), sigma times
U(0.8, 1.2) for the sigma models, pixels Poisson around
bg + N Phi_x(j) Phi_y(i) (:161-165).  The fit is fit(fitter, data) with 20
iterations (:313-321), and validate_fitting_results (:171-296) checks, per
parameter, |mean error| < bias_tol and |1 - std(error) / mean(CRLB)| <
std_tol, with the tolerances of :341-350; positions come back through the
dummy corners as x_roi = x_um / pixel - corner + 2 (:27-58).

The scenarios and their extra bounds follow comprehensive_cpu_tests.jl:
5-167: the three 2-D models at N = 1000, bg = 5; N = 200 (bg 5) with x std
above 0.08 px; N = 5000 (bg 20) with x std below 0.03 px; sigma 1.0 px
with x std below 0.06 and 2.0 px with x std above 0.07.

Differences, on purpose: numpy's seeded generator instead of Julia's seed
42, and 4,000 blobs per scenario instead of 200-500, so that the +-10 %
std/CRLB band is about nine standard errors of the ratio wide rather than
two (the reference's pass depends on its particular random stream).

The CPU tests run the Float32 oracle on the same data (the reference runs
these on its CPU path); the GPU tests go through fit(fitter, data).
"""
import numpy as np
import pytest
from scipy.special import erf

import gaussmle_amd as g
from gaussmle_amd import synth
from gaussmle_amd.models import (GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, default_constraints,
                                 to_pixel_units)

F32 = np.float32
BOX = 11
N_BLOBS = 4000
PIXEL = 0.1   # um, fit(fitter, Array)'s camera (src/interface.jl:191)

TOL = {   # validation_utils.jl:341-350
    "x": (0.15, 0.1), "y": (0.15, 0.1), "photons": (100.0, 0.1), "background": (2.0, 0.1),
    "sigma": (0.05, 0.1), "sigma_x": (0.05, 0.1), "sigma_y": (0.05, 0.1),
}
PARAMS = {
    "xynb": ["x", "y", "photons", "background"],
    "xynbs": ["x", "y", "photons", "background", "sigma"],
    "xynbsxsy": ["x", "y", "photons", "background", "sigma_x", "sigma_y"],
}


def integral_gaussian_1d(k, mu, sigma):
    """src/gausslib/baselibrary.jl:8-14."""
    s = np.sqrt(2.0) * sigma
    return 0.5 * (erf((k - mu + 0.5) / s) - erf((k - mu - 0.5) / s))


def generate_test_data(model, n, n_photons=1000.0, background=5.0, sigma=1.3, position_std=0.5, seed=42):
    """validation_utils.jl:68-160 (numpy's generator)."""
    rng = np.random.default_rng(seed)
    x = (BOX / 2 + position_std * rng.standard_normal(n)).astype(F32)
    y = (BOX / 2 + position_std * rng.standard_normal(n)).astype(F32)
    ph = (n_photons * (0.8 + 0.4 * rng.random(n))).astype(F32)
    bg = (background * (0.8 + 0.4 * rng.random(n))).astype(F32)
    truth = {"x": x, "y": y, "photons": ph, "background": bg}
    if model == "xynb":
        sx = sy = np.full(n, sigma, F32)
    elif model == "xynbs":
        sx = sy = truth["sigma"] = (sigma * (0.8 + 0.4 * rng.random(n))).astype(F32)
    else:
        sx = truth["sigma_x"] = (sigma * (0.8 + 0.4 * rng.random(n))).astype(F32)
        sy = truth["sigma_y"] = (sigma * (0.8 + 0.4 * rng.random(n))).astype(F32)
    k = np.arange(1, BOX + 1, dtype=np.float64)
    px = integral_gaussian_1d(k[:, None], x[None, :].astype(np.float64), sx[None, :].astype(np.float64))  # (j, n)
    py = integral_gaussian_1d(k[:, None], y[None, :].astype(np.float64), sy[None, :].astype(np.float64))  # (i, n)
    mu = bg[None, None, :] + ph[None, None, :] * py[:, None, :] * px[None, :, :]   # [i, j, n]
    data = rng.poisson(mu).astype(F32)
    return np.asfortranarray(data), truth


def psf_of(model, sigma):
    if model == "xynb":
        return GaussianXYNB(sigma * PIXEL)   # microns, as the reference passes them
    return GaussianXYNBS() if model == "xynbs" else GaussianXYNBSXSY()


def fitted_gpu(model, psf, data):
    """fit(fitter, data) -> ROI-local values and CRLBs in pixels (validation_utils.jl:27-58, 193-260)."""
    smld = g.fit(g.GaussMLEFitter(psf_model=psf, device="gpu", iterations=20), data)
    es = smld.emitters
    n = len(es)
    corner_x = 1 + np.arange(n) * BOX
    vals = {"x": np.array([e.x / PIXEL for e in es]) - corner_x + 2,
            "y": np.array([e.y / PIXEL for e in es]) - 1 + 2,
            "photons": np.array([e.photons for e in es]), "background": np.array([e.bg for e in es])}
    errs = {"x": np.array([e.σ_x / PIXEL for e in es]), "y": np.array([e.σ_y / PIXEL for e in es]),
            "photons": np.array([e.σ_photons for e in es]), "background": np.array([e.σ_bg for e in es])}
    if model == "xynbs":
        vals["sigma"] = np.array([e.σ / PIXEL for e in es])
        errs["sigma"] = np.array([e.σ_σ / PIXEL for e in es])
    elif model == "xynbsxsy":
        vals["sigma_x"] = np.array([e.σx / PIXEL for e in es])
        vals["sigma_y"] = np.array([e.σy / PIXEL for e in es])
        errs["sigma_x"] = np.array([e.σ_σx / PIXEL for e in es])
        errs["sigma_y"] = np.array([e.σ_σy / PIXEL for e in es])
    return vals, errs


def fitted_oracle(model, psf, data):
    """The Float32 oracle on the same ROIs, in ROI pixels (the fit's own units)."""
    from oracle import refcpu
    pp = to_pixel_units(psf, PIXEL)
    w = synth.Workload("validation", pp, BOX, data, np.zeros((len(pp), data.shape[2])),
                       default_constraints(pp, 11), 20)
    r = refcpu.fit_workload(w, refcpu.FP32)
    names = PARAMS[model]
    return ({k: r["params"][:, i] for i, k in enumerate(names)}, {k: r["crlb"][:, i] for i, k in enumerate(names)})


def validate(model, vals, errs, truth):
    """validation_utils.jl:263-296 for every parameter of the model."""
    out = {}
    for p in PARAMS[model]:
        bias_tol, std_tol = TOL[p]
        e = vals[p].astype(np.float64) - truth[p].astype(np.float64)
        bias, std = e.mean(), e.std(ddof=1)
        ratio = std / errs[p].astype(np.float64).mean()
        out[p] = dict(bias=bias, empirical_std=std, std_ratio=ratio,
                      bias_pass=abs(bias) < bias_tol, std_pass=abs(1.0 - ratio) < std_tol)
    return out


# (name, model, generate_test_data keywords, extra check on the x results)
SCENARIOS = [
    ("xynb", "xynb", {}, None),
    ("xynbs", "xynbs", {}, None),
    ("xynbsxsy", "xynbsxsy", {}, None),
    ("low_photons", "xynb", dict(n_photons=200.0, background=5.0), lambda r: r["x"]["empirical_std"] > 0.08),
    ("high_photons", "xynb", dict(n_photons=5000.0, background=20.0), lambda r: r["x"]["empirical_std"] < 0.03),
    ("narrow_psf", "xynb", dict(sigma=1.0), lambda r: r["x"]["empirical_std"] < 0.06),
    ("wide_psf", "xynb", dict(sigma=2.0), lambda r: r["x"]["empirical_std"] > 0.07),
]


def run(scenario, fitted):
    name, model, kw, extra = scenario
    data, truth = generate_test_data(model, N_BLOBS, **kw)
    vals, errs = fitted(model, psf_of(model, kw.get("sigma", 1.3)), data)
    res = validate(model, vals, errs, truth)
    for p, r in res.items():
        print(f"[validation] {name} {p}: bias {r['bias']:+.4f} std {r['empirical_std']:.4f} "
              f"std/CRLB {r['std_ratio']:.3f}", flush=True)
        assert r["bias_pass"], f"{name} {p} bias {r['bias']}"
        assert r["std_pass"], f"{name} {p} std/CRLB {r['std_ratio']}"
    if extra is not None:
        assert extra(res), f"{name}: x std {res['x']['empirical_std']}"


@pytest.mark.parametrize("scenario", SCENARIOS, ids=[s[0] for s in SCENARIOS])
def test_oracle_model_validation(scenario):
    run(scenario, fitted_oracle)


@pytest.mark.gpu
@pytest.mark.parametrize("scenario", SCENARIOS, ids=[s[0] for s in SCENARIOS])
def test_gpu_model_validation(scenario):
    run(scenario, fitted_gpu)
