"""GPU parity: the HIP kernels against the CPU oracle on identical seeded inputs.

Every model x camera x precision and the lanes-per-ROI variants, through both
C-ABI entry points (host arrays via gmle_fit, device-resident arrays via
gmle_fit_device), plus edge cases (smallest / largest box, one ROI, ragged
batches, zero iterations, out-of-range sCMOS corners).  Tolerances and the
outlier rule are in tests/parity.py.
"""
import sys

import numpy as np
import pytest

from parity import assert_parity, compare, summary

pytestmark = pytest.mark.gpu

import gaussmle_amd as g
from gaussmle_amd import _lib, synth
from gaussmle_amd.models import GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY

ASTIG = synth.ASTIG_CONFIG_PSF


def _cases():
    return {
        "xynb_ideal_11": lambda: synth.config0(n=4096, seed=100),
        "xynbs_scmos_13": lambda: synth.config1(n=4096, seed=101),
        "astig_ideal_15": lambda: synth.config2(n=2048, seed=102),
        "xynbsxsy_scmos_9": lambda: synth.generic(GaussianXYNBSXSY(), 2048, 9, seed=103, scmos=True,
                                                  sigma_range=(1.0, 1.8)),
        "xynbs_ideal_7": lambda: synth.generic(GaussianXYNBS(), 2048, 7, seed=104),
        "xynb_scmos_13": lambda: synth.generic(GaussianXYNB(1.3), 2048, 13, seed=105, scmos=True),
        "astig_scmos_11": lambda: synth.generic(ASTIG, 2048, 11, seed=106, scmos=True, n_range=(1000, 10000)),
        "xynbsxsy_ideal_16": lambda: synth.generic(GaussianXYNBSXSY(), 1024, 16, seed=107),
        "xynb_ideal_3": lambda: synth.generic(GaussianXYNB(1.0), 1024, 3, seed=108, n_range=(200, 800)),
        "xynbs_ideal_12": lambda: synth.generic(GaussianXYNBS(), 1024, 12, seed=109),
    }


_W = {}
_REF = {}


def workload(name):
    if name not in _W:
        _W[name] = _cases()[name]()
    return _W[name]


def reference(name, prec):
    """The oracle's fit of a case; prec "fp64p" is the Float64 oracle on the
    ulp-perturbed data (the conditioning check of tests/parity.py)."""
    from oracle import refcpu
    from parity import perturbed_workload
    key = (name, prec)
    if key not in _REF:
        if prec == "fp64p":
            _REF[key] = refcpu.fit_workload(perturbed_workload(workload(name)), refcpu.FP64)
        else:
            _REF[key] = refcpu.fit_workload(workload(name), prec)
    return _REF[key]


def lanes_for(box):
    return [l for l in (1, 2, 4, 8, 16) if _lib.load().gmle_lanes_supported(box, l)]


def gpu_fit_host(w, precision, lanes):
    f = g.GaussMLEFitter.__new__(g.GaussMLEFitter)
    f.psf_model, f.iterations, f.constraints = w.psf_pixels, w.iterations, w.constraints
    f.precision, f.lanes, f.batch_size, f.device = precision, lanes, 10_000, g.GPU()
    out = g.fit_arrays(f, w.data, w.psf_pixels, w.var_map, w.x_corners, w.y_corners)
    res = {"params": out.params.T.astype(np.float64), "crlb": out.crlb.T.astype(np.float64),
           "llr": out.llr.astype(np.float64)}
    if precision == "mixed":
        res["crlb"] = out.crlb64.T
        res["llr"] = out.llr64
    return res


def gpu_fit_device(w, precision, lanes):
    import torch
    from gaussmle_amd import device as gd
    d = gd.upload_workload(w)
    ny, nx = (0, 0) if w.var_map is None else w.var_map.shape
    out = gd.fit_device(w.psf_pixels, w.box, d["data"], w.constraints, w.iterations, precision, lanes,
                        d.get("var"), ny, nx, d.get("xc"), d.get("yc"))
    torch.cuda.synchronize()
    res = {"params": out["params"].double().cpu().numpy(), "crlb": out["crlb"].double().cpu().numpy(),
           "llr": out["llr"].double().cpu().numpy()}
    if precision == "mixed":
        res["crlb"] = out["crlb64"].cpu().numpy()
        res["llr"] = out["llr64"].cpu().numpy()
    return res


def _report(label, rep, model_id=None):
    """Prints the summary; with GMLE_PARITY_REPORT set, also appends the
    per-output error distributions to that JSON-lines file."""
    import json
    import os
    print(f"[parity] {label} {summary(rep)}", flush=True)
    path = os.environ.get("GMLE_PARITY_REPORT")
    if path and model_id is not None:
        from parity import distributions
        with open(path, "a") as f:
            f.write(json.dumps({"case": label, "n": rep["n"], "ill_posed": rep["n_ill"], "outliers": rep["n_bad"],
                                "errors": distributions(model_id, rep)}) + "\n")


@pytest.mark.parametrize("name", list(_cases()))
@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_parity_host_path(name, precision):
    from oracle import refcpu
    w = workload(name)
    r32, r64, r64p = reference(name, refcpu.FP32), reference(name, refcpu.FP64), reference(name, "fp64p")
    for lanes in lanes_for(w.box):
        print(f"[case] {name} {precision} lanes={lanes} n={w.n}", flush=True)
        rmix = reference(name, refcpu.MIXED) if precision == "mixed" else None
        rep = compare(w.psf_pixels.model_id, gpu_fit_host(w, precision, lanes), r32, r64, precision, r64p, rmix)
        _report(f"{name} {precision} L={lanes}", rep, w.psf_pixels.model_id)
        assert_parity(rep, f"{name} {precision} L={lanes}")


@pytest.mark.parametrize("name", ["xynb_ideal_11", "xynbs_scmos_13", "astig_ideal_15"])
def test_parity_device_path(name):
    from oracle import refcpu
    w = workload(name)
    r32, r64, r64p = reference(name, refcpu.FP32), reference(name, refcpu.FP64), reference(name, "fp64p")
    for lanes in lanes_for(w.box):
        print(f"[case] device {name} lanes={lanes}", flush=True)
        a = gpu_fit_device(w, "fp32", lanes)
        b = gpu_fit_host(w, "fp32", lanes)
        np.testing.assert_array_equal(a["params"], b["params"])   # same kernel, same bits
        np.testing.assert_array_equal(a["crlb"], b["crlb"])
        rep = compare(w.psf_pixels.model_id, a, r32, r64, "fp32", r64p)
        assert_parity(rep, f"device {name} L={lanes}")


@pytest.mark.parametrize("name", ["astig_ideal_15", "xynbs_scmos_13", "xynb_ideal_11"])
def test_final_pass_against_oracle_at_same_estimate(name):
    """The epilogue alone: the oracle's final pass (LLR, Fisher, inversion,
    CRLB) evaluated at the GPU's own estimates.  Mixed precision must match the
    Float64 final pass to near Float64 rounding; the Float32 mode matches the
    Float32 final pass and is reported against the Float64 one."""
    from oracle import refcpu
    w = workload(name)
    gm = gpu_fit_host(w, "mixed", 0)
    f64 = refcpu.final_pass(w, gm["params"], refcpu.FP64)
    e_crlb = np.abs(gm["crlb"] / f64["crlb"] - 1).max(axis=1)
    e_llr = np.abs(gm["llr"] - f64["llr"]) / np.maximum(1.0, np.abs(f64["llr"]))
    g32 = gpu_fit_host(w, "fp32", 0)
    f32 = refcpu.final_pass(w, g32["params"], refcpu.FP32)
    f64b = refcpu.final_pass(w, g32["params"], refcpu.FP64)
    e32 = np.abs(g32["crlb"] / f32["crlb"] - 1).max(axis=1)
    e32_64 = np.abs(g32["crlb"] / f64b["crlb"] - 1).max(axis=1)
    print(f"[final] {name} mixed vs fp64 final pass: crlb max={e_crlb.max():.2e} p99.9={np.quantile(e_crlb, .999):.2e} "
          f"llr max={e_llr.max():.2e}; fp32 vs fp32 final pass crlb p99.9={np.quantile(e32, .999):.2e}; "
          f"fp32 vs fp64 final pass crlb max={e32_64.max():.2e} p99.9={np.quantile(e32_64, .999):.2e}", flush=True)
    assert e_crlb.max() < 1e-9 and e_llr.max() < 1e-9
    assert np.quantile(e32, 0.999) < 1e-4


@pytest.mark.parametrize("n", [1, 63, 65, 129])
def test_ragged_batches(n):
    """Batch sizes that leave partial waves / groups."""
    from oracle import refcpu
    w = synth.config0(n=n, seed=200 + n)
    r32, r64 = refcpu.fit_workload(w, refcpu.FP32), refcpu.fit_workload(w, refcpu.FP64)
    for lanes in lanes_for(w.box):
        rep = compare(0, gpu_fit_host(w, "fp32", lanes), r32, r64, "fp32")
        assert_parity(rep, f"ragged n={n} L={lanes}")


def test_zero_iterations_reproduce_initialisation():
    from oracle import refcpu
    for name in ["xynb_ideal_11", "astig_ideal_15", "xynbsxsy_scmos_9"]:
        w = workload(name).subset(512)
        w.iterations = 0
        r32 = refcpu.fit_workload(w, refcpu.FP32)
        for lanes in lanes_for(w.box):
            gp = gpu_fit_host(w, "fp32", lanes)["params"]
            # Float32 sums in a different order; z amplifies the moment difference
            np.testing.assert_allclose(gp, r32["params"], rtol=1e-4, atol=1e-4, err_msg=f"{name} L={lanes}")


def test_statistical_efficiency_at_scale():
    """Reference validation setup (test/validation_utils.jl:68-160, box 15,
    N ~ 1000, bg ~ 5) at 200k ROIs: bias and std/CRLB within the reference's
    tolerances, with the GPU CRLB."""
    rng = np.random.default_rng(300)
    n, box = 200_000, 15
    th = np.stack([box / 2 + 0.5 * rng.standard_normal(n), box / 2 + 0.5 * rng.standard_normal(n),
                   1000 * (0.8 + 0.4 * rng.random(n)), 5 * (0.8 + 0.4 * rng.random(n))])
    psf = GaussianXYNB(1.3)
    data = np.asfortranarray(rng.poisson(synth.expected_counts(psf, th, box)).astype(np.float32))
    w = synth.Workload("stat", psf, box, data, th, g.default_constraints(psf, 11), 20)
    r = gpu_fit_host(w, "fp32", 0)
    err = r["params"] - th.T
    for k, tol in enumerate([0.15, 0.15, 100.0, 2.0]):
        ratio = err[:, k].std() / r["crlb"][:, k].mean()
        print(f"[stat] param {k} bias={err[:, k].mean():.4g} std/crlb={ratio:.4f}", flush=True)
        assert abs(err[:, k].mean()) < tol
        assert abs(ratio - 1) < 0.05


def test_out_of_range_corner_device_path_gives_nan():
    """gmle_fit_device cannot scan device corners on the host: an sCMOS ROI
    outside the map must come back NaN, never read out of bounds."""
    import torch
    from gaussmle_amd import device as gd
    w = workload("xynb_scmos_13").subset(128)
    w.x_corners = w.x_corners.copy()
    w.x_corners[5] = w.var_map.shape[1] - 5   # runs past the last column
    w.y_corners = w.y_corners.copy()
    w.y_corners[9] = -3
    d = gd.upload_workload(w)
    ny, nx = w.var_map.shape
    out = gd.fit_device(w.psf_pixels, w.box, d["data"], w.constraints, 20, "fp32", 0, d["var"], ny, nx,
                        d["xc"], d["yc"])
    torch.cuda.synchronize()
    p = out["params"].cpu().numpy()
    assert np.all(np.isnan(p[[5, 9]])) and np.all(np.isfinite(p[[0, 1, 2, 127]]))


def test_fit_api_end_to_end():
    """fit(fitter, ROIBatch) -> BasicSMLD on the GPU, for both cameras."""
    w = workload("xynb_ideal_11").subset(256)
    cam = g.IdealCamera.square(256, 256, 0.1)
    xc = np.arange(256, dtype=np.int32) % 200 + 1
    yc = np.arange(256, dtype=np.int32) % 150 + 1
    batch = g.ROIBatch(w.data, xc, yc, None, cam)
    fitter = g.GaussMLEFitter(psf_model=g.GaussianXYNB(0.13), device="gpu")
    smld = g.fit(fitter, batch)
    assert len(smld.emitters) == 256
    e = smld.emitters[3]
    assert np.isfinite(e.x) and e.σ_x > 0 and 0 <= e.pvalue <= 1
    em = g.fit(fitter, w.data[:, :, 7])
    assert abs(em.photons - w.truth[2, 7]) < 5 * 100
