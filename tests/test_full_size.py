"""The benchmark workloads at their full sizes, checked through properties that
do not depend on size (the oracle cannot fit millions of ROIs inside a test):

  - headline: 1,048,576 ROIs of 11x11 GaussianXYNB, Poisson, 20 iterations
    (synth.config0, the benchmark's default line), one gmle_fit_device launch
    with the library's lanes-per-ROI choice for that size;
  - precision split: 1,000,000 ROIs of 13x13 GaussianXYNBS on an sCMOS sensor
    (synth.config1), Float32 iterations with the Float64 final pass;
  - astigmatic 3-D: 4,000,000 ROIs of 15x15 AstigmaticXYZNB, 30 iterations
    (synth.config2), Float32 iterations with the Float64 final pass;
  - uneven batch: 4,000,000 ROIs of sizes {7, ..., 15} and three models on an
    sCMOS sensor (batch.config3), one gmle_fit_batch_device call.

Properties:
  - sample parity: a strided sample of the full launch's ROIs meets the oracle
    at the tolerances and outlier rule of tests/parity.py;
  - batch invariance: the sampled ROIs fitted on their own, with the same lanes
    per ROI, are bit-identical to their results inside the full launch (each
    ROI is fitted alone, so nothing may depend on the batch around it or on
    its position in it);
  - statistics (headline, precision split): over the whole launch, the
    normalised errors (fit - truth) / CRLB of x and y have the mean and spread
    that the Float64 oracle gives on the sample, within four standard errors
    of the sample.
"""
import numpy as np
import pytest

from parity import OUTLIER_FRACTION, assert_parity, compare, perturbed_workload, summary

pytestmark = pytest.mark.gpu

from gaussmle_amd import _lib, batch as gb, synth


def _sample_workload(w, idx):
    sc = w.var_map is not None
    return synth.Workload(w.name, w.psf_pixels, w.box, np.asfortranarray(w.data[:, :, idx]), w.truth[:, idx],
                          w.constraints, w.iterations, w.var_map, w.x_corners[idx].copy() if sc else None,
                          w.y_corners[idx].copy() if sc else None)


def _device_fit(w, lanes, precision="fp32"):
    """params / crlb (n, P) and llr (n,); in mixed precision crlb and llr
    are the Float64 final pass's."""
    import torch
    from gaussmle_amd import device as gd
    d = gd.upload_workload(w)
    ny, nx = (0, 0) if w.var_map is None else w.var_map.shape
    out = gd.fit_device(w.psf_pixels, w.box, d["data"], w.constraints, w.iterations, precision, lanes,
                        d.get("var"), ny, nx, d.get("xc"), d.get("yc"))
    torch.cuda.synchronize()
    keys = {"params": "params", "crlb": "crlb", "llr": "llr"}
    if precision == "mixed":
        keys.update(crlb="crlb64", llr="llr64")
    res = {k: out[v].cpu().numpy() for k, v in keys.items()}
    del d, out
    torch.cuda.empty_cache()
    return res


def _check_workload(label, w, precision, stride, stats):
    """Full launch, batch invariance and sample parity for one workload."""
    from oracle import refcpu
    n = w.n
    lanes = _lib.load().gmle_default_lanes(w.psf_pixels.model_id, w.camera, w.box, _lib.PRECISION_IDS[precision], n)
    print(f"[full] {label} n={n} lanes={lanes} {precision}", flush=True)
    full = _device_fit(w, lanes, precision)
    fin = np.all(np.isfinite(full["params"]), axis=1)
    print(f"[full] {label}: {int((~fin).sum())} ROIs with a non-finite parameter", flush=True)

    idx = np.arange(0, n, stride)
    sub = _sample_workload(w, idx)
    alone = _device_fit(sub, lanes, precision)
    for k in ("params", "crlb", "llr"):
        np.testing.assert_array_equal(alone[k], full[k][idx], err_msg=f"{label} batch invariance: {k}")

    r32, r64 = refcpu.fit_workload(sub, refcpu.FP32), refcpu.fit_workload(sub, refcpu.FP64)
    r64p = refcpu.fit_workload(perturbed_workload(sub), refcpu.FP64)
    rmix = refcpu.fit_workload(sub, refcpu.MIXED) if precision == "mixed" else None
    gpu = {k: full[k][idx].astype(np.float64) for k in ("params", "crlb", "llr")}
    rep = compare(w.psf_pixels.model_id, gpu, r32, r64, precision, r64p, rmix)
    print(f"[parity] {label} sample {summary(rep)}", flush=True)
    assert_parity(rep, f"{label} sample")
    # every ROI the Float32 oracle fits to finite values is finite here too
    assert np.all(fin[idx] | ~np.all(np.isfinite(r32["params"]), axis=1))

    if stats:
        for k, name in ((0, "x"), (1, "y")):
            z_full = (full["params"][:, k].astype(np.float64) - w.truth[k]) / full["crlb"][:, k]
            z_ref = (r64["params"][:, k] - sub.truth[k]) / r64["crlb"][:, k]
            z_full, z_ref = z_full[np.isfinite(z_full)], z_ref[np.isfinite(z_ref)]
            ns = len(z_ref)
            print(f"[full] {label} {name}: (fit - truth) / CRLB mean {z_full.mean():+.4f} std {z_full.std():.4f} "
                  f"over {len(z_full)}; oracle on the sample mean {z_ref.mean():+.4f} std {z_ref.std():.4f} over {ns}",
                  flush=True)
            assert abs(z_full.mean() - z_ref.mean()) < 4.0 * z_ref.std() / np.sqrt(ns)
            assert abs(z_full.std() - z_ref.std()) < 4.0 * z_ref.std() / np.sqrt(2.0 * ns)


def test_headline_full_size():
    w = synth.config0(n=1_048_576, seed=0)
    _check_workload("headline", w, "fp32", 256, stats=True)


def test_precision_split_full_size():
    w = synth.config1(n=1_000_000, seed=1)
    _check_workload("precision split", w, "mixed", 250, stats=True)


def test_astigmatic_full_size():
    w = synth.config2(n=4_000_000, seed=2)
    _check_workload("astigmatic", w, "mixed", 1000, stats=False)


def test_uneven_batch_full_size():
    import torch
    from oracle import refcpu
    n = 4_000_000
    mb = gb.config3(n)
    counts = mb.shape_counts()
    lib = _lib.load()
    table = np.zeros(gb.NUM_SHAPES, dtype=np.int32)
    for s, c in enumerate(counts):
        if c:
            m, b = gb.shape_of(s)
            table[s] = lib.gmle_default_lanes(m, mb.camera, b, _lib.PRECISION_IDS["fp32"], int(c))
    print(f"[full] uneven batch n={n} shapes={int((counts > 0).sum())} lanes={sorted(set(table[table > 0]))}",
          flush=True)
    d = gb.upload_batch(mb)
    out = gb.fit_batch_device(mb, d, "fp32", table, 4)
    torch.cuda.synchronize()
    full = {k: v.cpu().numpy() for k, v in out.items()}   # params / crlb (n, 6), llr (n,)
    del d, out
    torch.cuda.empty_cache()

    idx = np.arange(0, n, 1000)
    sub = mb.subset(idx)
    alone = gb.fit_batch(sub, "fp32", table, streams=4)
    np.testing.assert_array_equal(alone.params, full["params"][idx].T, err_msg="batch invariance: params")
    np.testing.assert_array_equal(alone.crlb, full["crlb"][idx].T, err_msg="batch invariance: crlb")
    np.testing.assert_array_equal(alone.llr, full["llr"][idx], err_msg="batch invariance: llr")

    n_well = n_bad = n_fin = 0
    for s, c in enumerate(sub.shape_counts()):
        if not c:
            continue
        m, b = gb.shape_of(s)
        j = sub.roi_indices(m, b)
        w = synth.Workload(f"full_{m}_{b}", sub.psfs[m], b, sub.roi_stack(j, b), sub.truth[:, j], sub.constraints[m],
                           sub.iterations, sub.var_map, sub.x_corners[j].copy(), sub.y_corners[j].copy())
        r32, r64 = refcpu.fit_workload(w, refcpu.FP32), refcpu.fit_workload(w, refcpu.FP64)
        r64p = refcpu.fit_workload(perturbed_workload(w), refcpu.FP64)
        P = len(sub.psfs[m])
        gpu = {"params": full["params"][idx[j], :P].astype(np.float64),
               "crlb": full["crlb"][idx[j], :P].astype(np.float64), "llr": full["llr"][idx[j]].astype(np.float64)}
        rep = compare(m, gpu, r32, r64, "fp32", r64p)
        print(f"[parity] uneven sample model={m} box={b} {summary(rep)}", flush=True)
        n_well += rep["n"] - rep["n_ill"]
        n_bad += rep["n_bad"]
        n_fin += rep["finite_mismatch"]
    print(f"[parity] uneven sample total well_posed={n_well} outliers={n_bad} finite_mismatch={n_fin}", flush=True)
    assert n_bad <= int(OUTLIER_FRACTION * n_well)
    assert n_fin == 0
