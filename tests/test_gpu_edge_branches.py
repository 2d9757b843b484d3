"""GPU parity at the reference's numerical edge branches.

(1) The masked Poisson branch.  Every other parity workload has bg >= 1, where
the kernel's fast path provably never meets the reference's 0.01 model
threshold or its 1e4 caps (src/camera_models.jl:23-32) and skips them.  The
workloads of edge_cases.lowbg_cases() (backgrounds of 0.002-0.5 photons/px,
N up to 1e5, interior hot pixels) fail that precondition, so the kernel runs
the branch structure of the reference; the oracle counts how often each
branch binds (printed) and the GPU must meet the parity rule of parity.py at
every lane count, in both precisions.

(2) The Fisher inversion (src/unified_kernel.jl:399-431).  The groups of
edge_cases.fisher_cases() are built so that the outcome does not depend on
rounding: Cholesky on all-zero and flat ROIs, the LU fallback with NaN
(all-zero Fisher matrix) and with zeros (negative definite), and the Inf
CRLB of an exactly singular block.  The GPU must give the oracle's pattern of
NaN / Inf / zero / finite CRLB value for value, the values the construction
fixes exactly, and, on ROIs the parity rule keeps, the oracle's numbers.
"""
import numpy as np
import pytest

import edge_cases as ec
from parity import assert_parity, compare, llr_rounding_scale, perturbed_workload, summary
from test_gpu_parity import gpu_fit_host, lanes_for

pytestmark = pytest.mark.gpu

_W = {}


def _dump(label, res):
    """With GMLE_DUMP_DIR set, keep the GPU outputs for offline analysis."""
    import os
    d = os.environ.get("GMLE_DUMP_DIR")
    if d:
        os.makedirs(d, exist_ok=True)
        np.savez_compressed(os.path.join(d, f"{label}.npz"), **{k: np.asarray(v) for k, v in res.items()})


def lowbg(name):
    if name not in _W:
        _W[name] = ec.lowbg_cases()[name]()
    return _W[name]


@pytest.mark.parametrize("name", list(ec.lowbg_cases()))
@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_masked_poisson_branch_parity(name, precision):
    from oracle import refcpu
    w = lowbg(name)
    br = refcpu.poisson_branches(w)
    r32, r64 = refcpu.fit_workload(w, refcpu.FP32), refcpu.fit_workload(w, refcpu.FP64)
    r64p = refcpu.fit_workload(perturbed_workload(w), refcpu.FP64)
    rmix = refcpu.fit_workload(w, refcpu.MIXED) if precision == "mixed" else None
    # Float32 mode: the LLR formula's rounding scale, and the Float32 final
    # pass's own sensitivity to rounding (parity.py, rule (4))
    scale = llr_rounding_scale(w)
    r32fp = refcpu.final_pass(perturbed_workload(w), r32["params"], refcpu.FP32) if precision == "fp32" else None
    for lanes in lanes_for(w.box):
        print(f"[case] {name} {precision} lanes={lanes} n={w.n}", flush=True)
        g = gpu_fit_host(w, precision, lanes)
        _dump(f"{name}_{precision}_L{lanes}", g)
        rep = compare(w.psf_pixels.model_id, g, r32, r64, precision, r64p, rmix, scale, r32fp)
        print(f"[parity] {name} {precision} L={lanes} {summary(rep)}", flush=True)
        assert_parity(rep, f"{name} {precision} L={lanes}")
    # branch coverage on the ROIs actually compared (well-posed)
    checked = rep["n"] - rep["n_ill"]
    print(f"[branches] {name} {precision}: compared {checked}/{rep['n']} ROIs; ROIs with threshold / cf cap / "
          f"df cap / masked iterations: {[(br[:, q] > 0).sum() for q in range(4)]} "
          f"(pixel terms {br[:, :3].sum(axis=0).tolist()}, iterations {int(br[:, 3].sum())})", flush=True)
    assert checked > 0.4 * rep["n"]


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_fisher_inversion_outcomes_match_oracle(precision):
    from oracle import refcpu
    counts = np.zeros(3, dtype=np.int64)
    for name, w in ec.fisher_cases().items():
        ref = refcpu.fit_workload(w, refcpu.FP32 if precision == "fp32" else refcpu.MIXED)
        counts += np.bincount(ref["path"], minlength=3)
        r32, r64 = refcpu.fit_workload(w, refcpu.FP32), refcpu.fit_workload(w, refcpu.FP64)
        r64p = refcpu.fit_workload(perturbed_workload(w), refcpu.FP64)
        rmix = refcpu.fit_workload(w, refcpu.MIXED) if precision == "mixed" else None
        for lanes in lanes_for(w.box):
            print(f"[case] fisher {name} {precision} lanes={lanes}", flush=True)
            g = gpu_fit_host(w, precision, lanes)
            pg, pr = ec.crlb_pattern(g["crlb"]), ec.crlb_pattern(ref["crlb"])
            assert np.array_equal(pg, pr), (name, lanes, np.unique(pg), np.unique(pr))
            assert np.array_equal(np.isfinite(g["params"]), np.isfinite(ref["params"])), (name, lanes)
            # values on ROIs the parity rule keeps: position-free groups (flat
            # and empty ROIs, whose x and y drift with rounding) are ill-posed
            rep = compare(w.psf_pixels.model_id, g, r32, r64, precision, r64p, rmix)
            assert rep["n_bad"] == 0, f"{name} L={lanes}: {summary(rep)}"
            # finite non-zero CRLBs against the same-precision oracle on those ROIs
            fin = (pr == "f") & rep["well"][:, None]
            ok = fin if precision == "mixed" else fin & (np.abs(r32["crlb"] / r64["crlb"] - 1) <= 1e-4)
            if ok.any():
                err = np.abs(g["crlb"][ok] / ref["crlb"][ok] - 1)
                assert err.max() <= 1e-4, (name, lanes, err.max())
            # values the construction fixes exactly
            if name == "point_fp_xynb_11":
                np.testing.assert_array_equal(g["params"], ref["params"])
                assert np.all(g["llr"] == 0)
            if name in ("no_signal_xynb_11", "neg_variance_xynb_11"):
                assert np.all(g["llr"] == 0)   # no pixel with model > 0 / total variance > 0
            if name == "no_signal_xynb_11":
                assert np.all(g["params"][:, 2:] == 0)
            if name == "zero_xynb_11":
                np.testing.assert_array_equal(g["params"][:, 2:], ref["params"][:, 2:])
        print(f"[fisher] {name} {precision}: paths (Cholesky, LU, Inf) "
              f"{np.bincount(ref['path'], minlength=3).tolist()}; CRLB pattern {sorted(set(map(''.join, pr)))}",
              flush=True)
    print(f"[fisher] {precision}: inversion paths over all groups (Cholesky, LU, Inf) = {counts.tolist()}",
          flush=True)
    if precision == "fp32":
        assert np.all(counts > 0), counts
    else:
        assert counts[0] > 0 and counts[1] > 0, counts
