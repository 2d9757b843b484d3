"""The crafted edge-branch workloads (tests/edge_cases.py) do what they are
built for, on the oracle (CPU): the masked Poisson branch, its 0.01
threshold and both 1e4 caps (src/camera_models.jl:23-32) are taken on ROIs
the GPU parity rule keeps, and the Fisher-matrix groups reach every
inversion outcome (src/unified_kernel.jl:399-431) in the predicted pattern.
The GPU tests (test_gpu_edge_branches.py) then compare the HIP kernels with
the oracle on the same inputs."""
import numpy as np
import pytest

import edge_cases as ec
from parity import _bad, _errors, perturbed_workload


def test_lowbg_workloads_drive_every_poisson_branch(oracle):
    totals = np.zeros(4, dtype=np.int64)
    for name, make in ec.lowbg_cases().items():
        w = make()
        br = oracle.poisson_branches(w)
        r32, r64 = oracle.fit_workload(w, oracle.FP32), oracle.fit_workload(w, oracle.FP64)
        r64p = oracle.fit_workload(perturbed_workload(w), oracle.FP64)
        mid = w.psf_pixels.model_id
        well = ~_bad(_errors(mid, r32, r32, r64, "fp32")) & ~_bad(_errors(mid, r64p, r64, r64, "mixed"))
        hit = np.array([(well & (br[:, q] > 0)).sum() for q in range(4)])
        print(f"[branches] {name}: well-posed {well.sum()}/{w.n}; well-posed ROIs with threshold / cf cap / "
              f"df cap / masked iterations: {hit.tolist()}")
        totals += hit
        assert hit[3] > 0.3 * w.n          # the GPU runs its masked path on a large share
    assert np.all(totals > 0), totals
    assert totals[1] >= 50 and totals[2] >= 50


def test_fisher_groups_reach_every_inversion_outcome(oracle):
    expect32 = {"zero_xynb_11": (0, "f"), "zero_xynbs_9": (0, "f"), "constant_xynb_11": (0, "f"),
                "point_fp_xynb_11": (2, "i"), "no_signal_xynb_11": (1, "n"), "neg_variance_xynb_11": (1, "0")}
    expect64 = dict(expect32, point_fp_xynb_11=(0, "f"))
    for name, w in ec.fisher_cases().items():
        for prec, expect in ((oracle.FP32, expect32), (oracle.MIXED, expect64), (oracle.FP64, expect64)):
            r = oracle.fit_workload(w, prec)
            path, pat = expect[name]
            assert np.all(r["path"] == path), (name, prec, np.bincount(r["path"]))
            assert np.all(ec.crlb_pattern(r["crlb"]) == pat), (name, prec)
    # the exact values the construction fixes
    cases = ec.fisher_cases()
    r = oracle.fit_workload(cases["point_fp_xynb_11"], oracle.FP32)
    w = cases["point_fp_xynb_11"]
    flat = w.data.reshape(-1, w.n, order="F").argmax(axis=0)
    np.testing.assert_array_equal(r["params"][:, 0], flat // w.box + 1)   # x = column of the hot pixel
    np.testing.assert_array_equal(r["params"][:, 1], flat % w.box + 1)    # y = its row
    assert np.all(r["params"][:, 2] == 256) and np.all(r["params"][:, 3] == 0) and np.all(r["llr"] == 0)
    rm = oracle.fit_workload(w, oracle.MIXED)
    reg = 1e-10 * 2.0 ** -8
    np.testing.assert_allclose(rm["crlb"][:, 0], 1 / np.sqrt(reg), rtol=1e-12)
    # [[a + r, a], [a, a + r]]^-1 has diagonal (a + r) / (2 a r + r^2)
    a = 2.0 ** -8
    np.testing.assert_allclose(rm["crlb"][:, 2], np.sqrt((a + reg) / (2 * a * reg + reg * reg)), rtol=1e-6)
    r = oracle.fit_workload(cases["no_signal_xynb_11"], oracle.FP32)
    assert np.all(r["params"][:, 2:] == 0) and np.all(r["llr"] == 0)
    r = oracle.fit_workload(cases["zero_xynb_11"], oracle.FP32)
    assert np.all(r["params"][:, 2] == 100) and np.all(r["params"][:, 3] == np.float32(0.01))
