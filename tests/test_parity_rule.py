"""The GPU parity checker (tests/parity.py) itself, exercised on the CPU with
mistakes planted in copies of the oracle's output."""
import numpy as np
import pytest

from parity import assert_parity, compare


@pytest.fixture(scope="module")
def refs(oracle):
    from gaussmle_amd import synth
    w = synth.config0(n=600, seed=77)
    return oracle.fit_workload(w, oracle.FP32), oracle.fit_workload(w, oracle.FP64)


def _copy(r):
    return {k: np.array(v, copy=True) for k, v in r.items()}


def test_reference_passes_itself(refs):
    r32, r64 = refs
    rep = compare(0, _copy(r32), r32, r64, "fp32")
    assert rep["n_bad"] == 0 and rep["finite_mismatch"] == 0
    assert_parity(rep)


@pytest.mark.parametrize("field,k,delta", [("params", 0, 2e-3), ("params", 1, -2e-3), ("params", 2, 3e-4),
                                           ("crlb", 3, 3e-4), ("llr", None, 3e-4)])
def test_planted_errors_are_flagged(refs, field, k, delta):
    r32, r64 = refs
    g = _copy(r32)
    rep0 = compare(0, g, r32, r64, "fp32")
    roi = int(np.where(~rep0["bad"])[0][0])
    if field == "llr":
        g["llr"][roi] += delta * max(1.0, abs(g["llr"][roi]))
    elif k < 2 and field == "params":
        g[field][roi, k] += delta
    else:
        g[field][roi, k] *= 1 + delta
    rep = compare(0, g, r32, r64, "fp32")
    assert rep["bad"][roi] and rep["n_bad"] == 1


def test_nan_where_reference_is_finite_fails(refs):
    r32, r64 = refs
    g = _copy(r32)
    g["params"][5, :] = np.nan
    rep = compare(0, g, r32, r64, "fp32")
    with pytest.raises(AssertionError):
        assert_parity(rep)


def test_perturb_ulp_moves_every_nonzero_value_by_that_many_ulp():
    from parity import perturb_ulp
    d = np.float32([0.0, 1.0, -3.5, 1e-3, 250.0])
    p = perturb_ulp(d, ulps=4, seed=1)
    assert p[0] == 0.0
    # same sign, so the distance in ulp is the distance of the bit patterns
    steps = np.abs(p[1:].view(np.int32).astype(np.int64) - d[1:].view(np.int32).astype(np.int64))
    np.testing.assert_array_equal(steps, 4)


def test_ill_conditioned_rois_are_excluded(refs):
    """An ROI whose perturbed Float64 fit moves beyond the tolerance is not
    held to it; others still are."""
    r32, r64 = refs
    p = _copy(r64)
    p["params"][7, 2] *= 1 + 5e-4          # the perturbed fit's N moved: ROI 7 is ill-conditioned
    g = _copy(r32)
    g["params"][7, 2] *= 1 + 5e-4
    g["params"][9, 2] *= 1 + 5e-4
    rep = compare(0, g, r32, r64, "fp32", p)
    assert not rep["bad"][7] and rep["bad"][9] and rep["n_ill"] >= 1


def test_rounding_limited_float32_crlb_at_the_reported_estimate_is_set_aside(refs):
    """Rule (5): where the oracle's Float32 final pass at the GPU's estimate
    misses the CRLB tolerance, that ROI is not held to it; an ROI where it
    meets the tolerance still is."""
    r32, r64 = refs
    g = _copy(r32)
    g["crlb"][11, 1] *= 1 + 3e-4
    g["crlb"][13, 1] *= 1 + 3e-4
    at_gpu = {"crlb": np.array(r64["crlb"], copy=True)}
    at_gpu["crlb"][11, 1] *= 1 + 5e-4     # the Float32 evaluation at ROI 11's estimate is itself off
    rep = compare(0, g, r32, r64, "fp32", ref32g=at_gpu)
    assert not rep["well"][11] and not rep["bad"][11]
    assert rep["well"][13] and rep["bad"][13]
    # without the evaluation at the GPU's estimate both count
    rep0 = compare(0, g, r32, r64, "fp32")
    assert rep0["bad"][11] and rep0["bad"][13]
