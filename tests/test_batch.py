"""Mixed-shape batches (gmle_fit_batch / gmle_fit_batch_device).

The batch path runs every shape's own fitting kernel over ROI indices grouped
on the device, so each ROI's result must be bit-identical to a plain gmle_fit
call on the same ROIs with the same lanes per ROI; on top of that the batch
is checked against the CPU oracle per shape with the tolerances of
tests/parity.py.  Also: host vs device entry points, stream counts, device-side
counting, mixed precision, unknown shapes and short shape counts.
"""
import numpy as np
import pytest

from parity import OUTLIER_FRACTION, compare, perturbed_workload, summary

pytestmark = pytest.mark.gpu

import gaussmle_amd as g
from gaussmle_amd import _lib, batch as gb

_B = {}


def mixed(name):
    if name not in _B:
        if name == "scmos":
            _B[name] = gb.mixed_workload(3000, seed=300, boxes=(7, 9, 11, 13, 15), models=(0, 1, 2), scmos=True)
        elif name == "ideal":
            _B[name] = gb.mixed_workload(2000, seed=301, boxes=(5, 8, 11, 14, 16), models=(0, 1, 2, 3), scmos=False)
        else:
            raise KeyError(name)
    return _B[name]


def lanes_table(mb, choice):
    """Per-shape lanes: 'default' (0 = library choice), 'max' or 'min' supported."""
    lib = _lib.load()
    t = np.zeros(gb.NUM_SHAPES, dtype=np.int32)
    for s, c in enumerate(mb.shape_counts()):
        if c == 0:
            continue
        m, b = gb.shape_of(s)
        ok = [L for L in (1, 2, 4, 8, 16) if lib.gmle_lanes_supported(b, L)]
        t[s] = 0 if choice == "default" else (max(ok) if choice == "max" else min(ok))
    return t


def resolved_lanes(mb, table, precision):
    lib = _lib.load()
    out = {}
    counts = mb.shape_counts()
    for s, c in enumerate(counts):
        if c:
            m, b = gb.shape_of(s)
            out[(m, b)] = int(table[s]) if table[s] else int(
                lib.gmle_default_lanes(m, mb.camera, b, _lib.PRECISION_IDS[precision], int(c)))
    return out


def plain_fit(mb, m, b, idx, precision, lanes):
    f = g.GaussMLEFitter.__new__(g.GaussMLEFitter)
    f.psf_model, f.iterations, f.constraints = mb.psfs[m], mb.iterations, mb.constraints[m]
    f.precision, f.lanes, f.batch_size, f.device = precision, lanes, 10_000, g.GPU()
    stack = mb.roi_stack(idx, b)
    if mb.camera:
        return g.fit_arrays(f, stack, mb.psfs[m], mb.var_map, mb.x_corners[idx], mb.y_corners[idx])
    return g.fit_arrays(f, stack, mb.psfs[m])


def device_result(out, precision):
    import torch
    torch.cuda.synchronize()
    r = {k: v.cpu().numpy().T for k, v in out.items() if v.dim() == 2}
    r.update({k: v.cpu().numpy() for k, v in out.items() if v.dim() == 1})
    return r


def assert_same_as_plain(mb, res, precision, lanes):
    """res: dict params/crlb (6, n) arrays and llr (n,) from a batch call."""
    for (m, b), L in lanes.items():
        idx = mb.roi_indices(m, b)
        P = len(mb.psfs[m])
        ref = plain_fit(mb, m, b, idx, precision, L)
        print(f"[batch] model={m} box={b} n={len(idx)} L={L}", flush=True)
        np.testing.assert_array_equal(res["params"][:P, idx], ref.params)
        np.testing.assert_array_equal(res["crlb"][:P, idx], ref.crlb)
        np.testing.assert_array_equal(res["llr"][idx], ref.llr)
        assert np.all(np.isnan(res["params"][P:, idx])) and np.all(np.isnan(res["crlb"][P:, idx]))
        if precision == "mixed":
            np.testing.assert_array_equal(res["crlb64"][:P, idx], ref.crlb64)
            np.testing.assert_array_equal(res["llr64"][idx], ref.llr64)


@pytest.mark.parametrize("name", ["scmos", "ideal"])
@pytest.mark.parametrize("choice", ["default", "max", "min"])
def test_batch_device_bit_identical_to_plain(name, choice):
    mb = mixed(name)
    table = lanes_table(mb, choice)
    d = gb.upload_batch(mb)
    out = gb.fit_batch_device(mb, d, "fp32", table, streams=4)
    assert_same_as_plain(mb, device_result(out, "fp32"), "fp32", resolved_lanes(mb, table, "fp32"))


@pytest.mark.parametrize("name", ["scmos", "ideal"])
def test_batch_host_mixed_precision_bit_identical_to_plain(name):
    mb = mixed(name)
    r = gb.fit_batch(mb, "mixed", None, streams=3)
    res = {"params": r.params, "crlb": r.crlb, "llr": r.llr, "crlb64": r.crlb64, "llr64": r.llr64}
    assert_same_as_plain(mb, res, "mixed", resolved_lanes(mb, np.zeros(gb.NUM_SHAPES, np.int32), "mixed"))


def test_batch_host_equals_device_streams_and_counting():
    mb = mixed("scmos")
    table = lanes_table(mb, "default")
    h = gb.fit_batch(mb, "fp32", table, streams=1)
    d = gb.upload_batch(mb)
    for streams, host_counts in ((1, True), (2, False), (4, True), (4, False)):
        print(f"[batch] streams={streams} host_counts={host_counts}", flush=True)
        r = device_result(gb.fit_batch_device(mb, d, "fp32", table, streams, host_counts=host_counts), "fp32")
        np.testing.assert_array_equal(r["params"], h.params)
        np.testing.assert_array_equal(r["crlb"], h.crlb)
        np.testing.assert_array_equal(r["llr"], h.llr)


@pytest.mark.parametrize("name", ["scmos", "ideal"])
def test_batch_against_oracle(name):
    from oracle import refcpu
    mb = mixed(name)
    r = gb.fit_batch(mb, "fp32", None, streams=4)
    # The outlier rule of tests/parity.py applies to the batch as a whole
    # (per shape, a few dozen ROIs would allow no outlier at all).
    n_well = n_bad = n_fin = 0
    for (m, b) in resolved_lanes(mb, np.zeros(gb.NUM_SHAPES, np.int32), "fp32"):
        idx = mb.roi_indices(m, b)
        w = g.synth.Workload(f"{name}_{m}_{b}", mb.psfs[m], b, mb.roi_stack(idx, b), mb.truth[:, idx],
                             mb.constraints[m], mb.iterations, mb.var_map,
                             None if mb.x_corners is None else mb.x_corners[idx].copy(),
                             None if mb.y_corners is None else mb.y_corners[idx].copy())
        r32, r64 = refcpu.fit_workload(w, refcpu.FP32), refcpu.fit_workload(w, refcpu.FP64)
        r64p = refcpu.fit_workload(perturbed_workload(w), refcpu.FP64)
        P = len(mb.psfs[m])
        res = {"params": r.params[:P, idx].T.astype(np.float64), "crlb": r.crlb[:P, idx].T.astype(np.float64),
               "llr": r.llr[idx].astype(np.float64)}
        rep = compare(m, res, r32, r64, "fp32", r64p)
        print(f"[parity] batch {name} model={m} box={b} {summary(rep)}", flush=True)
        n_well += rep["n"] - rep["n_ill"]
        n_bad += rep["n_bad"]
        n_fin += rep["finite_mismatch"]
    print(f"[parity] batch {name} total well_posed={n_well} outliers={n_bad} finite_mismatch={n_fin}", flush=True)
    assert n_bad <= int(OUTLIER_FRACTION * n_well)
    assert n_fin == 0


def test_batch_unknown_shapes_and_short_counts_give_nan():
    """Device entry point: a ROI of an unknown box size or model, and ROIs
    beyond their shape's count, come back as NaN; the others are unaffected."""
    import torch
    mb = mixed("ideal")
    models = mb.models.copy()
    bad = np.array([3, 17, 40])
    models[bad] = 7                                   # unknown model
    mb2 = gb.MixedBatch(models, mb.boxes, mb.data, mb.offsets, mb.psfs, mb.constraints, mb.iterations,
                        truth=mb.truth)
    d = gb.upload_batch(mb2)
    counts = mb.shape_counts().copy()
    for k in bad:                                     # unknown-model ROIs are not in any shape
        counts[gb.shape_index(int(mb.models[k]), int(mb.boxes[k]))] -= 1
    s_short = int(np.argmax(counts))
    counts[s_short] -= 2                              # two ROIs of the largest shape left over
    d["counts"] = counts
    table = np.zeros(gb.NUM_SHAPES, np.int32)      # fixed lanes: the counts must not change the choice
    for (m, b), L in resolved_lanes(mb, table, "fp32").items():
        table[gb.shape_index(m, b)] = L
    r = device_result(gb.fit_batch_device(mb2, d, "fp32", table, 4), "fp32")
    ref = gb.fit_batch(mb, "fp32", table, 4)
    nan_rows = np.all(np.isnan(r["params"]), axis=0) & np.isnan(r["llr"])
    assert nan_rows[bad].all()
    m_s, b_s = gb.shape_of(s_short)
    in_shape = (mb.models == m_s) & (mb.boxes == b_s)
    assert nan_rows[in_shape].sum() == 2
    keep = ~nan_rows
    keep[bad] = False
    np.testing.assert_array_equal(r["params"][:, keep], ref.params[:, keep])
    np.testing.assert_array_equal(r["llr"][keep], ref.llr[keep])


@pytest.mark.parametrize("name", ["scmos", "ideal"])
@pytest.mark.parametrize("choice", ["default", "min", "max"])
def test_batch_one_launch_mode_matches_plain(name, choice):
    """streams = 0: every shape in one multi-shape launch, each with the
    lanes gmle_one_launch_lanes reports; bit-identical to per-shape launches."""
    lib = _lib.load()
    mb = mixed(name)
    table = lanes_table(mb, choice)
    d = gb.upload_batch(mb)
    out = gb.fit_batch_device(mb, d, "fp32", table, streams=0)
    res = device_result(out, "fp32")
    used = {}
    for s, c in enumerate(mb.shape_counts()):
        if c:
            m, b = gb.shape_of(s)
            L = lib.gmle_one_launch_lanes(m, mb.camera, b, int(table[s]))
            used[(m, b)] = L if L > 0 else resolved_lanes(mb, table, "fp32")[(m, b)]
    print(f"[batch] one launch {name} {choice}: {used}", flush=True)
    assert_same_as_plain(mb, res, "fp32", used)
    h = gb.fit_batch(mb, "fp32", table, streams=0)
    np.testing.assert_array_equal(h.params, res["params"])
    np.testing.assert_array_equal(h.llr, res["llr"])
