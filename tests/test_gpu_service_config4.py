"""The resident service on the streaming workload at full size: frames of a
2048 x 2048 uint16 sCMOS sensor, each with a log-uniform 50-5000 ROIs of 9, 11,
13 or 15 px, GaussianXYNB and GaussianXYNBS, 20 iterations.

Each request's results must be bit-identical alone and in a mix (all frames
submitted at once), and the service's own outputs are held to the oracle under
the parity rule of parity.py, shape by shape, on the ROIs converted to
electrons as the reference's to_electrons does (src/camera_models.jl:110-113
with per-pixel maps; bit-identical to the service's device conversion, which
test_service.py checks) and fitted with the reference's box-11 default
constraints (src/interface.jl:141)."""
import numpy as np
import pytest

from parity import OUTLIER_FRACTION, compare, llr_rounding_scale, perturbed_workload, summary

import gaussmle_amd as g
from gaussmle_amd import service as sv

pytestmark = pytest.mark.gpu

N_FRAMES = 24
SAMPLE_STRIDE = 1          # every ROI of every frame goes to the oracle
_STATE = {}


def workload():
    if "frames" not in _STATE:
        maps = sv.sensor_maps(seed=70, ny=2048, nx=2048)
        frs = sv.make_frames(N_FRAMES, seed=71, maps=maps, rois=(50, 5000))
        _STATE["maps"], _STATE["frames"] = maps, frs
    return _STATE["maps"], _STATE["frames"]


def _results(reqs):
    out = [r.results() for r in reqs]
    return (np.concatenate([o[0] for o in out], axis=1), np.concatenate([o[1] for o in out], axis=1),
            np.concatenate([o[2] for o in out]))


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_service_config4_against_oracle(precision):
    from oracle import refcpu
    maps, frs = workload()
    n_total = sum(f.n for f in frs)
    print(f"[case] service config4 {precision} frames={len(frs)} rois={n_total} "
          f"counts={[f.n for f in frs]}", flush=True)
    assert min(f.n for f in frs) >= 50 and max(f.n for f in frs) <= 5000
    psfs, cons, table = sv.default_setup(lanes=2)
    s = sv.Service(psfs, cons, table, hbm_budget=2 << 30, slots=3, precision=precision)
    reqs = [sv.Request(f) for f in frs]
    try:
        s.calibrate(*maps)
        # all frames at once (continuous batching mixes them into batches)
        for t in [s.submit(r) for r in reqs]:
            s.wait(t)
        mix = [r.results() for r in reqs]
        # each frame alone
        for i, r in enumerate(reqs):
            s.wait(s.submit(r))
            for a, b, name in zip(r.results(), mix[i], ("params", "crlb", "llr")):
                np.testing.assert_array_equal(a, b, err_msg=f"frame {i} alone vs mix: {name}")
        params, crlb, llr = _results(reqs)
        info = s.info()
        print(f"[service] {info}", flush=True)
    finally:
        for r in reqs:
            r.close()
        s.close()

    mb = sv.frames_batch(frs, maps, psfs, cons, 20)
    sample = np.arange(0, mb.n, SAMPLE_STRIDE)
    n_well = n_bad = n_fin = n_seen = 0
    for m in (0, 1):
        for b in (9, 11, 13, 15):
            idx = sample[(mb.models[sample] == m) & (mb.boxes[sample] == b)]
            if len(idx) == 0:
                continue
            w = g.synth.Workload(f"config4_{m}_{b}", mb.psfs[m], b, mb.roi_stack(idx, b), mb.truth[:, idx],
                                 mb.constraints[m], mb.iterations, mb.var_map, mb.x_corners[idx].copy(),
                                 mb.y_corners[idx].copy())
            r32, r64 = refcpu.fit_workload(w, refcpu.FP32), refcpu.fit_workload(w, refcpu.FP64)
            r64p = refcpu.fit_workload(perturbed_workload(w), refcpu.FP64)
            rmix = refcpu.fit_workload(w, refcpu.MIXED) if precision == "mixed" else None
            r32fp = refcpu.final_pass(perturbed_workload(w), r32["params"], refcpu.FP32) \
                if precision == "fp32" else None
            P = len(mb.psfs[m])
            res = {"params": params[:P, idx].T.astype(np.float64), "crlb": crlb[:P, idx].T.astype(np.float64),
                   "llr": llr[idx].astype(np.float64)}
            rep = compare(m, res, r32, r64, precision, r64p, rmix, llr_rounding_scale(w), r32fp)
            print(f"[parity] service config4 {precision} model={m} box={b} {summary(rep)}", flush=True)
            n_seen += rep["n"]
            n_well += rep["n"] - rep["n_ill"]
            n_bad += rep["n_bad"]
            n_fin += rep["finite_mismatch"]
    print(f"[parity] service config4 {precision} total compared={n_seen} well_posed={n_well} outliers={n_bad} "
          f"finite_mismatch={n_fin}", flush=True)
    assert n_seen == len(sample)
    assert n_well >= 0.95 * n_seen
    assert n_bad <= int(OUTLIER_FRACTION * n_well)
    assert n_fin == 0
