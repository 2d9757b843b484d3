"""The host-array entry points' upload pipeline (gmle_fit, gmle_fit_batch).

Inputs larger than one pinned staging buffer (32 MiB) stream through the
ring in several chunks, each fitted as soon as it lands.  Results must be
bit-identical to the device-array entry points on the same ROIs with the
same lanes; a NaN or Inf in the last chunk must still reject the call with
GMLE_ERR_DATA and leave the caller's output arrays untouched (the
reference validates before fitting, src/interface.jl:26-29)."""
import numpy as np
import pytest

import gaussmle_amd as g
from gaussmle_amd import _lib, synth
from gaussmle_amd import batch as gb

pytestmark = pytest.mark.gpu


def _host(w, lanes, precision="fp32"):
    f = g.GaussMLEFitter.__new__(g.GaussMLEFitter)
    f.psf_model, f.iterations, f.constraints = w.psf_pixels, w.iterations, w.constraints
    f.precision, f.lanes, f.batch_size, f.device = precision, lanes, 10_000, g.GPU()
    return g.fit_arrays(f, w.data, w.psf_pixels, w.var_map, w.x_corners, w.y_corners)


@pytest.mark.parametrize("case", ["xynb_ideal_11", "xynbs_scmos_13_mixed"])
def test_chunked_host_path_equals_device_path(case):
    import torch
    from gaussmle_amd import device as gd
    if case == "xynb_ideal_11":
        w, precision = synth.config0(n=200_000, seed=700), "fp32"    # 97 MB: 3 chunks
    else:
        w, precision = synth.config1(n=120_000, seed=701), "mixed"   # 81 MB: 3 chunks
    lanes = _lib.load().gmle_default_lanes(w.psf_pixels.model_id, w.camera, w.box, _lib.PRECISION_IDS[precision], w.n)
    print(f"[case] host pipeline {case} n={w.n} lanes={lanes}", flush=True)
    h = _host(w, lanes, precision)
    d = gd.upload_workload(w)
    ny, nx = (0, 0) if w.var_map is None else w.var_map.shape
    o = gd.fit_device(w.psf_pixels, w.box, d["data"], w.constraints, w.iterations, precision, lanes, d.get("var"),
                      ny, nx, d.get("xc"), d.get("yc"))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(h.params.T, o["params"].cpu().numpy())
    np.testing.assert_array_equal(h.crlb.T, o["crlb"].cpu().numpy())
    np.testing.assert_array_equal(h.llr, o["llr"].cpu().numpy())
    if precision == "mixed":
        np.testing.assert_array_equal(h.crlb64.T, o["crlb64"].cpu().numpy())
        np.testing.assert_array_equal(h.llr64, o["llr64"].cpu().numpy())


def test_nonfinite_value_in_last_chunk_rejects_and_leaves_outputs():
    lib = _lib.load()
    w = synth.config0(n=200_000, seed=702)
    c = w.constraints
    for bad_value in (np.nan, np.inf, -np.inf):
        data = np.asfortranarray(w.data.copy())
        data[5, 6, w.n - 3] = bad_value
        flat = data.reshape(-1, order="F")
        out = [np.full(w.n * 4, -7.0, np.float32), np.full(w.n * 4, -7.0, np.float32),
               np.full(w.n, -7.0, np.float32)]
        rc = lib.gmle_fit(0, _lib.fptr(w.psf_pixels.psf_array()), 11, w.n, _lib.fptr(flat), 0, None, 0, 0, None,
                          None, _lib.fptr(c.lower), _lib.fptr(c.upper), _lib.fptr(c.max_step), 20, 0, 0,
                          *map(_lib.fptr, out), None, None)
        assert rc == _lib.GMLE_ERR_DATA, rc
        assert "NaN or Inf" in _lib.last_error()
        assert all(np.all(o == -7.0) for o in out)
    # and the next call on the same buffers fits normally
    h = _host(w.subset(5000), 1)
    assert np.all(np.isfinite(h.params))


def test_chunked_host_batch_equals_device_batch():
    import torch
    mb = gb.mixed_workload(150_000, seed=703, boxes=(7, 9, 11, 13, 15), box_weights=(1, 2, 4, 2, 1))
    print(f"[case] host pipeline batch n={mb.n} pixels={mb.data.size}", flush=True)
    h = gb.fit_batch(mb, "fp32", None, 4)
    d = gb.upload_batch(mb)
    o = gb.fit_batch_device(mb, d, "fp32", None, 4)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(h.params, o["params"].cpu().numpy().T)
    np.testing.assert_array_equal(h.crlb, o["crlb"].cpu().numpy().T)
    np.testing.assert_array_equal(h.llr, o["llr"].cpu().numpy())
