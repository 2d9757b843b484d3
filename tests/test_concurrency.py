"""Re-entrancy of the C ABI: the reference's host code may call the library
from several Julia threads at once (ccall runs on the calling thread).  Four
threads call gmle_fit (host arrays; the library serialises its cached device
buffers behind a mutex) and gmle_fit_device (caller-owned buffers, one
stream per thread) at the same time; every result must equal the same call
made alone, bit for bit.  ctypes releases the GIL during the calls, so they
overlap in the library."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import gaussmle_amd as g
from gaussmle_amd import synth
from gaussmle_amd.models import GaussianXYNBS


def workloads():
    return [synth.config0(n=3000, seed=300), synth.config1(n=2000, seed=301),
            synth.generic(GaussianXYNBS(), 2500, 9, seed=302), synth.config0(n=1500, seed=303)]


def host_fit(w):
    f = g.GaussMLEFitter.__new__(g.GaussMLEFitter)
    f.psf_model, f.iterations, f.constraints = w.psf_pixels, w.iterations, w.constraints
    f.precision, f.lanes, f.batch_size, f.device = "fp32", 0, 10_000, g.GPU()
    out = g.fit_arrays(f, w.data, w.psf_pixels, w.var_map, w.x_corners, w.y_corners)
    return out.params.copy(), out.crlb.copy(), out.llr.copy()


def run_threads(fn, items):
    results = [None] * len(items)
    errors = []
    start = threading.Barrier(len(items))

    def body(i):
        try:
            start.wait()
            results[i] = fn(items[i])
        except Exception as e:   # reported after the join
            errors.append(e)

    ts = [threading.Thread(target=body, args=(i,)) for i in range(len(items))]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not errors, errors
    assert all(not t.is_alive() for t in ts)
    return results


def test_concurrent_host_fits_match_serial():
    ws = workloads()
    serial = [host_fit(w) for w in ws]
    for rnd in range(3):
        par = run_threads(host_fit, ws)
        for i, (a, b) in enumerate(zip(par, serial)):
            for x, y, name in zip(a, b, ("params", "crlb", "llr")):
                np.testing.assert_array_equal(x, y, err_msg=f"round {rnd} workload {i} {name}")


def test_concurrent_device_fits_match_serial():
    import torch
    from gaussmle_amd import device as gd
    ws = workloads()
    up = [gd.upload_workload(w) for w in ws]
    torch.cuda.synchronize()

    def dev_fit(i, stream=None):
        w, d = ws[i], up[i]
        ny, nx = (0, 0) if w.var_map is None else w.var_map.shape
        s = stream or torch.cuda.Stream()
        out = gd.fit_device(w.psf_pixels, w.box, d["data"], w.constraints, w.iterations, "fp32", 0,
                            d.get("var"), ny, nx, d.get("xc"), d.get("yc"), stream=s)
        s.synchronize()
        return out["params"].cpu().numpy(), out["crlb"].cpu().numpy(), out["llr"].cpu().numpy()

    serial = [dev_fit(i) for i in range(len(ws))]
    par = run_threads(dev_fit, list(range(len(ws))))
    for i, (a, b) in enumerate(zip(par, serial)):
        for x, y, name in zip(a, b, ("params", "crlb", "llr")):
            np.testing.assert_array_equal(x, y, err_msg=f"workload {i} {name}")


def test_concurrent_batch_device_fits_with_streams_match_serial():
    """gmle_fit_batch_device with streams = 4 from four threads at once: the
    shared fork / join streams are created once (under their own lock) and
    each call holds them for its whole enqueue, so every result equals the
    serial call bit for bit."""
    import torch
    from gaussmle_amd import batch as gb
    mbs = [gb.mixed_workload(3000 + 500 * i, seed=310 + i, boxes=(7, 9, 11), box_weights=(1, 1, 1))
           for i in range(4)]
    ups = [gb.upload_batch(mb) for mb in mbs]
    torch.cuda.synchronize()

    def batch_fit(i):
        s = torch.cuda.Stream()
        out = gb.fit_batch_device(mbs[i], ups[i], "fp32", None, 4, stream=s)
        s.synchronize()
        return out["params"].cpu().numpy(), out["crlb"].cpu().numpy(), out["llr"].cpu().numpy()

    serial = [batch_fit(i) for i in range(4)]
    for rnd in range(3):
        par = run_threads(batch_fit, list(range(4)))
        for i, (a, b) in enumerate(zip(par, serial)):
            for x, y, name in zip(a, b, ("params", "crlb", "llr")):
                np.testing.assert_array_equal(x, y, err_msg=f"round {rnd} batch {i} {name}")
