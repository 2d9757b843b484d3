"""gmle_fit_batch_device_noalloc (the allocation-free batch entry point)
against gmle_fit_batch_device with a caller workspace: bit-identical results,
on two calls in a row (the second with fewer ROIs, so the static workspace
holds the first call's contents), in the one-launch and in-order modes and in
mixed precision; oversize batches and stream counts above 1 are rejected
before any device work."""
import numpy as np
import pytest

from gaussmle_amd import _lib
from gaussmle_amd import batch as gb

pytestmark = pytest.mark.gpu


def _out(n, mixed, dev="cuda:0"):
    import torch
    o = {"params": torch.full((n, gb.STRIDE), -7.0, dtype=torch.float32, device=dev),
         "crlb": torch.full((n, gb.STRIDE), -7.0, dtype=torch.float32, device=dev),
         "llr": torch.full((n,), -7.0, dtype=torch.float32, device=dev)}
    if mixed:
        o["crlb64"] = torch.full((n, gb.STRIDE), -7.0, dtype=torch.float64, device=dev)
        o["llr64"] = torch.full((n,), -7.0, dtype=torch.float64, device=dev)
    return o


def _same(a, b, label):
    for k in a:
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        assert np.array_equal(x, y, equal_nan=True), f"{label}: {k}"


@pytest.mark.parametrize("precision,streams,host_counts", [("fp32", 0, True), ("fp32", 1, True),
                                                           ("mixed", 1, True), ("fp32", 1, False)])
def test_noalloc_matches_workspace_entry_point(precision, streams, host_counts):
    import torch
    print(f"[case] noalloc precision={precision} streams={streams} host_counts={host_counts}", flush=True)
    big = gb.mixed_workload(30_000, seed=610, boxes=(7, 9, 11, 13), box_weights=(1, 2, 2, 1))
    small = gb.mixed_workload(9_000, seed=611, boxes=(9, 11, 15), box_weights=(1, 1, 1))
    for label, mb in (("first call", big), ("second, smaller call", small)):
        d = gb.upload_batch(mb)
        ref = gb.fit_batch_device(mb, d, precision, None, 1 if streams else 0, host_counts=host_counts)
        got = gb.fit_batch_device_noalloc(mb, d, precision, None, streams, out=_out(mb.n, precision == "mixed"),
                                          host_counts=host_counts)
        torch.cuda.synchronize()
        _same(ref, got, label)
        p = got["params"].cpu().numpy()
        assert np.isfinite(p[:, :4]).all(), label


def test_noalloc_rejects_oversize_and_streams():
    import torch
    lib = _lib.load()
    cap = 4_194_176
    assert lib.gmle_batch_workspace_bytes(cap) <= 16 << 20 < lib.gmle_batch_workspace_bytes(cap + 1)
    mb = gb.mixed_workload(1000, seed=612, boxes=(9,), box_weights=(1,))
    d = gb.upload_batch(mb)
    out = _out(mb.n, False)
    lo, up, ms = (t.reshape(-1) for t in mb.constraint_tables())
    psf = mb.psf_table().reshape(-1)
    # n beyond the static workspace (checked before any array is read), and
    # stream counts that would need the library's extra streams
    for n, streams in ((cap + 1, 1), (mb.n, 2), (mb.n, 4)):
        rc = lib.gmle_fit_batch_device_noalloc(
            n, d["models"].data_ptr(), d["boxes"].data_ptr(), d["offsets"].data_ptr(), d["data"].data_ptr(),
            _lib.fptr(psf), _lib.fptr(lo), _lib.fptr(up), _lib.fptr(ms),
            1, d["var"].data_ptr(), 256, 256, d["xc"].data_ptr(), d["yc"].data_ptr(), 20, 0, None, streams,
            None, out["params"].data_ptr(), out["crlb"].data_ptr(), out["llr"].data_ptr(), None, None,
            torch.cuda.current_stream().cuda_stream)
        assert rc == -1, (n, streams, rc)
        msg = _lib.last_error()
        assert ("at most" in msg) if n > cap else ("streams" in msg), msg
    # out untouched by the rejected calls
    torch.cuda.synchronize()
    assert np.all(out["params"].cpu().numpy() == -7.0)
