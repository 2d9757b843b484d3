"""The resident fitting service (gmle_service_*).

Results must be bit-identical to the batch path on the same ROIs converted to
electrons on the host ((adu - offset) * gain in Float32), and each request's
results must be the same alone and in a mix: submitted one at a time, all at
once, and replayed at a rate, with every graph replay checked (the service
re-launches each slot's HIP graph for every batch, with eager work between).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from gaussmle_amd import batch as gb
from gaussmle_amd import service as sv

_STATE = {}


def frames():
    if "frames" not in _STATE:
        maps = sv.sensor_maps(seed=40, ny=512, nx=640)
        _STATE["maps"] = maps
        _STATE["frames"] = sv.make_frames(10, seed=41, maps=maps, rois=(20, 600))
    return _STATE["maps"], _STATE["frames"]


LANE_TABLES = {
    "lanes_1_2_4": {(0, 9): 1, (0, 11): 2, (0, 13): 4, (0, 15): 2, (1, 9): 2, (1, 11): 4, (1, 13): 2, (1, 15): 4},
    "lanes_8_16": {(0, 9): 16, (0, 11): 8, (0, 13): 16, (0, 15): 8, (1, 9): 16, (1, 11): 16, (1, 13): 8,
                   (1, 15): 16},
}


@pytest.fixture(params=list(LANE_TABLES))
def service(request):
    maps, _ = frames()
    psfs, cons, table = sv.default_setup(lanes=LANE_TABLES[request.param])
    s = sv.Service(psfs, cons, table, hbm_budget=1 << 30, slots=3)
    s.calibrate(*maps)
    yield s
    s.close()


def reference(s, frs):
    """The batch path on host-converted electrons, with the lanes the service
    uses: Float32 services fit every shape in one multi-shape kernel (the
    batch path's one-launch mode), mixed-precision ones per shape."""
    maps, _ = frames()
    mb = sv.frames_batch(frs, maps, s.psfs, s.constraints, s.iterations)
    if s.precision == "fp32":
        return gb.fit_batch(mb, "fp32", s.table, streams=0)
    return gb.fit_batch(mb, s.precision, s.table, streams=2)


def split(res, frs):
    out, k = [], 0
    for f in frs:
        out.append((res.params[:, k:k + f.n], res.crlb[:, k:k + f.n], res.llr[k:k + f.n]))
        k += f.n
    return out


def assert_same(got, want, label):
    for a, b, name in zip(got, want, ("params", "crlb", "llr")):
        np.testing.assert_array_equal(a, b, err_msg=f"{label} {name}")


def test_service_matches_batch_path_alone_and_in_mix(service):
    maps, frs = frames()
    want = split(reference(service, frs), frs)
    reqs = [sv.Request(f) for f in frs]
    try:
        # one at a time
        for i, r in enumerate(reqs):
            service.wait(service.submit(r))
            assert_same(r.results(), want[i], f"alone {i}")
        # all at once, three rounds (graph replays), with eager work between
        for rnd in range(3):
            ts = [service.submit(r) for r in reqs]
            for t in ts:
                service.wait(t)
            for i, r in enumerate(reqs):
                assert_same(r.results(), want[i], f"mix round {rnd} request {i}")
            if rnd == 1:
                again = gb.fit_batch(sv.frames_batch(frs[:2], maps, service.psfs, service.constraints), "fp32",
                                     service.table, streams=0)
                np.testing.assert_array_equal(again.params, np.concatenate([w[0] for w in want[:2]], axis=1))
        # replayed at a rate: arrivals 0.2 ms apart
        lat, el = service.replay(reqs, np.arange(len(reqs)) * 2e-4)
        assert np.all(lat > 0) and el > 0
        for i, r in enumerate(reqs):
            assert_same(r.results(), want[i], f"replay request {i}")
        info = service.info()
        print(f"[service] {info} replay latency p50={np.median(lat) * 1e3:.3f} ms max={lat.max() * 1e3:.3f} ms",
              flush=True)
        assert info["batches"] >= len(reqs) + 3
    finally:
        for r in reqs:
            r.close()


def test_service_rejects_bad_requests(service):
    from gaussmle_amd import _lib
    _, frs = frames()
    f = frs[0]
    # buffers that are not gmle_host_alloc memory
    import ctypes as C
    host = {k: np.ascontiguousarray(getattr(f, k)) for k in ("models", "boxes", "xc", "yc", "adu")}
    out = np.zeros(6 * f.n, np.float32)
    bad = _lib.Request(f.n, host["models"].ctypes.data, host["boxes"].ctypes.data, host["xc"].ctypes.data,
                       host["yc"].ctypes.data, host["adu"].ctypes.data, out.ctypes.data, out.ctypes.data,
                       out.ctypes.data)
    assert _lib.load().gmle_service_submit(service.handle, C.byref(bad)) == _lib.GMLE_ERR_ARG
    # an unserved shape, and a corner outside the sensor
    g = sv.Frame(f.models.copy(), f.boxes.copy(), f.xc.copy(), f.yc.copy(), f.adu.copy())
    g.models[0] = 2
    r = sv.Request(g)
    try:
        with pytest.raises(_lib.GaussMLEError):
            service.submit(r)
    finally:
        r.close()
    g = sv.Frame(f.models.copy(), f.boxes.copy(), f.xc.copy(), f.yc.copy(), f.adu.copy())
    g.xc[3] = 640
    r = sv.Request(g)
    try:
        with pytest.raises(_lib.ArgumentError):
            service.submit(r)
    finally:
        r.close()


def test_service_mixed_precision_matches_batch_path():
    maps, frs = frames()
    psfs, cons, table = sv.default_setup(lanes=2)
    s = sv.Service(psfs, cons, table, hbm_budget=1 << 30, slots=2, precision="mixed")
    s.calibrate(*maps)
    reqs = [sv.Request(f) for f in frs[:4]]
    try:
        ts = [s.submit(r) for r in reqs]
        for t in ts:
            s.wait(t)
        want = split(reference(s, frs[:4]), frs[:4])
        for i, r in enumerate(reqs):
            assert_same(r.results(), want[i], f"mixed precision request {i}")
    finally:
        for r in reqs:
            r.close()
        s.close()


@pytest.mark.parametrize("pooled", [False, True], ids=["own_buffers", "count_pool"])
@pytest.mark.parametrize("mode", ["direct", "staged", "dma", "auto"])
def test_service_ingest_modes_give_identical_results(mode, pooled, monkeypatch):
    """Raw counts read in place by the ingest kernel, packed on the host into
    the slot's staging buffer and copied by one DMA, copied by one DMA per run
    of requests whose counts are adjacent in host memory, or per run by its
    size: the pixels the fit sees are the same, so the results are too.  With
    the counts in one pool, a batch's requests form one run, and dma ingest
    moves each batch with one copy."""
    maps, frs = frames()
    monkeypatch.setenv("GMLE_SERVICE_INGEST", mode)
    psfs, cons, table = sv.default_setup(lanes=LANE_TABLES["lanes_1_2_4"])
    s = sv.Service(psfs, cons, table, hbm_budget=1 << 30, slots=2)
    pool = sv.CountPool(sum(f.adu.size for f in frs)) if pooled else None
    reqs = [sv.Request(f, pool=pool) for f in frs]
    try:
        s.calibrate(*maps)
        want = split(reference(s, frs), frs)
        for t in [s.submit(r) for r in reqs]:
            s.wait(t)
        for i, r in enumerate(reqs):
            assert_same(r.results(), want[i], f"{mode} request {i}")
        st = s.stats()
        print(f"[ingest] {mode} pooled={pooled} {st}", flush=True)
        if mode == "staged":
            assert st["batches_staged"] > 0
        elif mode != "auto":
            assert st["batches_staged"] == 0, (mode, st)
        if mode == "direct":
            assert st["dma_fraction"] == 0.0 and st["dma_copies_per_batch"] == 0.0, st
        if mode == "dma":
            assert st["dma_fraction"] == 1.0, st
            if pooled:
                assert st["dma_copies_per_batch"] == 1.0, st   # one run per batch
            else:
                assert st["dma_copies_per_batch"] == st["requests_per_batch"], st
        if mode == "auto" and not pooled:
            # every request of this sensor is far below 1 MiB of counts
            assert st["dma_fraction"] == 0.0, st
    finally:
        for r in reqs:
            r.close()
        if pool is not None:
            pool.close()
        s.close()
