"""Resident fitting service (gmle_service_* in include/gaussmle.h).

Requests are camera frames: one frame's ROIs in raw uint16 counts with their
corners on the sensor, in pinned host buffers the GPU reads and writes
directly.  The service keeps the sensor's offset / gain / variance maps on the
device, converts counts to electrons on the device and batches whatever is
pending into free slots (continuous batching).  See the header for the
semantics; this module wraps it for Python and generates seeded frames.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import batch as gb
from .models import GaussianXYNB, GaussianXYNBS, ParameterConstraints, default_constraints

F32 = np.float32


class PinnedArray:
    """A numpy array over gmle_host_alloc memory (freed by close())."""

    def __init__(self, shape, dtype):
        lib = _lib.load()
        self.dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * self.dtype.itemsize
        self.ptr = lib.gmle_host_alloc(max(n, 1))
        if not self.ptr:
            raise _lib.GaussMLEError(_lib.GMLE_ERR_NOMEM, _lib.last_error())
        buf = (C.c_uint8 * max(n, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=np.uint8, count=n).view(self.dtype).reshape(shape)

    def close(self):
        if self.ptr:
            _lib.check(_lib.load().gmle_host_free(self.ptr))
            self.ptr = None
            self.array = None


class CountPool:
    """One pinned allocation that hands out raw-count buffers back to back.

    Requests whose counts come from one pool in submission order lie next to
    each other in host memory, so the service moves a batch of them to the
    GPU with one DMA (dma ingest, or auto for runs of 1 MiB and more) instead
    of reading them in place."""

    def __init__(self, n_counts: int):
        self._buf = PinnedArray((max(int(n_counts), 1),), np.uint16)
        self._at = 0

    def take(self, n: int):
        """(view, address) of the next n counts."""
        if self._at + n > self._buf.array.size:
            raise ValueError(f"count pool exhausted: {n} more counts, {self._buf.array.size - self._at} left")
        view = self._buf.array[self._at:self._at + n]
        addr = self._buf.ptr + self._at * 2
        self._at += n
        return view, addr

    def close(self):
        self._buf.close()


@dataclass
class Frame:
    """One frame's ROIs (host numpy arrays)."""

    models: np.ndarray   # (n,) int8
    boxes: np.ndarray    # (n,) int8
    xc: np.ndarray       # (n,) int32, 1-based sensor column
    yc: np.ndarray       # (n,) int32, 1-based sensor row
    adu: np.ndarray      # (sum box^2,) uint16, ROIs back to back
    truth: np.ndarray | None = None   # (6, n)

    @property
    def n(self) -> int:
        return len(self.models)


class Request:
    """A frame copied into pinned buffers, with pinned outputs.  With a pool,
    the raw counts go into the pool's next slice (the pool owns them)."""

    def __init__(self, frame: Frame, pool: CountPool | None = None):
        n = frame.n
        self.frame = frame
        self._bufs = {}
        ptrs = {}
        for name, arr in (("models", frame.models.astype(np.int8)), ("boxes", frame.boxes.astype(np.int8)),
                          ("xc", frame.xc.astype(np.int32)), ("yc", frame.yc.astype(np.int32)),
                          ("adu", frame.adu.astype(np.uint16))):
            if name == "adu" and pool is not None:
                view, ptrs[name] = pool.take(arr.size)
                view[...] = arr
                continue
            b = PinnedArray(arr.shape, arr.dtype)
            b.array[...] = arr
            self._bufs[name] = b
        for name, shape in (("params", (n, gb.STRIDE)), ("crlb", (n, gb.STRIDE)), ("llr", (n,))):
            b = PinnedArray(shape, F32)
            b.array[...] = np.nan
            self._bufs[name] = b
        p = {k: v.ptr for k, v in self._bufs.items()}
        p.update(ptrs)
        self.struct = _lib.Request(n, p["models"], p["boxes"], p["xc"], p["yc"], p["adu"], p["params"], p["crlb"],
                                   p["llr"])

    def results(self):
        """Copies of (params (6, n), crlb (6, n), llr (n,))."""
        return (self._bufs["params"].array.T.copy(), self._bufs["crlb"].array.T.copy(),
                self._bufs["llr"].array.copy())

    def close(self):
        for b in self._bufs.values():
            b.close()


class Service:
    def __init__(self, psfs: dict, constraints: dict, lanes: dict, hbm_budget: int = 2 << 30, slots: int = 4,
                 precision: str = "fp32", iterations: int = 20, device: int = 0):
        lib = _lib.load()
        self.psfs, self.constraints, self.lanes = psfs, constraints, dict(lanes)
        self.precision, self.iterations = precision, iterations
        mb = gb.MixedBatch(np.zeros(0, np.int8), np.zeros(0, np.int8), np.zeros(0, F32), np.zeros(0, np.int64),
                           psfs, constraints, iterations)
        self._psf = mb.psf_table()
        self._lo, self._up, self._ms = mb.constraint_tables()
        self.table = gb._lanes_table(self.lanes)
        h = C.c_void_p()
        _lib.check(lib.gmle_service_create(device, hbm_budget, slots, _lib.PRECISION_IDS[precision], iterations,
                                           _lib.fptr(self._psf.reshape(-1)), _lib.fptr(self._lo.reshape(-1)),
                                           _lib.fptr(self._up.reshape(-1)), _lib.fptr(self._ms.reshape(-1)),
                                           _lib.iptr(self.table), C.byref(h)))
        self.handle = h

    def calibrate(self, offset: np.ndarray, gain: np.ndarray, var: np.ndarray):
        """Sensor maps (ny, nx) indexed [row = y, col = x]."""
        f = lambda a: np.asfortranarray(a, dtype=F32).reshape(-1, order="F")  # noqa: E731
        self.maps = (np.asarray(offset, F32), np.asarray(gain, F32), np.asarray(var, F32))
        ny, nx = np.shape(var)
        o, g, v = f(offset), f(gain), f(var)
        _lib.check(_lib.load().gmle_service_calibrate(self.handle, _lib.fptr(o), _lib.fptr(g), _lib.fptr(v), ny, nx))

    def submit(self, req: Request) -> int:
        t = _lib.load().gmle_service_submit(self.handle, C.byref(req.struct))
        if t < 0:
            _lib.check(int(t))
        return int(t)

    def wait(self, ticket: int, timeout: float = 60.0):
        rc = _lib.load().gmle_service_wait(self.handle, ticket, timeout)
        if rc == 1:
            raise TimeoutError(f"ticket {ticket} not done after {timeout}s")
        _lib.check(rc)

    def replay(self, reqs: list, arrivals: np.ndarray):
        """Submit reqs[i] at arrivals[i] s; returns (latencies (n,), elapsed s)."""
        n = len(reqs)
        arr = (_lib.Request * n)(*[r.struct for r in reqs])
        at = np.ascontiguousarray(arrivals, dtype=np.float64)
        lat = np.zeros(n, dtype=np.float64)
        el = C.c_double()
        _lib.check(_lib.load().gmle_service_replay(self.handle, n, arr, at.ctypes.data_as(_lib._DP),
                                                   lat.ctypes.data_as(_lib._DP), C.byref(el)))
        return lat, el.value

    def info(self) -> dict:
        a, b, d = C.c_int64(), C.c_int64(), C.c_int64()
        c = C.c_int()
        _lib.check(_lib.load().gmle_service_info(self.handle, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return {"slot_rois": a.value, "slot_pixels": b.value, "slots": c.value, "batches": d.value}

    def stats(self) -> dict:
        v = np.zeros(12)
        _lib.check(_lib.load().gmle_service_stats(self.handle, v.ctypes.data_as(_lib._DP), 12))
        keys = ("batches", "rois_per_batch", "requests_per_batch", "host_launch_us", "gpu_us", "launch_to_seen_us",
                "launch_to_start_us", "gpu_max_us", "launch_to_seen_max_us", "batches_staged", "dma_fraction",
                "dma_copies_per_batch")
        return {k: round(float(x), 2) for k, x in zip(keys, v)}

    def close(self):
        if self.handle:
            _lib.check(_lib.load().gmle_service_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------------------
# Seeded frames
# ---------------------------------------------------------------------------
def sensor_maps(seed: int, ny: int = 2048, nx: int = 2048):
    """Offset 100 ADU, gain ~ U(0.4, 0.6) e-/ADU, read variance ~ LogNormal(ln 2, 0.5) e-^2."""
    from .synth import scmos_maps
    return scmos_maps(np.random.default_rng(seed), ny, nx)


def make_frames(n_frames: int, seed: int, maps, boxes=(9, 11, 13, 15), models=(0, 1), rois=(50, 5000),
                chunk: int = 1 << 15):
    """Seeded frames: ROI count log-uniform in `rois`, box sizes and models
    uniform over `boxes` / `models`; x, y within 1.5 px of the ROI centre,
    N ~ U(500, 5000), bg ~ U(1, 20), sigma ~ U(1.0, 1.6) (XYNBS); raw counts
    round((Poisson(mu) + N(0, var)) / gain + offset) clipped to uint16."""
    from .synth import expected_counts
    offset, gain, var = maps
    ny, nx = var.shape
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(rois[0]), np.log(rois[1]), n_frames)).astype(np.int64)
    n = int(counts.sum())
    box_arr = np.asarray(boxes, np.int8)[rng.integers(0, len(boxes), n)]
    model_arr = np.asarray(models, np.int8)[rng.integers(0, len(models), n)]
    xc = np.zeros(n, np.int32)
    yc = np.zeros(n, np.int32)
    truth = np.full((gb.STRIDE, n), np.nan)
    psfs = {0: GaussianXYNB(1.3), 1: GaussianXYNBS()}
    sizes = box_arr.astype(np.int64) ** 2
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])   # packed in ROI order
    adu = np.empty(int(sizes.sum()), np.uint16)
    for m in models:
        for b in boxes:
            all_idx = np.flatnonzero((model_arr == m) & (box_arr == b))
            for c0 in range(0, len(all_idx), chunk):
                idx = all_idx[c0:c0 + chunk]
                k = len(idx)
                c = (b + 1) / 2
                th = [rng.uniform(c - 1.5, c + 1.5, k), rng.uniform(c - 1.5, c + 1.5, k), rng.uniform(500, 5000, k),
                      rng.uniform(1, 20, k)]
                if m == 1:
                    th.append(rng.uniform(1.0, 1.6, k))
                th = np.stack(th)
                truth[:th.shape[0], idx] = th
                mu = expected_counts(psfs[m], th, b)               # (b, b, k)
                xs = rng.integers(1, nx - b + 2, k).astype(np.int32)
                ys = rng.integers(1, ny - b + 2, k).astype(np.int32)
                xc[idx], yc[idx] = xs, ys
                ii = np.arange(b)
                flat = (((ys.astype(np.int64) - 1)[None, None, :] + ii[:, None, None]) * nx
                        + (xs.astype(np.int64) - 1)[None, None, :] + ii[None, :, None])
                v = np.take(var.ravel(), flat).astype(np.float64)
                g = np.take(gain.ravel(), flat).astype(np.float64)
                o = np.take(offset.ravel(), flat).astype(np.float64)
                e = rng.poisson(mu) + rng.standard_normal(mu.shape) * np.sqrt(v)
                a = np.clip(np.rint(e / g + o), 0, 65535).astype(np.uint16)
                dst = offs[idx][:, None] + np.arange(b * b)[None, :]
                adu[dst] = a.reshape(b * b, k, order="F").T
    frames = []
    start = 0
    for cnt in counts:
        sl = slice(start, start + int(cnt))
        p0, p1 = int(offs[start]), int(offs[start + cnt - 1] + sizes[start + cnt - 1])
        frames.append(Frame(model_arr[sl].copy(), box_arr[sl].copy(), xc[sl].copy(), yc[sl].copy(),
                            adu[p0:p1].copy(), truth[:, sl].copy()))
        start += int(cnt)
    return frames


def electrons(frame: Frame, maps) -> np.ndarray:
    """(adu - offset) * gain in Float32 per pixel, as the service computes it
    (src/camera_models.jl:110-113 with per-pixel maps)."""
    offset, gain, _ = maps
    out = np.empty(frame.adu.shape, F32)
    p = 0
    for k in range(frame.n):
        b = int(frame.boxes[k])
        ii = np.arange(b)
        rows = (frame.yc[k] - 1 + ii)[:, None] + 0 * ii[None, :]
        cols = 0 * ii[:, None] + (frame.xc[k] - 1 + ii)[None, :]
        o = offset[rows, cols].reshape(-1, order="F")
        g = gain[rows, cols].reshape(-1, order="F")
        a = frame.adu[p:p + b * b].astype(F32)
        out[p:p + b * b] = (a - o) * g
        p += b * b
    return out


def frames_batch(frames: list, maps, psfs: dict, constraints: dict, iterations: int = 20) -> gb.MixedBatch:
    """The frames' ROIs as one MixedBatch in electrons (for the plain batch path)."""
    models = np.concatenate([f.models for f in frames]).astype(np.int8)
    boxes = np.concatenate([f.boxes for f in frames]).astype(np.int8)
    data = np.concatenate([electrons(f, maps) for f in frames]).astype(F32)
    sizes = boxes.astype(np.int64) ** 2
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    return gb.MixedBatch(models, boxes, data, offsets, psfs, constraints, iterations, maps[2],
                         np.concatenate([f.xc for f in frames]).astype(np.int32),
                         np.concatenate([f.yc for f in frames]).astype(np.int32),
                         np.concatenate([f.truth for f in frames], axis=1))


def default_setup(lanes: int | dict = 2):
    """PSFs, constraints and a lanes table for frames of XYNB / XYNBS ROIs of
    9-15 px (every served shape gets `lanes` lanes per ROI unless a dict)."""
    psfs = {0: GaussianXYNB(1.3), 1: GaussianXYNBS()}
    cons = {m: default_constraints(p, 11) for m, p in psfs.items()}
    if isinstance(lanes, dict):
        table = dict(lanes)
    else:
        table = {(m, b): lanes for m in (0, 1) for b in (9, 11, 13, 15)}
    return psfs, cons, table
