"""Mixed-shape batches: ROIs of different box sizes and PSF models in one call.

GaussMLE.jl fits one box size and one model per `fit` call
(src/interface.jl:177-291; the kernel reads a single box size,
src/unified_kernel.jl:301), so a mixed set of ROIs takes one call per
(model, box).  `fit_batch` / `fit_batch_device` hand the whole set to the HIP
library (gmle_fit_batch / gmle_fit_batch_device), which groups the ROIs by
shape on the device and runs every shape's own kernel, spread over several
HIP streams.

Layout of a batch (`MixedBatch`): ROI k has model models[k] and box size
boxes[k]; its box*box Float32 pixels (column-major, the reference's per-ROI
layout) start at data[offsets[k]]; results come back as (6, n) params / crlb
with each ROI's P values followed by NaN, and llr (n,).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .models import (AstigmaticXYZNB, GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, ParameterConstraints,
                     PSFModel, default_constraints)

F32 = np.float32
NUM_SHAPES = _lib.NUM_SHAPES
STRIDE = _lib.BATCH_STRIDE
MIN_BOX, MAX_BOX = 3, 16
MODEL_CLASSES = (GaussianXYNB, GaussianXYNBS, GaussianXYNBSXSY, AstigmaticXYZNB)


def shape_index(model: int, box: int) -> int:
    """Index of (model, box) in the per-shape tables (gmle_shape_index)."""
    if not (0 <= model <= 3 and MIN_BOX <= box <= MAX_BOX):
        return -1
    return model * (MAX_BOX - MIN_BOX + 1) + (box - MIN_BOX)


def shape_of(index: int) -> tuple[int, int]:
    nb = MAX_BOX - MIN_BOX + 1
    return index // nb, MIN_BOX + index % nb


@dataclass
class MixedBatch:
    """A concatenated batch of ROIs of mixed box sizes and models."""

    models: np.ndarray          # (n,) int8
    boxes: np.ndarray           # (n,) int8
    data: np.ndarray            # (sum box^2,) float32
    offsets: np.ndarray         # (n,) int64
    psfs: dict                  # model id -> PSFModel in pixel units
    constraints: dict           # model id -> ParameterConstraints
    iterations: int = 20
    var_map: np.ndarray | None = None     # (ny, nx) float32, sCMOS only
    x_corners: np.ndarray | None = None
    y_corners: np.ndarray | None = None
    truth: np.ndarray | None = None       # (6, n) float64, NaN-padded (synthetic batches)
    name: str = "mixed"

    @property
    def n(self) -> int:
        return len(self.models)

    @property
    def camera(self) -> int:
        return 0 if self.var_map is None else 1

    def shape_counts(self) -> np.ndarray:
        keys = self.models.astype(np.int64) * (MAX_BOX - MIN_BOX + 1) + (self.boxes.astype(np.int64) - MIN_BOX)
        return np.bincount(keys, minlength=NUM_SHAPES).astype(np.int64)

    def psf_table(self) -> np.ndarray:
        t = np.zeros((4, 8), dtype=F32)
        for m, p in self.psfs.items():
            t[m] = p.psf_array()
        return t

    def constraint_tables(self):
        lo = np.zeros((4, 6), dtype=F32)
        up = np.zeros((4, 6), dtype=F32)
        ms = np.zeros((4, 6), dtype=F32)
        for m, c in self.constraints.items():
            P = len(c)
            lo[m, :P], up[m, :P], ms[m, :P] = c.lower, c.upper, c.max_step
        return lo, up, ms

    def roi_indices(self, model: int, box: int) -> np.ndarray:
        return np.flatnonzero((self.models == model) & (self.boxes == box))

    def roi_stack(self, idx: np.ndarray, box: int) -> np.ndarray:
        """(box, box, len(idx)) Fortran-ordered stack of the given ROIs (one shape)."""
        gather = self.offsets[idx][:, None] + np.arange(box * box)[None, :]
        flat = self.data[gather]                       # (k, box*box)
        # flat[k, i + box*j] is pixel (row i, column j) of ROI k
        return np.asfortranarray(flat.reshape(len(idx), box, box).transpose(2, 1, 0))

    def subset(self, idx: np.ndarray) -> "MixedBatch":
        """The ROIs idx, repacked in that order."""
        idx = np.asarray(idx, dtype=np.int64)
        boxes = self.boxes[idx]
        sizes = boxes.astype(np.int64) ** 2
        offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        # pixel q of the repacked buffer comes from ROI r(q) at position q - offsets_new[r(q)]
        owner = np.repeat(np.arange(len(idx)), sizes)
        gather = self.offsets[idx][owner] + (np.arange(int(sizes.sum())) - offsets[owner])
        return MixedBatch(self.models[idx].copy(), boxes.copy(), self.data[gather].astype(F32), offsets, self.psfs,
                          self.constraints, self.iterations, self.var_map,
                          None if self.x_corners is None else self.x_corners[idx].copy(),
                          None if self.y_corners is None else self.y_corners[idx].copy(),
                          None if self.truth is None else self.truth[:, idx].copy(), self.name)


@dataclass
class BatchResult:
    params: np.ndarray     # (6, n) float32
    crlb: np.ndarray       # (6, n) float32
    llr: np.ndarray        # (n,) float32
    crlb64: np.ndarray | None = None
    llr64: np.ndarray | None = None


def _lanes_table(lanes) -> np.ndarray | None:
    """lanes: None, an int (every shape), a dict {(model, box): L} or a table."""
    if lanes is None:
        return None
    if isinstance(lanes, (int, np.integer)):
        return np.full(NUM_SHAPES, int(lanes), dtype=np.int32)
    if isinstance(lanes, dict):
        t = np.zeros(NUM_SHAPES, dtype=np.int32)
        for (m, b), L in lanes.items():
            t[shape_index(m, b)] = L
        return t
    t = np.ascontiguousarray(lanes, dtype=np.int32)
    assert t.shape == (NUM_SHAPES,)
    return t


def fit_batch(batch: MixedBatch, precision: str = "fp32", lanes=None, streams: int = 4) -> BatchResult:
    """gmle_fit_batch on host arrays (synchronous).  That entry point takes
    the ROIs packed back to back in ROI order; other layouts are repacked."""
    lib = _lib.load()
    sizes = batch.boxes.astype(np.int64) ** 2
    if not np.array_equal(batch.offsets, np.concatenate([[0], np.cumsum(sizes)[:-1]])):
        batch = batch.subset(np.arange(batch.n))
    n = batch.n
    params = np.empty(n * STRIDE, dtype=F32)
    crlb = np.empty(n * STRIDE, dtype=F32)
    llr = np.empty(n, dtype=F32)
    mixed = precision == "mixed"
    crlb64 = np.empty(n * STRIDE, dtype=np.float64) if mixed else None
    llr64 = np.empty(n, dtype=np.float64) if mixed else None
    lo, up, ms = batch.constraint_tables()
    psf = batch.psf_table()
    vflat = None
    ny = nx = 0
    xc = yc = None
    if batch.camera:
        v = np.asfortranarray(batch.var_map, dtype=F32)
        ny, nx = v.shape
        vflat = v.reshape(-1, order="F")
        xc = np.ascontiguousarray(batch.x_corners, dtype=np.int32)
        yc = np.ascontiguousarray(batch.y_corners, dtype=np.int32)
    data = np.ascontiguousarray(batch.data, dtype=F32)
    rc = lib.gmle_fit_batch(n, _lib.i8ptr(np.ascontiguousarray(batch.models, np.int8)),
                            _lib.i8ptr(np.ascontiguousarray(batch.boxes, np.int8)), _lib.fptr(data),
                            _lib.fptr(psf.reshape(-1)), _lib.fptr(lo.reshape(-1)), _lib.fptr(up.reshape(-1)),
                            _lib.fptr(ms.reshape(-1)), batch.camera, _lib.fptr(vflat), ny, nx, _lib.iptr(xc),
                            _lib.iptr(yc), batch.iterations, _lib.PRECISION_IDS[precision],
                            _lib.iptr(_lanes_table(lanes)), streams, _lib.fptr(params), _lib.fptr(crlb),
                            _lib.fptr(llr), _lib.dptr(crlb64), _lib.dptr(llr64))
    _lib.check(rc)
    shape = (STRIDE, n)
    return BatchResult(params.reshape(shape, order="F"), crlb.reshape(shape, order="F"), llr,
                       None if crlb64 is None else crlb64.reshape(shape, order="F"), llr64)


def upload_batch(batch: MixedBatch, device="cuda:0") -> dict:
    """Device copies (torch tensors) of a batch's arrays, plus its workspace."""
    import torch

    lib = _lib.load()
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    d = {"models": t(batch.models.astype(np.int8)), "boxes": t(batch.boxes.astype(np.int8)),
         "offsets": t(batch.offsets.astype(np.int64)), "data": t(batch.data.astype(F32)),
         "counts": batch.shape_counts()}
    if batch.camera:
        d["var"] = t(np.asfortranarray(batch.var_map, dtype=F32).reshape(-1, order="F"))
        d["xc"] = t(batch.x_corners.astype(np.int32))
        d["yc"] = t(batch.y_corners.astype(np.int32))
    ws = int(lib.gmle_batch_workspace_bytes(batch.n))
    d["workspace"] = torch.empty(ws, dtype=torch.uint8, device=device)
    return d


def fit_batch_device(batch: MixedBatch, d: dict, precision: str = "fp32", lanes=None, streams: int = 4,
                     out=None, stream=None, host_counts: bool = True) -> dict:
    """gmle_fit_batch_device on tensors from upload_batch, enqueued on torch's
    current stream (or `stream`)."""
    import torch

    lib = _lib.load()
    n = batch.n
    dev = d["data"].device
    if out is None:
        out = {"params": torch.empty((n, STRIDE), dtype=torch.float32, device=dev),
               "crlb": torch.empty((n, STRIDE), dtype=torch.float32, device=dev),
               "llr": torch.empty((n,), dtype=torch.float32, device=dev)}
        if precision == "mixed":
            out["crlb64"] = torch.empty((n, STRIDE), dtype=torch.float64, device=dev)
            out["llr64"] = torch.empty((n,), dtype=torch.float64, device=dev)
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    lo, up, ms = batch.constraint_tables()
    psf = batch.psf_table()
    ny, nx = (batch.var_map.shape if batch.camera else (0, 0))

    def ptr(x):
        return None if x is None else x.data_ptr()

    counts = np.ascontiguousarray(d["counts"], dtype=np.int64) if host_counts else None
    rc = lib.gmle_fit_batch_device(
        n, ptr(d["models"]), ptr(d["boxes"]), ptr(d["offsets"]), ptr(d["data"]), _lib.fptr(psf.reshape(-1)),
        _lib.fptr(lo.reshape(-1)), _lib.fptr(up.reshape(-1)), _lib.fptr(ms.reshape(-1)), batch.camera,
        ptr(d.get("var")), ny, nx, ptr(d.get("xc")), ptr(d.get("yc")), batch.iterations,
        _lib.PRECISION_IDS[precision], _lib.iptr(_lanes_table(lanes)), streams, _lib.i64ptr(counts),
        ptr(out["params"]), ptr(out["crlb"]), ptr(out["llr"]), ptr(out.get("crlb64")), ptr(out.get("llr64")),
        ptr(d["workspace"]), d["workspace"].numel(), s.cuda_stream)
    _lib.check(rc)
    return out


def fit_batch_device_noalloc(batch: MixedBatch, d: dict, precision: str = "fp32", lanes=None, streams: int = 1,
                             out: dict | None = None, stream=None, host_counts: bool = True) -> dict:
    """gmle_fit_batch_device_noalloc: fit_batch_device with the library's
    static device workspace instead of d["workspace"]; `out` must be given
    (the call itself allocates nothing).  streams 0 or 1."""
    import torch

    lib = _lib.load()
    s = stream if stream is not None else torch.cuda.current_stream(d["data"].device)
    lo, up, ms = batch.constraint_tables()
    psf = batch.psf_table()
    ny, nx = (batch.var_map.shape if batch.camera else (0, 0))

    def ptr(x):
        return None if x is None else x.data_ptr()

    counts = np.ascontiguousarray(d["counts"], dtype=np.int64) if host_counts else None
    rc = lib.gmle_fit_batch_device_noalloc(
        batch.n, ptr(d["models"]), ptr(d["boxes"]), ptr(d["offsets"]), ptr(d["data"]), _lib.fptr(psf.reshape(-1)),
        _lib.fptr(lo.reshape(-1)), _lib.fptr(up.reshape(-1)), _lib.fptr(ms.reshape(-1)), batch.camera,
        ptr(d.get("var")), ny, nx, ptr(d.get("xc")), ptr(d.get("yc")), batch.iterations,
        _lib.PRECISION_IDS[precision], _lib.iptr(_lanes_table(lanes)), streams, _lib.i64ptr(counts),
        ptr(out["params"]), ptr(out["crlb"]), ptr(out["llr"]), ptr(out.get("crlb64")), ptr(out.get("llr64")),
        s.cuda_stream)
    _lib.check(rc)
    return out


# ---------------------------------------------------------------------------
# Synthetic mixed batches
# ---------------------------------------------------------------------------
def mixed_workload(n: int, seed: int, boxes=(7, 9, 11, 13, 15), box_weights=(0.1, 0.2, 0.4, 0.2, 0.1),
                   models=(0, 1, 2), model_weights=None, scmos: bool = True, cam: int = 256,
                   sigma_range=(1.0, 1.8), iterations: int = 20, chunk: int = 1 << 16) -> MixedBatch:
    """A seeded mixed batch.  Default: box sizes {7, 9, 11, 13, 15} with
    weights {.1, .2, .4, .2, .1}; models uniform over XYNB (sigma 1.3 px),
    XYNBS and XYNBSXSY (widths ~ U(1.0, 1.8) px); x, y within 1.5 px of the
    ROI centre, N ~ U(500, 5000), bg ~ U(1, 20) (astigmatic ROIs: z ~
    U(-300, 300), N ~ U(1000, 10000), bg ~ U(2, 30)); sCMOS on a 256x256
    sensor (offset 100 ADU, gain ~ U(0.4, 0.6), read variance
    LogNormal(ln 2, 0.5)) with random corners; 20 iterations.  Default
    constraints use box 11 for every ROI, as the reference's fitter does
    (src/interface.jl:141).

    ROIs are interleaved in random shape order; the pixel buffer holds them
    grouped by shape (offsets[k] locates ROI k), which is how a caller that
    cuts ROIs shape by shape would fill it."""
    from .synth import ASTIG_CONFIG_PSF, astig_constraints, expected_counts, scmos_maps, scmos_stack

    rng = np.random.default_rng(seed)
    bsel = rng.choice(len(boxes), size=n, p=np.asarray(box_weights) / np.sum(box_weights))
    mw = np.ones(len(models)) if model_weights is None else np.asarray(model_weights, dtype=np.float64)
    msel = rng.choice(len(models), size=n, p=mw / mw.sum())
    box_arr = np.asarray(boxes, dtype=np.int8)[bsel]
    model_arr = np.asarray(models, dtype=np.int8)[msel]
    psfs = {}
    cons = {}
    for m in models:
        psfs[m] = GaussianXYNB(1.3) if m == 0 else ASTIG_CONFIG_PSF if m == 3 else MODEL_CLASSES[m]()
        cons[m] = astig_constraints() if m == 3 else default_constraints(psfs[m], 11)
    if scmos:
        offset, gain, var = scmos_maps(rng, cam, cam)
        xc = np.zeros(n, dtype=np.int32)
        yc = np.zeros(n, dtype=np.int32)
    offsets = np.zeros(n, dtype=np.int64)
    data = np.empty(int((box_arr.astype(np.int64) ** 2).sum()), dtype=F32)
    truth = np.full((STRIDE, n), np.nan)
    pos = 0
    for m in models:
        for b in boxes:
            all_idx = np.flatnonzero((model_arr == m) & (box_arr == b))
            for c0 in range(0, len(all_idx), chunk):
                idx = all_idx[c0:c0 + chunk]
                k = len(idx)
                c = (b + 1) / 2
                th = [rng.uniform(c - 1.5, c + 1.5, k), rng.uniform(c - 1.5, c + 1.5, k)]
                if m == 3:
                    th += [rng.uniform(-300, 300, k), rng.uniform(1000, 10000, k), rng.uniform(2, 30, k)]
                else:
                    th += [rng.uniform(500, 5000, k), rng.uniform(1, 20, k)]
                if m == 1:
                    th.append(rng.uniform(*sigma_range, k))
                elif m == 2:
                    th += [rng.uniform(*sigma_range, k), rng.uniform(*sigma_range, k)]
                th = np.stack(th)
                truth[:th.shape[0], idx] = th
                mu = expected_counts(psfs[m], th, b)
                if scmos:
                    xs = rng.integers(1, cam - b + 2, k).astype(np.int32)
                    ys = rng.integers(1, cam - b + 2, k).astype(np.int32)
                    xc[idx], yc[idx] = xs, ys
                    stack = scmos_stack(rng, mu, var, gain, offset, xs, ys)
                else:
                    stack = np.empty(mu.shape, dtype=F32, order="F")
                    stack[...] = rng.poisson(mu)
                data[pos:pos + k * b * b] = stack.reshape(-1, order="F")   # ROI after ROI
                offsets[idx] = pos + np.arange(k, dtype=np.int64) * (b * b)
                pos += k * b * b
    return MixedBatch(model_arr, box_arr, data, offsets, psfs, cons, iterations,
                      var if scmos else None, xc if scmos else None, yc if scmos else None, truth,
                      f"mixed_{'scmos' if scmos else 'ideal'}_{n}")


def config3(n: int = 4_000_000, seed: int = 3) -> MixedBatch:
    """Uneven batch: box sizes {7, 9, 11, 13, 15} (10/20/40/20/10 %), models
    uniform over {XYNB, XYNBS, XYNBSXSY}, widths ~ U(1.0, 1.8) px, sCMOS as the
    13x13 XYNBS workload, 20 iterations."""
    return mixed_workload(n, seed)
