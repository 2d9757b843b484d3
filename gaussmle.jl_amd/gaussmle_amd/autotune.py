"""Autotuner for mixed-shape batches.

Each (model, box) shape has kernels for 1, 2, 4, 8 or 16 lanes per ROI (64 down
to 4 ROIs per wavefront; a lane holds a block of ceil(box / L) pixel columns of its
ROI in VGPRs, which is the kernel's tile).  The best choice depends on the
shape's register footprint and on how many ROIs of it a launch has, and in a
batch also on what runs beside it.  The tuner:
  1. times every supported L of each shape alone, on that shape's ROIs of the
     batch (plain launches, device-resident);
  2. refines in the batch itself: coordinate descent over the shapes, timing
     the whole batched call, keeping a change only when it is faster;
  3. picks the number of streams the per-shape launches are spread over, or
     the one-launch mode (every shape in one multi-shape kernel, 4 lanes).
All timings are HIP-event times on the device, median of `reps` runs.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _lib
from . import batch as gb


def _median_ms(fn, reps: int, warmup: int = 1) -> float:
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ts = []
    s = torch.cuda.current_stream()
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        fn()
        b.record(s)
        b.synchronize()
        ts.append(a.elapsed_time(b))
    return float(np.median(ts))


def shape_stacks(mb: gb.MixedBatch, d: dict) -> dict:
    """Per shape: contiguous device copies of its ROIs (data and corners), as a
    plain per-shape launch would take them."""
    import torch
    out = {}
    counts = mb.shape_counts()
    for s in np.flatnonzero(counts):
        m, b = gb.shape_of(int(s))
        idx = mb.roi_indices(m, b)
        ti = torch.from_numpy(idx).to(d["data"].device)
        offs = d["offsets"][ti]
        pix = (offs[:, None] + torch.arange(b * b, device=offs.device)[None, :]).reshape(-1)
        e = {"idx": idx, "data": d["data"][pix].contiguous()}
        if mb.camera:
            e["xc"] = d["xc"][ti].contiguous()
            e["yc"] = d["yc"][ti].contiguous()
        out[(m, b)] = e
    return out


def plain_launch(mb: gb.MixedBatch, d: dict, stacks: dict, key, lanes: int, precision: str, outs: dict):
    from .device import fit_device
    m, b = key
    e = stacks[key]
    ny, nx = (mb.var_map.shape if mb.camera else (0, 0))
    outs[key] = fit_device(mb.psfs[m], b, e["data"], mb.constraints[m], mb.iterations, precision, lanes,
                           d.get("var"), ny, nx, e.get("xc"), e.get("yc"), out=outs.get(key))


@dataclass
class Tuning:
    lanes: dict                      # (model, box) -> L
    streams: int
    alone_ms: dict = field(default_factory=dict)     # (model, box) -> {L: ms}
    batch_ms: float = 0.0
    one_launch_ms: float = 0.0
    history: list = field(default_factory=list)

    def table(self) -> np.ndarray:
        t = np.zeros(gb.NUM_SHAPES, dtype=np.int32)
        for (m, b), L in self.lanes.items():
            t[gb.shape_index(m, b)] = L
        return t

    def to_json(self) -> dict:
        return {"lanes": {f"{m}:{b}": L for (m, b), L in self.lanes.items()}, "streams": self.streams,
                "batch_ms": self.batch_ms, "one_launch_ms": self.one_launch_ms,
                "alone_ms": {f"{m}:{b}": {str(L): t for L, t in v.items()} for (m, b), v in self.alone_ms.items()}}

    def save(self, path):
        Path(path).write_text(json.dumps(self.to_json(), indent=1))


def tune(mb: gb.MixedBatch, d: dict, precision: str = "fp32", reps: int = 3, stacks: dict | None = None,
         refine: bool = True, stream_choices=(1, 2, 4)) -> Tuning:
    lib = _lib.load()
    stacks = stacks or shape_stacks(mb, d)
    outs = {}
    alone = {}
    best = {}
    for key in stacks:
        m, b = key
        ok = [L for L in (1, 2, 4, 8, 16) if lib.gmle_lanes_supported(b, L)]
        alone[key] = {L: _median_ms(lambda: plain_launch(mb, d, stacks, key, L, precision, outs), reps)
                      for L in ok}
        best[key] = min(alone[key], key=alone[key].get)
    tuning = Tuning(dict(best), max(stream_choices), alone)

    out = {}

    def batch_ms(lanes, streams):
        t = np.zeros(gb.NUM_SHAPES, dtype=np.int32)
        for (m, b), L in lanes.items():
            t[gb.shape_index(m, b)] = L
        return _median_ms(lambda: out.update(gb.fit_batch_device(mb, d, precision, t, streams, out=out or None)),
                          reps)

    cur = batch_ms(tuning.lanes, tuning.streams)
    tuning.history.append(("start", cur))
    if refine:
        # largest shapes first: they decide most of the time
        order = sorted(stacks, key=lambda k: -len(stacks[k]["idx"]) * k[1] ** 2)
        for key in order:
            for L in alone[key]:
                if L == tuning.lanes[key]:
                    continue
                trial = dict(tuning.lanes)
                trial[key] = L
                t = batch_ms(trial, tuning.streams)
                tuning.history.append((f"{key}:L{L}", t))
                if t < cur * 0.99:
                    tuning.lanes, cur = trial, t
    for s in stream_choices:
        if s == tuning.streams:
            continue
        t = batch_ms(tuning.lanes, s)
        tuning.history.append((f"streams{s}", t))
        if t < cur * 0.99:
            tuning.streams, cur = s, t
    if precision == "fp32":
        # one launch for every shape (4 lanes per ROI; the lanes table is unused)
        t = batch_ms(tuning.lanes, 0)
        tuning.history.append(("one_launch", t))
        tuning.one_launch_ms = t
        if t < cur * 0.99:
            tuning.streams, cur = 0, t
    tuning.batch_ms = cur
    return tuning
