#!/usr/bin/env python3
"""Throughput benchmark of the MI355X per-ROI MLE fit (one GPU).

A step is one launch of the fitting kernel over one batch of device-resident,
seeded synthetic ROIs.  The default workload is the headline one: 11x11
GaussianXYNB (sigma 1.3 px), Poisson camera, 20 Newton iterations, with the
parameter draws x, y = centre +- 1.5 px, N ~ U(500, 5000), bg ~ U(1, 20); the
batch is 2^20 ROIs.  Prints one JSON line with fits/s, the kernel's roofline
position (FP32 VALU, algorithmic FLOPs per fit as flops_per_fit counts them)
and a CPU baseline (the oracle, timed on this host's cores on a bounded
sample).  Beside the headline the same line reports, each with its own
roofline: the 13x13 sCMOS XYNBS and 15x15 astigmatic configurations at full
size in both precisions, the uneven batch of 4,000,000 ROIs, the host-array
entry point on the headline workload, the 10,000-ROI launch, the result
epilogue, the streaming service, ROI extraction and the simulator.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload xynb11|config1|config2|mixed|...]
    python bench.py --trace-summary kernel_trace.csv     # per-kernel durations of a rocprofv3 trace
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "gaussmle.jl_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "ROI fits/sec (11×11 XYNB, 20 iters) plain vs batched; p99 latency under load"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# GPU fits/s quoted by the reference for 11x11 GaussianXYNB (RTX 4090, batch 50K).
REFERENCE_GPU_FITS_PER_S = 1.0e7


def flops_per_fit(model: str, n: int, iters: int, camera: str) -> int:
    """Separable-algorithm FP32 operation count per fit (add/mul = 1, FMA = 2;
    transcendentals not counted): I (n^2 c_pix + 2 n c_axis) + n^2 (c_fin + c_init)."""
    c_pix = {"xynb": 32, "xynbs": 45, "xynbsxsy": 50, "astig": 45}[model] - (1 if camera == "scmos" else 0)
    c_axis = 13 if model == "xynb" else 23
    P = {"xynb": 4, "xynbs": 5, "xynbsxsy": 6, "astig": 5}[model]
    c_fin = {4: 33, 5: 45, 6: 59}[P]
    c_init = 12 if model == "astig" else 6
    return iters * (n * n * c_pix + 2 * n * c_axis) + n * n * (c_fin + c_init)


def bytes_per_fit(n: int, P: int) -> int:
    """data + corners + params/CRLB + LLR."""
    return 4 * n * n + 8 + 8 * P + 8


WORKLOADS = {
    "xynb11": dict(label="GaussianXYNB 11x11, Poisson, 20 iterations", model="xynb", box=11, iters=20,
                   camera="ideal", rois=1 << 20, gen="config0"),
    "xynbs13": dict(label="GaussianXYNBS 13x13, sCMOS 256x256 maps, 20 iterations", model="xynbs", box=13,
                    iters=20, camera="scmos", rois=1 << 20, gen="config1"),
    "astig15": dict(label="AstigmaticXYZNB 15x15, Poisson, 30 iterations", model="astig", box=15, iters=30,
                    camera="ideal", rois=1 << 20, gen="config2"),
}


MIXED_LABEL = ("uneven batch: box sizes {7,9,11,13,15} (10/20/40/20/10%), models XYNB/XYNBS/XYNBSXSY uniform, "
               "sCMOS 256x256 maps, 20 iterations")
MODEL_KEYS = {0: "xynb", 1: "xynbs", 2: "xynbsxsy", 3: "astig"}


def mixed_flops(mb) -> float:
    """Total algorithmic FLOPs of one pass over a mixed batch."""
    from gaussmle_amd import batch as gb
    cam = "scmos" if mb.camera else "ideal"
    tot = 0.0
    for s, c in enumerate(mb.shape_counts()):
        if c:
            m, b = gb.shape_of(s)
            tot += c * flops_per_fit(MODEL_KEYS[m], b, mb.iterations, cam)
    return tot


def mixed_bytes(mb) -> float:
    P = np.array([4, 5, 6, 5])[mb.models.astype(np.int64)]
    b = mb.boxes.astype(np.int64)
    return float(np.sum(4 * b * b + 8 + 8 * P + 8))


def run_mixed(n: int, steps: int, warmup: int, precision: str, tune_reps: int = 3, seed: int = 3,
              batched_only: bool = False):
    """Config-3 style uneven batch: autotune, then time the batched call
    against per-shape plain launches (each shape's ROIs contiguous, its best
    lanes per ROI, launched one after another on one stream).  The batch is
    drawn on the device (synth_device.config3_device)."""
    import torch
    from gaussmle_amd import autotune
    from gaussmle_amd import batch as gb
    from gaussmle_amd import synth_device

    t0 = time.perf_counter()
    mb, d = synth_device.config3_device(n=n, seed=seed)
    torch.cuda.synchronize()
    print(f"[bench] generated mixed batch of {n} ROIs on the device in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    stacks = autotune.shape_stacks(mb, d)
    t0 = time.perf_counter()
    tuning = autotune.tune(mb, d, precision, reps=tune_reps, stacks=stacks)
    print(f"[bench] autotuned in {time.perf_counter() - t0:.1f}s: streams={tuning.streams} "
          f"lanes={ {f'{k[0]}:{k[1]}': v for k, v in tuning.lanes.items()} }", file=sys.stderr, flush=True)
    table = tuning.table()
    out = {}
    plain_outs = {}
    best_alone = {k: min(v, key=v.get) for k, v in tuning.alone_ms.items()}

    def batched(streams):
        out.update(gb.fit_batch_device(mb, d, precision, table, streams, out=out or None))

    def plain():
        for key in stacks:
            autotune.plain_launch(mb, d, stacks, key, best_alone[key], precision, plain_outs)

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        s = torch.cuda.current_stream()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t = time.perf_counter()
        for a, b in ev:
            a.record(s)
            fn()
            b.record(s)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        return wall, float(np.mean([a.elapsed_time(b) for a, b in ev])) / 1e3

    wall_b, dev_b = timed(lambda: batched(tuning.streams))
    rate = lambda wall: None if wall is None else round(steps * n / wall, 1)  # noqa: E731
    wall_1 = wall_0 = wall_p = dev_p = None
    if not batched_only:
        wall_1, _ = timed(lambda: batched(1))
        wall_0, _ = timed(lambda: batched(0))
        wall_p, dev_p = timed(plain)
    res = {
        "workload": MIXED_LABEL, "rois": n, "precision": precision,
        "batched_fits_per_s": rate(wall_b),
        "batched_one_stream_fits_per_s": rate(wall_1),
        "one_launch_fits_per_s": rate(wall_0),
        "per_shape_plain_fits_per_s": rate(wall_p),
        "batched_over_plain": None if wall_p is None else round(wall_p / wall_b, 4),
        "batched_ms": round(dev_b * 1e3, 4), "plain_ms": None if dev_p is None else round(dev_p * 1e3, 4),
        "streams": tuning.streams,
        "lanes": {f"{k[0]}:{k[1]}": v for k, v in tuning.lanes.items()},
        "plain_lanes": {f"{k[0]}:{k[1]}": v for k, v in best_alone.items()},
        "shape_counts": {f"{gb.shape_of(int(s))[0]}:{gb.shape_of(int(s))[1]}": int(c)
                         for s, c in enumerate(mb.shape_counts()) if c},
        "flops": mixed_flops(mb), "bytes": mixed_bytes(mb),
    }
    achieved = res["flops"] / dev_b / 1e12
    res["roofline"] = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                       "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                       "kernel_ms": round(dev_b * 1e3, 4), "flops_per_fit": round(res["flops"] / n, 1),
                       "note": "whole batched call (grouping kernels + every shape's fit kernel), HIP events"}
    return mb, d, res, dev_b


# Configurations reported beside the headline, drawn on the device.
CONFIGS = {
    "config1": dict(label="GaussianXYNBS 13x13, sCMOS 256x256 maps, 20 iterations", model="xynbs", box=13,
                    iters=20, camera="scmos", rois=1_000_000, gen="config1_device"),
    "config2": dict(label="AstigmaticXYZNB 15x15, Poisson, 30 iterations", model="astig", box=15, iters=30,
                    camera="ideal", rois=4_000_000, gen="config2_device"),
    # boxes above 16 px (8 lanes per ROI)
    "box21": dict(label="GaussianXYNBS 21x21 (sigma 1.5-2.5 px), Poisson, 20 iterations", model="xynbs", box=21,
                  iters=20, camera="ideal", rois=1_000_000, gen="large_box_device"),
}


def time_launches(fn, steps: int, warmup: int):
    """Wall time of `steps` calls (synchronised on both sides) and the mean
    HIP-event time per call on torch's current stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t = time.perf_counter()
    for a, b in ev:
        a.record(s)
        fn()
        b.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    return wall, float(np.mean([a.elapsed_time(b) for a, b in ev])) / 1e3


def run_config(key: str, steps: int, warmup: int, n: int | None = None, lanes_override: int = 0) -> dict:
    """One of CONFIGS at its full size, in Float32 (one fit kernel) and in
    mixed precision (fit kernel + Float64 final pass), with the roofline
    position of each (algorithmic FLOPs as for the headline)."""
    import torch
    from gaussmle_amd import _lib
    from gaussmle_amd import device as gd
    from gaussmle_amd import synth_device
    spec = CONFIGS[key]
    t0 = time.perf_counter()
    dw = getattr(synth_device, spec["gen"])(n=n or spec["rois"])
    torch.cuda.synchronize()
    print(f"[bench] {key}: generated {dw.n} ROIs ({spec['label']}) on the device in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    lib = _lib.load()
    ny, nx = (0, 0) if dw.var_map is None else dw.var_map.shape
    F = flops_per_fit(spec["model"], dw.box, dw.iterations, spec["camera"])
    res = {"workload": spec["label"], "rois": dw.n, "data": "synthetic: seeded parameters (numpy PCG64), pixels "
           "drawn by the on-device simulator (Philox4x32-10), device-resident before timing"}
    for precision in ("fp32", "mixed"):
        lanes = lanes_override or lib.gmle_default_lanes(dw.psf_pixels.model_id, 0 if dw.camera == "ideal" else 1,
                                                         dw.box, _lib.PRECISION_IDS[precision], dw.n)
        out = {}

        def step():
            out.update(gd.fit_device(dw.psf_pixels, dw.box, dw.d["data"], dw.constraints, dw.iterations, precision,
                                     lanes, dw.d.get("var"), ny, nx, dw.d.get("xc"), dw.d.get("yc"),
                                     out=out or None))

        wall, dev_s = time_launches(step, steps, warmup)
        achieved = F * dw.n / dev_s / 1e12
        res[precision] = {
            "fits_per_s": round(steps * dw.n / wall, 1), "ms_per_launch": round(dev_s * 1e3, 4), "lanes": lanes,
            "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                         "kernel_ms": round(dev_s * 1e3, 4), "flops_per_fit": F,
                         "kernels": "fit_kernel" if precision == "fp32" else "fit_kernel + final_pass_kernel"}}
        print(f"[bench] {key} {precision}: {res[precision]}", file=sys.stderr, flush=True)
    del dw
    torch.cuda.empty_cache()
    return res


def pcie_h2d_gbs(nbytes: int = 256 << 20, reps: int = 5) -> float:
    """Host-to-device rate of one large copy from pinned memory: the ceiling
    of any path that moves host arrays to the GPU."""
    import torch
    src = torch.empty(nbytes, dtype=torch.uint8).pin_memory()
    dst = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0")
    dst.copy_(src, non_blocking=True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src, non_blocking=True)
    torch.cuda.synchronize()
    rate = reps * nbytes / (time.perf_counter() - t) / 1e9
    del src, dst
    return rate


def run_host_path(w, reps: int = 5) -> dict:
    """The headline workload through gmle_fit on host arrays (the entry point
    GaussMLE.jl's ccall uses): pageable input through the pinned upload
    pipeline, fit, results back into the caller's arrays, per call wall time."""
    import gaussmle_amd as g
    from gaussmle_amd import _lib
    f = g.GaussMLEFitter.__new__(g.GaussMLEFitter)
    f.psf_model, f.iterations, f.constraints = w.psf_pixels, w.iterations, w.constraints
    f.precision, f.lanes, f.batch_size, f.device = "fp32", 0, 10_000, g.GPU()
    data = np.asfortranarray(w.data)
    g.fit_arrays(f, data, w.psf_pixels)      # warm-up: buffers, pinned ring, worker threads
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        g.fit_arrays(f, data, w.psf_pixels)
        ts.append(time.perf_counter() - t)
    dt = float(np.median(ts))
    in_bytes = 4 * w.box * w.box
    out_bytes = 4 * (2 * len(w.psf_pixels) + 1)
    h2d = pcie_h2d_gbs()
    bound = h2d * 1e9 / in_bytes
    return {"entry_point": "gmle_fit (host arrays, synchronous)", "rois": w.n, "ms_per_call": round(dt * 1e3, 3),
            "fits_per_s": round(w.n / dt, 1), "bytes_in_per_roi": in_bytes, "bytes_out_per_roi": out_bytes,
            "pcie_h2d_gbs": round(h2d, 2), "pcie_bound_fits_per_s": round(bound, 1),
            "fraction_of_pcie_bound": round(w.n / dt / bound, 4),
            "vs_reference_gpu_host_call": round(w.n / dt / REFERENCE_GPU_FITS_PER_S, 3),
            "host_threads": int(os.environ.get("GMLE_HOST_THREADS", min(8, os.cpu_count() or 1))),
            "note": "median of 5 calls; pageable numpy arrays in and out"}


def trace_summary(path: str, drop_first: int = 1, last: int = 0) -> int:
    """Per-kernel duration statistics of a rocprofv3 kernel trace (CSV), one
    line per (kernel, grid size), leaving out each line's first `drop_first`
    dispatches (cold caches, code-object load), or keeping only its last
    `last` dispatches when `last` > 0 (the timed steps at the end of a run)."""
    import csv
    rows = {}
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            name = r.get("Kernel_Name") or r.get("KernelName") or ""
            grid = "x".join(r[k] for k in ("Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z") if k in r) \
                or r.get("Grid_Size", "")
            t = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            rows.setdefault((name, grid), []).append((int(r["Start_Timestamp"]), t))
    out = []
    for (name, grid), ts in rows.items():
        ts.sort()
        if last > 0:
            kept = ts[-last:]
        else:
            kept = ts[drop_first:] if len(ts) > drop_first else ts
        a = np.asarray([t for _, t in kept], dtype=np.float64)
        out.append({"kernel": name[:160], "grid": grid, "dispatches": len(ts), "used": len(kept),
                    "mean_ns": round(a.mean(), 1), "median_ns": round(float(np.median(a)), 1),
                    "min_ns": int(a.min()), "max_ns": int(a.max())})
    out.sort(key=lambda r: -r["mean_ns"] * r["used"])
    for r in out:
        print(json.dumps(r))
    return 0


BOXER = {"ny": 2048, "nx": 2048, "boxsize": 11, "overlap": 2.0, "sigma_small": 1.0, "sigma_large": 2.0,
         "minval": 5.0, "spots_per_frame": 1000}


def run_boxer(n_frames: int = 32, reps: int = 10, cpu: bool = True) -> dict:
    """ROI extraction on 2048 x 2048 frames (one call over the whole stack),
    then the pipeline frames -> boxes -> fits (XYNB 11 x 11, 20 iterations)
    with the detection count read back between the two."""
    import torch
    from gaussmle_amd import boxer as gbx
    from gaussmle_amd import synth
    from gaussmle_amd import device as gd
    from gaussmle_amd.models import GaussianXYNB, default_constraints
    c = BOXER
    ny, nx = c["ny"], c["nx"]
    frames, _ = synth.spot_frames(ny, nx, n_frames, c["spots_per_frame"], seed=5)
    d = torch.from_numpy(frames.reshape(-1, order="F")).cuda()
    cap = 2 * n_frames * c["spots_per_frame"]
    args = (ny, nx, n_frames, c["boxsize"], c["overlap"], c["sigma_small"], c["sigma_large"], c["minval"])
    out = gbx.find_boxes_device(d, *args, max_boxes=cap)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(reps):
        gbx.find_boxes_device(d, *args, max_boxes=cap, out=out)
    b.record(s)
    b.synchronize()
    t = a.elapsed_time(b) / reps / 1e3
    n = int(out["count"].item())
    npix = ny * nx * n_frames
    win = gbx.suppression_radius(c["boxsize"], c["overlap"])
    # algorithmic HBM bytes: frames read, filtered stack written, candidate /
    # survivor words (8 B per 64 pixels, written and read twice), each
    # detection's window re-read and its box read and written
    per_pix = 4 + 4 + 4 * 8 / 64
    per_det = (2 * win + 1) ** 2 * 4 + 2 * c["boxsize"] ** 2 * 4 + 16
    algo = per_pix * npix + per_det * n
    res = {"frames": n_frames, "frame_shape": [ny, nx], "detections": n, "ms_per_call": round(t * 1e3, 4),
           "frames_per_s": round(n_frames / t, 1), "pixels_per_s": round(npix / t, 1),
           "algorithmic_bytes_per_pixel": round(algo / npix, 3), "hbm_gbs": round(algo / t / 1e9, 1),
           "frac_hbm": round(algo / t / 1e9 / PEAK_HBM_GBS, 4),
           "params": {k: v for k, v in c.items() if k not in ("ny", "nx")},
           "data": "synthetic: seeded pixel-integrated Gaussian spots, flat background, Poisson-limit noise"}

    # frames -> boxes -> fits
    psf = GaussianXYNB(1.3)
    cons = default_constraints(psf, 11)
    fit_out = None

    def pipeline():
        nonlocal fit_out
        gbx.find_boxes_device(d, *args, max_boxes=cap, out=out)
        m = min(int(out["count"].item()), cap)
        fit_out = gd.fit_device(psf, 11, out["rois"][:m].reshape(-1), cons, 20)
        return m

    pipeline()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        m = pipeline()
    torch.cuda.synchronize()
    tp = (time.perf_counter() - t0) / reps
    res["pipeline"] = {"ms_per_call": round(tp * 1e3, 4), "frames_per_s": round(n_frames / tp, 1),
                       "fits_per_s": round(m / tp, 1),
                       "note": "find boxes, read the count back, fit every box; wall clock"}
    if cpu:
        from oracle import refboxer
        f1 = np.asfortranarray(frames[:, :, :2])
        refboxer.find(f1[:, :, :1], c["boxsize"], c["overlap"], c["sigma_small"], c["sigma_large"], c["minval"])
        t0 = time.perf_counter()
        refboxer.find(f1, c["boxsize"], c["overlap"], c["sigma_small"], c["sigma_large"], c["minval"])
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(2 / dt, 2), "unit": "frames/s", "cores": 1, "kind": "port",
                               "sample": f"2 frames of the same stack, oracle/refboxer.cpp, {dt:.2f}s"}
    return res


def run_simulator(n: int = 1_048_576, reps: int = 10, cpu: bool = True) -> dict:
    """The on-device simulator: n ROIs of 11 x 11 XYNB (the headline shape),
    Poisson noise, per-ROI parameters resident on the device."""
    import torch
    from gaussmle_amd import simulator as gsim
    from gaussmle_amd.models import GaussianXYNB
    psf = GaussianXYNB(1.3)
    rng = np.random.default_rng(9)
    params = np.stack([rng.uniform(4.5, 7.5, n), rng.uniform(4.5, 7.5, n), rng.uniform(500, 5000, n),
                       rng.uniform(1, 20, n)], axis=1).astype(np.float32)
    d = torch.from_numpy(params).cuda()
    out = gsim.simulate_device(psf, 11, d, 1)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    for r in range(reps):
        gsim.simulate_device(psf, 11, d, 2 + r, out=out)
    b.record(s)
    b.synchronize()
    t = a.elapsed_time(b) / reps / 1e3
    res = {"rois": n, "box": 11, "model": "GaussianXYNB", "camera": "ideal", "ms_per_call": round(t * 1e3, 4),
           "rois_per_s": round(n / t, 1), "pixels_per_s": round(n * 121 / t, 1),
           "hbm_gbs": round((n * 121 * 4 + n * 16) / t / 1e9, 1)}
    if cpu:
        from oracle import refsim
        m = 20_000
        t0 = time.perf_counter()
        refsim.simulate(psf, 11, params[:m], seed=1)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(m / dt, 1), "unit": "ROIs/s", "cores": 1, "kind": "port",
                               "sample": f"{m} ROIs, oracle/refsim.cpp, {dt:.2f}s"}
    return res


def run_service(n_frames: int = 300, submissions: int = 3000, loads=(0.1, 0.5, 0.9, 1.0), slots: int = 4,
                lanes=None, seed: int = 4, peak_repeat: int = 10):
    """Streaming service: seeded frames (ROI count log-uniform 50-5000, box
    sizes {9,11,13,15}, XYNB / XYNBS) of a 2048x2048 uint16 sCMOS sensor,
    their raw counts back to back in one pinned pool (a camera's frame ring)
    and submitted in ring order; the peak with every frame queued at once,
    peak_repeat passes over the ring; Poisson arrivals at fractions of that
    peak; p50 / p99 latency from scheduled arrival to completion, and
    sustained fits/s."""
    from gaussmle_amd import batch as gb
    from gaussmle_amd import service as sv
    import torch

    t0 = time.perf_counter()
    maps = sv.sensor_maps(seed=seed, ny=2048, nx=2048)
    frames = sv.make_frames(n_frames, seed=seed + 1, maps=maps)
    rois = np.array([f.n for f in frames])
    print(f"[bench] generated {n_frames} frames ({rois.sum()} ROIs) in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    psfs, cons, table = sv.default_setup(lanes=lanes if lanes is not None else 2)
    # frames' raw counts back to back in one pinned pool, as a camera's frame
    # ring would hold them: consecutive requests then reach the GPU by one DMA
    pool = sv.CountPool(sum(f.adu.size for f in frames))
    reqs = [sv.Request(f, pool=pool) for f in frames]
    rng = np.random.default_rng(seed + 2)
    res = {"workload": "frames of 50-5000 ROIs (log-uniform), box 9-15, XYNB/XYNBS, 2048x2048 uint16 sCMOS",
           "frames": n_frames, "mean_rois_per_frame": round(float(rois.mean()), 1), "slots": slots,
           "lanes": {f"{m}:{b}": L for (m, b), L in table.items()}}
    with sv.Service(psfs, cons, table, hbm_budget=8 << 30, slots=slots) as svc:
        svc.calibrate(*maps)
        # the peak: every frame queued at once, peak_repeat times over, so the
        # pipeline's fill and drain weigh less
        burst = reqs * peak_repeat
        lat, el = svc.replay(burst, np.zeros(len(burst)))          # warm-up
        peaks = []
        for _ in range(3):
            lat, el = svc.replay(burst, np.zeros(len(burst)))
            peaks.append(peak_repeat * rois.sum() / el)
        peak = max(peaks)
        res["peak_fits_per_s"] = round(peak, 1)
        # batch statistics of the warm-up and peak replays alone (all requests queued at once)
        res["peak_batch_stats"] = svc.stats()
        print(f"[bench] service peak {peak:.4g} fits/s {res['peak_batch_stats']}", file=sys.stderr, flush=True)
        res["loads"] = []
        for frac in loads:
            order = np.arange(submissions) % len(reqs)   # frames in ring order
            rate = frac * peak / rois.mean()                         # requests / s
            arrivals = np.cumsum(rng.exponential(1.0 / rate, submissions))
            arrivals -= arrivals[0]
            lat, el = svc.replay([reqs[i] for i in order], arrivals)
            res["loads"].append({"load": frac, "offered_fits_per_s": round(float(rate * rois.mean()), 1),
                                 "sustained_fits_per_s": round(float(rois[order].sum() / el), 1),
                                 "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 4),
                                 "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 4),
                                 "max_ms": round(float(lat.max()) * 1e3, 4)})
            print(f"[bench] service load {frac}: {res['loads'][-1]} {svc.stats()}", file=sys.stderr, flush=True)
        res["batches"] = svc.info()["batches"]
        res["batch_stats"] = svc.stats()
    for r in reqs:
        r.close()
    pool.close()
    # PCIe host-to-device rate of one large pinned copy: the ceiling for any
    # path that must move the frames' raw counts to the GPU
    src = torch.empty(256 << 20, dtype=torch.uint8).pin_memory()
    dst = torch.empty(256 << 20, dtype=torch.uint8, device="cuda:0")
    dst.copy_(src, non_blocking=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(5):
        dst.copy_(src, non_blocking=True)
    torch.cuda.synchronize()
    res["pcie_h2d_gbs"] = round(5 * (256 << 20) / (time.perf_counter() - t1) / 1e9, 2)
    res["raw_bytes_per_roi"] = round(float(sum(f.adu.nbytes for f in frames)) / rois.sum(), 1)
    del src, dst
    # the plain batch kernel on the same ROIs (electrons, device-resident)
    mb = sv.frames_batch(frames, maps, psfs, cons)
    d = gb.upload_batch(mb)
    t = gb._lanes_table(table)
    out = gb.fit_batch_device(mb, d, "fp32", t, 4)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(5):
        out = gb.fit_batch_device(mb, d, "fp32", t, 4, out=out)
    torch.cuda.synchronize()
    res["plain_batch_fits_per_s"] = round(5 * mb.n / (time.perf_counter() - t1), 1)
    # the same with host arrays: upload of Float32 electrons, fit, download
    gb.fit_batch(mb, "fp32", t, 4)
    t1 = time.perf_counter()
    for _ in range(3):
        gb.fit_batch(mb, "fp32", t, 4)
    res["plain_host_path_fits_per_s"] = round(3 * mb.n / (time.perf_counter() - t1), 1)
    # the service's peak against the batch kernel on device-resident ROIs and
    # against the PCIe ceiling of its raw counts
    pcie_bound = res["pcie_h2d_gbs"] * 1e9 / res["raw_bytes_per_roi"]
    res["pcie_bound_fits_per_s"] = round(pcie_bound, 1)
    res["peak_over_plain_batch"] = round(res["peak_fits_per_s"] / res["plain_batch_fits_per_s"], 4)
    res["peak_over_pcie_bound"] = round(res["peak_fits_per_s"] / pcie_bound, 4)
    return res


def make_workload(spec, n):
    from gaussmle_amd import synth
    return getattr(synth, spec["gen"])(n=n)


def traffic_from_profiles(workload: str, n_rois: int):
    """HBM bytes per launch measured offline with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE) for this workload, if recorded under profiles/."""
    p = ROOT / "profiles" / f"traffic_{workload}.json"
    if not p.exists():
        return None
    try:
        rec = json.loads(p.read_text())
        return float(rec["bytes_per_roi"]) * n_rois
    except (KeyError, ValueError):
        return None


def host_cpu() -> dict:
    """CPU model and core counts of the host the baseline runs on."""
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": usable, "cpu_count": os.cpu_count()}


def instruction_mix(workload: str):
    """VALU / transcendental instructions per fit measured offline with
    rocprofv3 --pmc (SQ_INSTS_VALU, SQ_INSTS_VALU_TRANS_F32), if recorded."""
    p = ROOT / "profiles" / f"traffic_{workload}.json"
    try:
        return json.loads(p.read_text())["instruction_mix_per_fit"]
    except (OSError, KeyError, ValueError):
        return None


def issue_rates():
    """Clocks per SIMD instruction at two resident waves per SIMD (the fit
    kernels' occupancy), measured by gmle_issue_probe, if recorded."""
    p = ROOT / "profiles" / "valu_issue_rates.json"
    try:
        return json.loads(p.read_text())["model_at_2_waves"]
    except (OSError, KeyError, ValueError):
        return None


def issue_model(mix, rates, waves_per_simd: float, kernel_ms: float):
    """Launch time the profiled instruction mix needs at the measured issue
    rates and clock, and the kernel's fraction of it (1 = issue-bound with
    no stall)."""
    if not mix or not rates or not mix.get("clock_ghz_under_load"):
        return None
    trans = mix["valu_trans_f32"]
    cycles = (mix["valu"] - trans) * rates["cycles_per_valu"] + trans * rates["cycles_per_trans"]
    ms = waves_per_simd * cycles / (mix["clock_ghz_under_load"] * 1e9) * 1e3
    return {"ms": round(ms, 4), "frac": round(ms / kernel_ms, 4), "cycles_per_valu": rates["cycles_per_valu"],
            "cycles_per_trans": rates["cycles_per_trans"], "clock_ghz": mix["clock_ghz_under_load"],
            "waves_per_simd": round(waves_per_simd, 3)}


def cpu_baseline(w, sample: int, threads: int, runs: int = 5):
    """Time the oracle (the reference's algorithm restated in C++, Float32) on
    the first `sample` ROIs with `threads` OpenMP threads: one warm-up run,
    then the median of `runs`.  Returns (fits/s of the fit loop, median s,
    fits/s end to end with the host p-value epilogue)."""
    from gaussmle_amd.fitter import pvalues
    from oracle import refcpu
    refcpu.build()
    sub = w.subset(min(sample, w.n))
    refcpu.fit_workload(sub, refcpu.FP32, nthreads=threads)   # warm-up
    ts, te = [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        r = refcpu.fit_workload(sub, refcpu.FP32, nthreads=threads)
        t1 = time.perf_counter()
        pvalues(r["llr"].astype(np.float32), sub.box, len(sub.psf_pixels))
        ts.append(t1 - t0)
        te.append(time.perf_counter() - t0)
    dt = float(np.median(ts))
    return sub.n / dt, dt, sub.n / float(np.median(te))


def run_epilogue(out: dict, w, d: dict, reps: int = 20) -> dict:
    """The on-device result epilogue (microns, p-values) over the step's
    results: HIP-event time per launch and its HBM roofline."""
    import torch
    from gaussmle_amd.device import epilogue_device
    n, P = out["params"].shape
    psx = float(np.float32(0.1))
    args = (out["params"], out["crlb"], out["llr"], psx, psx, w.psf_pixels.model_id, w.box)
    kw = {"d_xc": d.get("xc"), "d_yc": d.get("yc")}
    o = epilogue_device(*args, **kw)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(reps):
        epilogue_device(*args, **kw, out=o)
    b.record(s)
    b.synchronize()
    t = a.elapsed_time(b) / reps / 1e3
    per_roi = 4 * (2 * P + 1) + (8 if "xc" in d else 0) + 4 * o.shape[1]
    gbs = per_roi * n / t / 1e9
    return {"rois": n, "kernel_us": round(t * 1e6, 2), "rois_per_s": round(n / t, 1), "bytes_per_roi": per_roi,
            "hbm_gbs": round(gbs, 1), "frac_hbm": round(gbs / PEAK_HBM_GBS, 4),
            "outputs": "x, y, sigma_x, sigma_y (um), p-value, widths and their sigmas (um), chi2"}


def copy_probe(nbytes: int) -> int:
    """A device-to-device copy of known size, to scale rocprofv3's FETCH_SIZE /
    WRITE_SIZE counters into bytes."""
    import torch
    n = nbytes // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "kernels": 3}), flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    # The GPU's clock ramps up over the first ~30 ms of sustained work (the
    # headline kernel's launches shorten from ~1.75 to ~1.47 ms over its first
    # 20 dispatches): the warm-up covers that ramp.
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="xynb11",
                    choices=sorted(WORKLOADS) + sorted(CONFIGS) + ["mixed", "serve", "boxer"])
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-lanes", type=int, default=2)
    ap.add_argument("--serve-peak-repeat", type=int, default=10,
                    help="--workload serve: the peak replays the frames this many times over, all queued at once")
    ap.add_argument("--serve-loads", default="0.1,0.5,0.9,1.0",
                    help="--workload serve: load levels (fractions of the measured peak, comma-separated; "
                         "empty: the peak runs only)")
    ap.add_argument("--mixed-rois", type=int, default=4_000_000,
                    help="ROIs of the uneven batch reported beside the headline line (0: skip)")
    ap.add_argument("--serve-frames", type=int, default=300,
                    help="frames of the streaming-service run reported beside the headline line (0: skip)")
    ap.add_argument("--boxer-frames", type=int, default=32,
                    help="2048x2048 frames of the ROI-extraction run reported beside the headline line (0: skip)")
    ap.add_argument("--sim-rois", type=int, default=1_048_576,
                    help="ROIs of the simulator run reported beside the headline line (0: skip)")
    ap.add_argument("--rois", type=int, default=0, help="ROIs per step (default: workload's)")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "mixed"])
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=100_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes (HBM counter calibration)")
    ap.add_argument("--configs", default="config1,config2,box21",
                    help="configurations reported beside the headline line (comma list of "
                         f"{','.join(CONFIGS)}; empty: none)")
    ap.add_argument("--trace-summary", default=None, metavar="CSV",
                    help="only summarise a rocprofv3 kernel trace (per-kernel durations, first dispatch dropped)")
    ap.add_argument("--trace-last", type=int, default=0, metavar="K",
                    help="with --trace-summary: keep only each kernel's last K dispatches")
    ap.add_argument("--batched-only", action="store_true",
                    help="--workload mixed: time only the tuned batched call (no one-stream, one-launch or "
                         "per-shape plain comparisons)")
    args = ap.parse_args(argv)
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.trace_summary:
        return trace_summary(args.trace_summary, last=args.trace_last)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the per-ROI fit shards trivially across GPUs (independent ROIs), "
              "and multi-GPU runs are not part of this build. Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch
    from gaussmle_amd import _lib
    from gaussmle_amd import device as gd

    if args.workload == "mixed":
        return main_mixed(args)
    if args.workload in CONFIGS:
        res = run_config(args.workload, args.steps, args.warmup, args.rois or None, args.lanes)
        print(json.dumps({"metric": METRIC, args.workload: res}), flush=True)
        return 0
    if args.workload == "boxer":
        print(json.dumps({"metric": METRIC, "roi_extraction": run_boxer(args.boxer_frames or 32, args.steps,
                                                                       cpu=not args.no_cpu_baseline)}), flush=True)
        return 0
    if args.workload == "serve":
        loads = tuple(float(x) for x in args.serve_loads.split(",") if x.strip())
        res = run_service(args.serve_frames or 300, slots=args.serve_slots, lanes=args.serve_lanes, loads=loads,
                          peak_repeat=args.serve_peak_repeat)
        print(json.dumps({"metric": METRIC, "serving": res}), flush=True)
        return 0
    spec = WORKLOADS[args.workload]
    n = args.rois or spec["rois"]
    t0 = time.perf_counter()
    w = make_workload(spec, n)
    print(f"[bench] generated {n} ROIs ({spec['label']}) in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    lib = _lib.load()
    lanes = args.lanes or lib.gmle_default_lanes(w.psf_pixels.model_id, w.camera, w.box,
                                                 _lib.PRECISION_IDS[args.precision], n)
    torch.cuda.init()
    dev = torch.device("cuda:0")
    d = gd.upload_workload(w, dev)
    ny, nx = (0, 0) if w.var_map is None else w.var_map.shape
    out = None

    def step():
        nonlocal out
        out = gd.fit_device(w.psf_pixels, w.box, d["data"], w.constraints, w.iterations, args.precision, lanes,
                            d.get("var"), ny, nx, d.get("xc"), d.get("yc"), out=out)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in ev]
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    value = args.steps * n / wall

    epilogue = run_epilogue(out, w, d)

    # Smaller batch: the reference's config-0 size (10,000 ROIs per launch).
    small = {}
    if args.workload == "xynb11":
        ws = w.subset(10_000)
        ds = gd.upload_workload(ws, dev)
        ls = lib.gmle_default_lanes(0, 0, 11, _lib.PRECISION_IDS[args.precision], ws.n)
        o = None
        for _ in range(3):
            o = gd.fit_device(ws.psf_pixels, 11, ds["data"], ws.constraints, 20, args.precision, ls, out=o)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        reps = 50
        for _ in range(reps):
            o = gd.fit_device(ws.psf_pixels, 11, ds["data"], ws.constraints, 20, args.precision, ls, out=o)
        torch.cuda.synchronize()
        small = {"rois_per_launch": ws.n, "lanes": ls, "fits_per_s": reps * ws.n / (time.perf_counter() - t1)}

    F = flops_per_fit(spec["model"], w.box, w.iterations, spec["camera"])
    achieved = F * n / avg_kernel_s / 1e12
    traffic = traffic_from_profiles(args.workload, n)
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                "kernel_ms": round(avg_kernel_s * 1e3, 4), "flops_per_fit": F,
                "hbm_gbs": round(bytes_per_fit(w.box, len(w.psf_pixels)) * n / avg_kernel_s / 1e9, 2)}
    mix = instruction_mix(args.workload)
    if mix and lanes == 1 and args.precision == "fp32":
        # lane instructions per fit (= wave instructions per wave of 64
        # one-lane fits), from the profile of this kernel
        roofline["valu_insts_per_fit"] = mix["valu"]
        roofline["trans_per_fit"] = mix["valu_trans_f32"]
        simds = torch.cuda.get_device_properties(dev).multi_processor_count * 4
        model = issue_model(mix, issue_rates(), n / 64 / simds, avg_kernel_s * 1e3)
        if model:
            roofline["issue_model"] = model

    cpu = None
    if not args.no_cpu_baseline:
        host = host_cpu()
        threads = min(host["nproc"], 16)
        rate, dt, rate_e2e = cpu_baseline(w, args.cpu_sample, threads)
        rate1, dt1, _ = cpu_baseline(w, max(2000, args.cpu_sample // 20), 1)
        cpu = {"value": round(rate, 1), "unit": "fits/s", "cores": threads, "kind": "port",
               "sample": f"first {min(args.cpu_sample, n)} ROIs of the same workload, Float32 oracle "
                         f"(oracle/refcpu.cpp, OpenMP dynamic over ROIs), median of 5 runs after a warm-up, "
                         f"{dt:.2f}s per run",
               "single_thread_fits_per_s": round(rate1, 1),
               "end_to_end_fits_per_s": round(rate_e2e, 1),
               "end_to_end": "fit loop + host p-value epilogue (chi-square CDF, scipy), median of 5",
               **host}

    host_path = run_host_path(w) if args.workload == "xynb11" else None
    configs = {}
    if args.workload == "xynb11":
        for key in [k for k in args.configs.split(",") if k]:
            configs[key] = run_config(key, max(3, args.steps // 4), 2)
    batched = None
    if args.workload == "xynb11" and args.mixed_rois > 0:
        _, _, batched, _ = run_mixed(args.mixed_rois, max(3, args.steps // 4), 2, args.precision)
    serving = None
    if args.workload == "xynb11" and args.serve_frames > 0:
        serving = run_service(args.serve_frames)
        # the same service with 16 lanes per ROI: shorter fits for small
        # batches (latency), at the cost of peak throughput
        low = run_service(args.serve_frames, loads=(0.1,), lanes=16)
        serving["low_latency"] = {k: low[k] for k in ("lanes", "peak_fits_per_s", "loads")}
    roi_extraction = None
    if args.workload == "xynb11" and args.boxer_frames > 0:
        roi_extraction = run_boxer(args.boxer_frames, cpu=not args.no_cpu_baseline)
    simulator = None
    if args.workload == "xynb11" and args.sim_rois > 0:
        simulator = run_simulator(args.sim_rois, cpu=not args.no_cpu_baseline)

    rec = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / REFERENCE_GPU_FITS_PER_S, 3) if args.workload == "xynb11" else None,
        "vs_baseline_note": ("value (inputs resident in HBM) over the reference's published ~1e7 fits/s for 11x11 "
                             "XYNB on an RTX 4090 through fit() on host arrays; the host-array call of this "
                             "library against the same figure is host_path.vs_reference_gpu_host_call")
        if args.workload == "xynb11" else None,
        "dtype": "f32" if args.precision == "fp32" else "f32+f64",
        "data": "synthetic: seeded Poisson ROIs (numpy PCG64), device-resident before timing",
        "config": {"workload": spec["label"], "model": type(w.psf_pixels).__name__, "box": w.box,
                   "iterations": w.iterations, "rois_per_step": n, "camera": spec["camera"],
                   "precision": args.precision, "lanes_per_roi": lanes, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "config0_batch": small or None,
        "host_path": host_path,
        **configs,
        "epilogue": epilogue,
        "batched": batched,
        "serving": serving,
        "roi_extraction": roi_extraction,
        "simulator": simulator,
    }
    print(json.dumps(rec), flush=True)
    return 0


def main_mixed(args) -> int:
    """--workload mixed: the uneven batch (4M ROIs by default) is the line."""
    n = args.rois or 4_000_000
    mb, d, res, dev_s = run_mixed(n, args.steps, args.warmup, args.precision, batched_only=args.batched_only)
    achieved = res["flops"] / dev_s / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                "kernel_ms": round(dev_s * 1e3, 4), "flops_per_fit": round(res["flops"] / n, 1),
                "hbm_gbs": round(res["bytes"] / dev_s / 1e9, 2),
                "note": "whole batched call (grouping kernels + every shape's fit kernel), HIP events"}
    cpu = None
    if not args.no_cpu_baseline:
        from gaussmle_amd import synth_device
        k = max(20_000, args.cpu_sample // 2)
        cpu = cpu_baseline_mixed(synth_device.host_batch(mb, d, k), k, min(os.cpu_count() or 1, 16))
    rec = {
        "metric": METRIC, "value": res["batched_fits_per_s"], "unit": "fits/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(n / res["batched_fits_per_s"] * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f32+f64",
        "data": "synthetic: seeded sCMOS ROIs (numpy PCG64), device-resident before timing",
        "config": {"workload": MIXED_LABEL, "rois_per_step": n, "iterations": 20, "camera": "scmos",
                   "precision": args.precision, "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "batched": res,
    }
    print(json.dumps(rec), flush=True)
    return 0


def cpu_baseline_mixed(mb, sample: int, threads: int):
    """The oracle on the first `sample` ROIs of the batch, shape by shape."""
    from gaussmle_amd import batch as gb
    from gaussmle_amd import synth
    from oracle import refcpu
    refcpu.build()
    sub = np.arange(min(sample, mb.n))
    total = 0.0
    for s, c in enumerate(mb.subset(sub).shape_counts()):
        if not c:
            continue
        m, b = gb.shape_of(s)
        idx = sub[(mb.models[sub] == m) & (mb.boxes[sub] == b)]
        w = synth.Workload("s", mb.psfs[m], b, mb.roi_stack(idx, b), mb.truth[:, idx], mb.constraints[m],
                           mb.iterations, mb.var_map, mb.x_corners[idx].copy(), mb.y_corners[idx].copy())
        t0 = time.perf_counter()
        refcpu.fit_workload(w, refcpu.FP32, nthreads=threads)
        total += time.perf_counter() - t0
    return {"value": round(len(sub) / total, 1), "unit": "fits/s", "cores": threads, "kind": "port",
            "sample": f"first {len(sub)} ROIs of the same batch, Float32 oracle shape by shape, {total:.1f}s"}


if __name__ == "__main__":
    sys.exit(main())
