#!/usr/bin/env python3
"""Throughput benchmark of the MI355X per-ROI MLE fit (one GPU).

A step is one launch of the fitting kernel over one batch of device-resident,
seeded synthetic ROIs.  The default workload is the headline one: 11x11
GaussianXYNB (sigma 1.3 px), Poisson camera, 20 Newton iterations, with the
parameter draws x, y = centre +- 1.5 px, N ~ U(500, 5000), bg ~ U(1, 20); the
batch is 2^20 ROIs.  Prints one JSON line with fits/s, the kernel's roofline
position (FP32 VALU, algorithmic FLOPs per fit as defined in DESIGN.md) and a
CPU baseline (the oracle, timed on this host's cores on a bounded sample).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload xynb11|xynbs13|astig15]
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


def cpu_baseline(w, sample: int, threads: int):
    """Time the oracle (the reference's algorithm restated in C++, Float32) on
    the first `sample` ROIs with `threads` OpenMP threads."""
    from oracle import refcpu
    refcpu.build()
    sub = w.subset(min(sample, w.n))
    refcpu.fit_workload(sub.subset(min(256, sub.n)), refcpu.FP32, nthreads=threads)   # warm-up
    t0 = time.perf_counter()
    refcpu.fit_workload(sub, refcpu.FP32, nthreads=threads)
    dt = time.perf_counter() - t0
    return sub.n / dt, dt


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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="xynb11", choices=sorted(WORKLOADS))
    ap.add_argument("--rois", type=int, default=0, help="ROIs per step (default: workload's)")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "mixed"])
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=100_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes (HBM counter calibration)")
    args = ap.parse_args(argv)
    if args.copy_probe:
        return copy_probe(args.copy_probe)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the per-ROI fit shards trivially across GPUs (independent ROIs), "
              "and multi-GPU runs are not part of this build. Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch
    from gaussmle_amd import _lib
    from gaussmle_amd import device as gd

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

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        rate, dt = cpu_baseline(w, args.cpu_sample, threads)
        rate1, dt1 = cpu_baseline(w, max(2000, args.cpu_sample // 20), 1)
        cpu = {"value": round(rate, 1), "unit": "fits/s", "cores": threads, "kind": "port",
               "sample": f"first {min(args.cpu_sample, n)} ROIs of the same workload, Float32 oracle "
                         f"(oracle/refcpu.cpp, OpenMP), {dt:.1f}s wall",
               "single_thread_fits_per_s": round(rate1, 1)}

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
        "dtype": "f32" if args.precision == "fp32" else "f32+f64",
        "data": "synthetic: seeded Poisson ROIs (numpy PCG64), device-resident before timing",
        "config": {"workload": spec["label"], "model": type(w.psf_pixels).__name__, "box": w.box,
                   "iterations": w.iterations, "rois_per_step": n, "camera": spec["camera"],
                   "precision": args.precision, "lanes_per_roi": lanes, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "config0_batch": small or None,
    }
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
