"""Build libgaussmle_hip.so (gfx950) from gaussmle.jl_amd/csrc.

Each kernel translation unit (one per PSF model x camera x precision) is
compiled separately with hipcc and in parallel, then linked with the C-ABI
layer into ``gaussmle.jl_amd/lib/libgaussmle_hip.so``.

    python gaussmle.jl_amd/build.py [--force] [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libgaussmle_hip.so"
PROBE_SRC = CSRC / "probe" / "issue_probe.hip"
PROBE = LIBDIR / "gmle_issue_probe"
INCLUDE = PKG.parent / "include"

ARCH = "gfx950"
# Packed FP32 (v_pk_*_f32) is switched off: the kernels are written for scalar
# FP32 issue and do not rely on packed forms.
HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
]

# Per-file flags.  inst_sched_ilp.hip and inst_sched_maxilp.hip hold the
# specialisations that run faster under another machine scheduler (each file
# lists them and the measurements).
SOURCE_FLAGS = {
    "inst_sched_ilp.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"],
    "inst_sched_maxilp.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
}


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; the ROCm toolchain is required to build the HIP library")
    return exe


def _headers_mtime() -> float:
    hs = list(CSRC.glob("*.hpp")) + list(INCLUDE.glob("*.h"))
    return max(p.stat().st_mtime for p in hs)


def _deps_mtime(obj: Path, fallback: float) -> float:
    """Newest mtime among the files the object's depfile lists (hipcc -MMD),
    or `fallback` (every header) when there is no usable depfile."""
    dep = obj.with_suffix(".d")
    try:
        text = dep.read_text().replace("\\\n", " ")
        files = text.split(":", 1)[1].split()
        return max(Path(f).stat().st_mtime for f in files)
    except (OSError, IndexError, ValueError):
        return fallback


def _compile(src: Path, force: bool, hdr_mtime: float) -> Path:
    obj = BUILD / (src.stem + ".o")
    if not force and obj.exists() and obj.stat().st_mtime >= max(src.stat().st_mtime, _deps_mtime(obj, hdr_mtime)):
        return obj
    cmd = [hipcc(), *HIPFLAGS, *SOURCE_FLAGS.get(src.name, []), "-I", str(INCLUDE), "-MMD", "-MF",
           str(obj.with_suffix(".d")), "-c", str(src), "-o", str(obj)]
    multi = src.name.startswith("inst_multi_")
    if multi:
        cmd.append("-Rpass-analysis=kernel-resource-usage")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr[-4000:]}")
    if multi:
        # The multi-shape kernel is sized to run two waves per SIMD: its tile
        # functions' register limits (max_box_within_256) assume this compiler.
        occ = re.findall(r"Occupancy \[waves/SIMD\]: (\d+)", r.stderr)
        if occ and min(int(o) for o in occ) < 2:
            print(f"[build] warning: {src.name}: multi-shape kernel at {min(int(o) for o in occ)} wave(s) per SIMD; "
                  "lower the limits in max_box_within_256", flush=True)
    return obj


def build(force: bool = False, jobs: int | None = None, verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + [CSRC / "gmle_api.cpp"]
    hdr = _headers_mtime()
    jobs = jobs or min(16, os.cpu_count() or 4)
    # Largest translation units first so the parallel build ends together.
    srcs.sort(key=lambda p: ("mixed" not in p.name, "astig" not in p.name, p.name))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_compile, s, force, hdr): s for s in srcs}
        for f in cf.as_completed(futs):
            objs.append(f.result())
            if verbose:
                print(f"[build] {futs[f].name}", flush=True)
    newest = max(o.stat().st_mtime for o in objs)
    if force or not LIB.exists() or LIB.stat().st_mtime < newest:
        tmp = LIB.with_suffix(".so.tmp")
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, sorted(objs))]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
        os.replace(tmp, LIB)
    if verbose:
        print(f"[build] {LIB}", flush=True)
    return LIB


def build_probe(force: bool = False, verbose: bool = True) -> Path:
    """The VALU / MFMA issue-rate probe (a stand-alone executable, not part of
    the library).  Its MFMA accumulators are kept in VGPRs so the mixed
    MFMA + FMA loop carries no AGPR copies."""
    if force or not PROBE.exists() or PROBE.stat().st_mtime < PROBE_SRC.stat().st_mtime:
        LIBDIR.mkdir(exist_ok=True)
        cmd = [hipcc(), *HIPFLAGS, "-mllvm", "-amdgpu-mfma-vgpr-form=1", str(PROBE_SRC), "-o", str(PROBE)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {PROBE_SRC.name}:\n{r.stderr[-4000:]}")
    if verbose:
        print(f"[build] {PROBE}", flush=True)
    return PROBE


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--force", action="store_true", help="recompile every translation unit")
    ap.add_argument("--jobs", type=int, default=None)
    args = ap.parse_args(argv)
    build(force=args.force, jobs=args.jobs)
    build_probe(force=args.force)
    return 0


if __name__ == "__main__":
    sys.exit(main())
