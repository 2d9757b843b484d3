"""The C-ABI library loads, exports every symbol include/gaussmle.h declares,
and rejects bad arguments before touching a device (CPU only)."""
import ctypes as C
import math
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "gaussmle.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(gmle_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_abi():
    names = declared_functions()
    for n in ["gmle_fit", "gmle_fit_device", "gmle_device_count", "gmle_last_error", "gmle_version"]:
        assert n in names


def test_library_exports_every_declared_symbol(hiplib):
    from gaussmle_amd import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (gmle_\w+)", out))
    for name in declared_functions():
        assert name in exported, name
        assert hasattr(hiplib, name)
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_library_needs_no_recent_hip_symbol_versions(hiplib):
    """The library binds to the HIP runtime torch bundles (loaded first, so
    one runtime serves the process), which is older than the image's ROCm.
    An entry point that exists only in a newer runtime (hipMemcpyBatchAsync
    carries hip_7.1) makes the library fail to load there, so every HIP
    symbol version it needs must be an old one."""
    from gaussmle_amd import _lib
    out = subprocess.run(["readelf", "-V", str(_lib.LIB_PATH)], capture_output=True, text=True, check=True).stdout
    needs = out[out.index("Version needs"):] if "Version needs" in out else ""
    versions = {tuple(int(x) for x in v.split(".")) for v in re.findall(r"Name: hip_([0-9.]+)", needs)}
    assert versions, "no HIP symbol versions found"
    assert max(versions) < (6, 0), sorted(versions)


def test_scalar_queries(hiplib):
    assert hiplib.gmle_version().decode().startswith("gaussmle-hip")
    assert [hiplib.gmle_n_params(m) for m in range(4)] == [4, 5, 6, 5]
    assert hiplib.gmle_n_params(9) < 0
    assert hiplib.gmle_max_box() == 32
    # large boxes: 8 or 16 lanes up to 24 px, 16 lanes above
    assert hiplib.gmle_lanes_supported(21, 8) == 1 and hiplib.gmle_lanes_supported(21, 16) == 1
    assert hiplib.gmle_lanes_supported(21, 4) == 0 and hiplib.gmle_lanes_supported(25, 8) == 0
    assert hiplib.gmle_lanes_supported(32, 16) == 1 and hiplib.gmle_lanes_supported(33, 16) == 0
    assert hiplib.gmle_default_lanes(0, 0, 21, 0, 10_000_000) == 8
    assert hiplib.gmle_default_lanes(3, 1, 21, 0, 10_000_000) == 16
    assert hiplib.gmle_default_lanes(1, 1, 28, 0, 10_000_000) == 16
    assert hiplib.gmle_lanes_supported(11, 1) == 1
    assert hiplib.gmle_lanes_supported(15, 1) == 0
    assert hiplib.gmle_lanes_supported(15, 4) == 1
    assert hiplib.gmle_lanes_supported(2, 1) == 0
    assert hiplib.gmle_lanes_supported(16, 8) == 1 and hiplib.gmle_lanes_supported(16, 16) == 1
    assert hiplib.gmle_lanes_supported(11, 32) == 0
    assert hiplib.gmle_default_lanes(0, 0, 11, 0, 10_000_000) == 1
    assert hiplib.gmle_default_lanes(0, 0, 11, 0, 100) == 16
    assert hiplib.gmle_default_lanes(0, 0, 11, 0, 10_000) == 4
    assert hiplib.gmle_default_lanes(0, 0, 11, 0, 30_000) == 2
    assert hiplib.gmle_device_count() >= 0


@pytest.mark.parametrize("model", [0, 1, 2, 3])
@pytest.mark.parametrize("box", [7, 11, 15])
def test_default_constraints_match_reference(hiplib, model, box):
    from gaussmle_amd import models
    P = hiplib.gmle_n_params(model)
    lo, up, ms = (np.zeros(P, np.float32) for _ in range(3))
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    assert hiplib.gmle_default_constraints(model, box, f(lo), f(up), f(ms)) == 0
    psf = [models.GaussianXYNB(0.13), models.GaussianXYNBS(), models.GaussianXYNBSXSY(),
           models.AstigmaticXYZNB(0.13, 0.13, 0.05, -0.05, 0.01, -0.01, 0.2, 0.5)][model]
    c = models.default_constraints(psf, box)
    np.testing.assert_array_equal(lo, c.lower)
    np.testing.assert_array_equal(up, c.upper)
    np.testing.assert_array_equal(ms, c.max_step)
    # src/constraints.jl:16-47 literal values
    assert lo[0] == -2 and up[0] == box + 2 and ms[0] == 1
    assert math.isinf(up[2 if model != 3 else 3])


def _call_fit(hiplib, data, box, n, camera=0, var=None, ny=0, nx=0, xc=None, yc=None, model=0, iterations=20,
              lanes=0):
    from gaussmle_amd import _lib
    P = hiplib.gmle_n_params(model) if 0 <= model <= 3 else 4
    psf = np.array([1.3, 1.3, 0.05, 0.05, 0.3, 0.3, 50, 100], np.float32)
    lo = np.full(P, -1e30, np.float32)
    up = np.full(P, 1e30, np.float32)
    ms = np.full(P, 1e30, np.float32)
    out = [np.zeros(max(n, 1) * P, np.float32), np.zeros(max(n, 1) * P, np.float32), np.zeros(max(n, 1), np.float32)]
    return hiplib.gmle_fit(model, _lib.fptr(psf), box, n, _lib.fptr(data), camera, _lib.fptr(var), ny, nx,
                           _lib.iptr(xc), _lib.iptr(yc), _lib.fptr(lo), _lib.fptr(up), _lib.fptr(ms), iterations, 0,
                           lanes, *map(_lib.fptr, out), None, None)


def test_fit_rejects_bad_input_before_any_device_work(hiplib):
    from gaussmle_amd import _lib
    d = np.ones(11 * 11 * 4, np.float32)
    assert _call_fit(hiplib, d, 11, 0) == _lib.GMLE_ERR_ARG                 # empty
    assert _call_fit(hiplib, d, 2, 4) == _lib.GMLE_ERR_ARG                  # box < 3
    assert _call_fit(hiplib, np.ones(33 * 33, np.float32), 33, 1) == _lib.GMLE_ERR_UNSUPPORTED
    assert _call_fit(hiplib, d, 11, 4, model=7) == _lib.GMLE_ERR_ARG
    assert _call_fit(hiplib, d, 11, 4, iterations=-1) == _lib.GMLE_ERR_ARG
    assert _call_fit(hiplib, d, 11, 4, lanes=3) == _lib.GMLE_ERR_UNSUPPORTED
    bad = d.copy()
    bad[77] = np.nan
    assert _call_fit(hiplib, bad, 11, 4) == _lib.GMLE_ERR_DATA
    bad[77] = np.inf
    assert _call_fit(hiplib, bad, 11, 4) == _lib.GMLE_ERR_DATA
    assert "NaN or Inf" in _lib.last_error()
    var = np.ones(20 * 20, np.float32)
    xc = np.array([1, 1, 10, 1], np.int32)   # ROI 3 ends at column 20: fits
    yc = np.array([1, 1, 1, 11], np.int32)   # ROI 4 ends at row 21: outside
    assert _call_fit(hiplib, d, 11, 4, camera=1, var=var, ny=20, nx=20, xc=xc, yc=yc) == _lib.GMLE_ERR_BOUNDS
    assert "outside" in _lib.last_error()
    assert _call_fit(hiplib, d, 11, 4, camera=1, var=None, ny=20, nx=20, xc=xc, yc=yc) == _lib.GMLE_ERR_ARG


@pytest.mark.skipif("__import__('gaussmle_amd')._lib.device_count() > 0")
def test_fit_without_device_reports_it(hiplib):
    from gaussmle_amd import _lib
    d = np.ones(11 * 11 * 4, np.float32)
    assert _call_fit(hiplib, d, 11, 4) == _lib.GMLE_ERR_NODEVICE


CSRC = ROOT / "gaussmle.jl_amd" / "csrc"
_SPEC = r"\(\s*(\w+)\s*,\s*(\w+)\s*,\s*(\w+)\s*,\s*(\d+)\s*,\s*(\d+)\s*\)"


def _specs(kind: str, files):
    out = []
    for f in files:
        for m in re.finditer(rf"^GMLE_{kind}_(FIT|KERNEL){_SPEC}", f.read_text(), flags=re.M):
            out.append((m.group(1),) + m.groups()[1:])
    return out


def test_kernels_built_apart_are_instantiated_once():
    """A finder declares some specialisations extern (GMLE_EXTERN_FIT /
    GMLE_EXTERN_KERNEL); each must be instantiated, with the same macro kind,
    in exactly one of the inst_sched_*.hip units, and each of those units has
    its own compiler flags in build.py."""
    sched = sorted(CSRC.glob("inst_sched_*.hip"))
    finders = [f for f in sorted(CSRC.glob("inst_*.hip")) if f not in sched]
    externs = _specs("EXTERN", finders)
    inst = _specs("INSTANTIATE", sched)
    assert externs, "no specialisation is built apart"
    assert sorted(externs) == sorted(inst)
    assert len(set(inst)) == len(inst)
    assert not _specs("INSTANTIATE", finders) and not _specs("EXTERN", sched)
    build = (ROOT / "gaussmle.jl_amd" / "build.py").read_text()
    for f in sched:
        assert f'"{f.name}": ["-mllvm", "-amdgpu-sched-strategy=' in build, f.name


def test_library_leaves_no_gmle_symbol_undefined(hiplib):
    """A kernel declared extern but built nowhere would be an undefined
    symbol of the shared library (-shared links it anyway)."""
    from gaussmle_amd import _lib
    out = subprocess.run(["nm", "-D", "-C", "--undefined-only", str(_lib.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    assert "gmle::" not in out, [line for line in out.splitlines() if "gmle::" in line]
