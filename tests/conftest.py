"""Shared test setup: import paths, the `gpu` marker, and common fixtures."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "gaussmle.jl_amd"
for p in (str(ROOT), str(PKG)):
    if p not in sys.path:
        sys.path.insert(0, p)


# torch ships its own HIP / HSA runtime with the same sonames as the system
# ROCm; loading torch first lets the HIP library bind to that one runtime.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import refcpu
    refcpu.build()
    return refcpu


@pytest.fixture(scope="session")
def hiplib():
    from gaussmle_amd import _lib
    return _lib.load()
