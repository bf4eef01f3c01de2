"""Shared test setup.

Marker `gpu`: the test needs an MI355X (run with `-m gpu`).  Everything else
runs on the CPU: the oracle against fixtures and closed forms, the scene
generator, host-side indexing and the C-ABI's symbols/error paths.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP device)")


def _ensure_built():
    lib = os.path.join(ROOT, "psba_amd", "libpsba.so")
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    dropin = os.path.join(ROOT, "tests", "dropin", "main_dropin")
    if not (os.path.exists(lib) and os.path.exists(orc) and os.path.exists(dropin)):
        from psba_amd import build
        build.build_oracle()
        build.build_lib()
        build.build_dropin()


_ensure_built()


@pytest.fixture(scope="session")
def P():
    import psba_amd
    return psba_amd


@pytest.fixture(scope="session")
def O():
    from oracle import binding
    return binding


@pytest.fixture(scope="session")
def gpu(P):
    if P.device_count() < 1:
        pytest.skip("no HIP device")
    return True


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)
