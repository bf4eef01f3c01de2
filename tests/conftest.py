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
    """Build when a library is missing or was built from other sources than
    the tree's (the stamp holds the source hash of the last full build)."""
    from psba_amd import build
    if not build.is_current():
        build.build_all()


_ensure_built()


def pytest_report_header(config):
    from psba_amd import build
    return "psba sources %s, libraries built from %s" % (build.source_hash(), build.built_hash())


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
