"""Shared test setup: import paths, the gpu marker, and helpers.

CPU tests (-m "not gpu") check the oracle against the reference's golden vectors, the host
logic and the C ABI's exports. GPU tests (-m gpu) run the HIP library through its C ABI and
compare it with the oracle.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "integer-schemes_amd", ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")
    config.addinivalue_line("markers", "slow: longer-running checks")


def normwise(a: np.ndarray, b: np.ndarray, den=None) -> np.ndarray:
    """Per-field normwise relative difference max|a-b| / max(max|b|, 1e-12) over the fields of
    axis 0. The 1e-12 floor is the reference oracle's (3D/src/bssn3d/oracle.py:187): a field
    whose exact right-hand side vanishes (e.g. At_xy of a 1D gauge wave) carries only round-off
    and must not be divided by it. ``den`` overrides max|b| (e.g. a full-field maximum when a
    and b are samples)."""
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    if den is None:
        den = np.max(np.abs(b), axis=1)
    den = np.maximum(np.asarray(den, dtype=np.float64), 1e-12)
    return np.max(np.abs(a - b), axis=1) / den


@pytest.fixture(scope="session")
def golden_points():
    return np.load(GOLDEN / "algebra_points.npz")


@pytest.fixture(scope="session")
def golden_grid():
    return np.load(GOLDEN / "grid_rhs.npz")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
