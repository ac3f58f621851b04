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


def normwise(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-field normwise relative difference max|a-b| / max|b| over axis 0 fields."""
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    den = np.max(np.abs(b), axis=1)
    den = np.where(den == 0.0, 1.0, den)
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
