"""Shared fixtures for the test suite.

Tests marked ``gpu`` need an AMD GPU (run with ``-m gpu``); everything else
runs on the CPU.  The oracle (oracle/, a CPU restatement of the reference) is
the checker of the GPU path; it is built here if missing.
"""

from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "darcy-cpp_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

from darcy.fields import gaussian_random_field, two_phase_permeability  # noqa: E402


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X)")
    config.addinivalue_line("markers", "slow: long-running test")


def two_phase(n: int, seed: int = 0, hi: float = 12.0, lo: float = 3.0) -> np.ndarray:
    """The benchmark's thresholded-GRF medium (reference tests/conftest.py:11-13)."""
    return two_phase_permeability(gaussian_random_field(n, seed=seed), hi=hi, lo=lo)


def _ensure_oracle() -> None:
    lib = os.path.join(ROOT, "oracle", "lib", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)


@pytest.fixture(scope="session")
def oracle():
    _ensure_oracle()
    from pyoracle import Oracle

    return Oracle()


@pytest.fixture
def rng() -> np.random.Generator:
    return np.random.default_rng(20240724)


def rel_l2(x: np.ndarray, y: np.ndarray) -> float:
    return float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))
