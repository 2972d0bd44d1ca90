"""Shared test fixtures.

Markers: ``gpu`` -- needs a real MI355X (run with ``-m gpu``); everything
else runs on the CPU.  The oracle (``oracle/``) is the checker for every GPU
test; it is never part of the product path.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

SCENES = ROOT / "pyro_amd" / "scenes"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: requires an AMD Instinct GPU (gfx950)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def pool_scene():
    import pyro_amd as pa

    return pa.Scene.load(str(SCENES / "pool.toml"))


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    if den == 0.0:
        return float(num)
    return float(num / den)
