"""Shared test set-up. GPU tests carry @pytest.mark.gpu; the CPU suite runs with -m "not gpu"."""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
