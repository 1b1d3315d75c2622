import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


@pytest.fixture
def rng():
    return np.random.default_rng(0)


# reference tests/conftest.py:166-210 fixtures (data only)
@pytest.fixture
def initial_conditions_positions():
    return np.array(
        [[1.0, 2.0], [3.0, 4.0], [5.0, 6.0], [7.0, 8.0],
         [9.0, 10.0], [11.0, 12.0], [13.0, 14.0], [15.0, 16.0]]
    )


@pytest.fixture
def jacobians_2x2():
    return np.array([[[2.0, 0.0], [0.0, 1.0]], [[1.5, 0.5], [0.5, 1.5]]], dtype=np.float64)


@pytest.fixture
def jacobians_3x3():
    return np.array([np.eye(3), [[1.0, 0.2, 0.0], [0.0, 1.5, 0.1], [0.0, 0.0, 1.2]]], dtype=np.float64)
