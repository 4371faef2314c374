"""Shared fixtures. GPU tests are marked ``gpu``; everything else runs on CPU."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "bcm-problems_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


@pytest.fixture(scope="session")
def B():
    import bcmhip

    bcmhip.lib()
    return bcmhip


@pytest.fixture(scope="session")
def O():
    import oracle

    oracle.orc()
    return oracle


@pytest.fixture(scope="session")
def ctx(B):
    return B.Context()


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel(order="F")
    b = np.asarray(b, dtype=np.float64).ravel(order="F")
    assert a.shape == b.shape, (a.shape, b.shape)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)
