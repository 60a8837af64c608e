import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as entry  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs through the C ABI")


@pytest.fixture(scope="session")
def orc():
    return entry.load_oracle()


@pytest.fixture(scope="session")
def xr():
    return entry.load_package()


@pytest.fixture(scope="session")
def ctx(xr):
    c = xr.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def ellipsoid_phantom(orc, xr, N, vs, seed, count=6, value_hi=0.05):
    params = xr.ellipsoid_params(seed, count, fov_radius=0.5 * N * vs, centre_frac=0.5, axis_lo=0.1,
                                 axis_hi=0.35, value_lo=0.005, value_hi=value_hi)
    return orc.phantom_ellipsoids(N, vs, params), params
