import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


@pytest.fixture(scope="session")
def tb():
    import __graft_entry__
    return __graft_entry__.package()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def backend(tb):
    be = tb.HIPBackend(0, pool_bytes=8 << 30)
    yield be
    be.close()


def rand_tropical(rng, shape, dtype, lo=0, hi=1 << 10, neg_frac=0.1):
    """Integer values in [lo, hi] with a fraction of tropical zeros."""
    x = rng.integers(lo, hi + 1, size=shape)
    x = x.astype(np.float32 if dtype == np.float32 else np.int32)
    mask = rng.random(shape) < neg_frac
    x[mask] = -np.inf if dtype == np.float32 else -(1 << 30)
    return np.asfortranarray(x)
