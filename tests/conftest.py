import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP library")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


_ORACLE_CACHE = {}


@pytest.fixture(scope="session")
def oracle_ctx():
    """Factory returning cached oracle contexts (one per parameter set)."""
    import oracle as O

    def get(**kw):
        key = tuple(sorted(kw.items()))
        if key not in _ORACLE_CACHE:
            kw.setdefault("num_threads", min(16, os.cpu_count() or 1))
            _ORACLE_CACHE[key] = O.Oracle(**kw)
        return _ORACLE_CACHE[key]

    return get


@pytest.fixture(scope="session")
def gpu():
    """The product on the GPU; fails (never skips) when no HIP device is visible."""
    import torch
    import upcgen_amd as U
    assert torch.cuda.is_available(), "gpu test run without a visible HIP device"
    U.load()
    return U


_GPU_CACHE = {}


@pytest.fixture(scope="session")
def gpu_ctx(gpu):
    def get(**kw):
        key = tuple(sorted(kw.items()))
        if key not in _GPU_CACHE:
            _GPU_CACHE[key] = gpu.Context(**kw)
        return _GPU_CACHE[key]

    return get
