"""Shared fixtures: the in-tree HIP library and the CPU oracle (test infrastructure)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle():
    import oracle_ctypes
    oracle_ctypes.lib()
    return oracle_ctypes


@pytest.fixture(scope="session")
def rt():
    import ray_tracing_amd
    ray_tracing_amd.lib()
    return ray_tracing_amd


@pytest.fixture(scope="session")
def device(rt):
    dev = rt.Device(0)
    yield dev
    dev.close()
