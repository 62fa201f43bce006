"""Shared pytest configuration: the `gpu` marker and helpers to reach the in-tree build."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def hip():
    """The HIP C-ABI binding with device 0 current; fails (never skips) when no device exists,
    so a GPU run cannot pass on a silent fallback."""
    from ptask_amd import hip as H
    n = H.device_count()
    assert n >= 1, "no HIP device visible to a test marked gpu"
    H.set_device(0)
    return H
