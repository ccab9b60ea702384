"""Shared fixtures.  Marker `gpu`: needs an MI355X (run with -m gpu)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs an AMD MI355X GPU")


def _ensure_oracle():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()
    return pyoracle


def _ensure_lib():
    import stencil3d_amd
    if not os.path.exists(stencil3d_amd.LIB_PATH):
        from stencil3d_amd import build
        build.build()
    return stencil3d_amd


@pytest.fixture(scope="session")
def orc():
    return _ensure_oracle()


@pytest.fixture(scope="session")
def sth():
    return _ensure_lib()


@pytest.fixture(scope="session")
def gpu(sth):
    n = sth.device_count()
    if n < 1:
        pytest.fail("no GPU visible to libstencil_hip")
    return sth
