"""Shared fixtures.  `gpu` tests need an MI355X; everything else runs on the CPU."""
import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

PKG_NAME = "flux-interpolated-advection_amd"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X) and the built libfia.so")


def load_pkg():
    return importlib.import_module(PKG_NAME)


def load_scenes():
    return importlib.import_module(PKG_NAME + ".scenes")


@pytest.fixture(scope="session")
def fia():
    return load_pkg()


@pytest.fixture(scope="session")
def scenes():
    return load_scenes()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx(fia):
    if fia.device_count() < 1:
        pytest.skip("no GPU visible")
    c = fia.Context(0)
    yield c
    c.close()


def rel_linf(a, b):
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
