"""Shared fixtures: loaders for the product package and the oracle, and the
`gpu` marker (tests that need an MI355X)."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "blocked-linearized-coordinate_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_pkg():
    if "blco_amd" in sys.modules:
        return sys.modules["blco_amd"]
    spec = importlib.util.spec_from_file_location("blco_amd", os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["blco_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def B():
    return load_pkg()


@pytest.fixture(scope="session")
def O():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def gpu(B):
    """Skips when no GPU is visible.  GPU tests never import torch: its
    bundled HIP/HSA runtime would load beside the one libblco.so links."""
    if B.device_count() < 1:
        pytest.skip("no GPU visible")
    return B


def random_coo(rng, dims, nnz, unique=True):
    """Uniform random coordinates (mode-major) with distinct coordinates."""
    dims = [int(d) for d in dims]
    c = np.stack([rng.integers(0, d, nnz, dtype=np.int64) for d in dims]).astype(np.uint64)
    if unique and nnz:
        order = np.lexsort(c[::-1])
        cs = c[:, order]
        keep = np.ones(nnz, bool)
        keep[1:] = np.any(cs[:, 1:] != cs[:, :-1], axis=0)
        first = order[keep]
        first.sort()
        c = c[:, first]
    v = rng.random(c.shape[1]) + 0.5
    return c, v
