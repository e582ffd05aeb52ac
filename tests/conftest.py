import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _load_pkg():
    if "eikonal_amd" in sys.modules:
        return sys.modules["eikonal_amd"]
    path = os.path.join(ROOT, "sci-solver_eikonal_amd", "eikonal_amd.py")
    spec = importlib.util.spec_from_file_location("eikonal_amd", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    sys.modules["eikonal_amd"] = mod
    return mod


@pytest.fixture(scope="session")
def E():
    mod = _load_pkg()
    mod.lib()  # fails loudly when the HIP library is not built
    return mod


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.lib()
    return oracle


def golden(name):
    return os.path.join(GOLDEN, name)


def center_vertex(n):
    c = n // 2
    return (c * (n + 1) + c) * (n + 1) + c


def lognormal_mult(nt, seed, sigma=0.5):
    """Per-tet metric multiplier 1/s^2 for a log-normal(0, sigma) speed s."""
    rng = np.random.default_rng(seed)
    s = np.exp(rng.normal(0.0, sigma, nt))
    return (1.0 / s ** 2).astype(np.float32)


def rel_err(T, ref, tau):
    """max |T - ref| / max(ref, tau), the accuracy measure used throughout."""
    T = np.asarray(T, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.max(np.abs(T - ref) / np.maximum(np.abs(ref), tau)))
