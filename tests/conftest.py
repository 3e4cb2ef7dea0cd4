import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "legendre-to-chebyshev_amd")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (gfx950)")


def load_l2c_amd():
    """Import the package directory (its name carries hyphens) as `l2c_amd`."""
    if "l2c_amd" in sys.modules:
        return sys.modules["l2c_amd"]
    spec = importlib.util.spec_from_file_location(
        "l2c_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["l2c_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def amd():
    return load_l2c_amd()


@pytest.fixture(scope="session")
def orc():
    from oracle import orc as o
    return o


def relerr(x, ref):
    """Relative max-norm error ||x - ref||_inf / ||ref||_inf."""
    import numpy as np
    den = float(np.max(np.abs(ref)))
    return float(np.max(np.abs(np.asarray(x) - np.asarray(ref)))) / (den if den > 0 else 1.0)


def announce(msg):
    print(f"[case] {msg}", flush=True)
