import importlib.util
import os
import sys

import pytest

# FMM plans of N <= 8192 run single transforms as direct sums by default
# (plan field direct_max_n); the suite's FMM cases at those sizes are there
# to test the FMM kernels, so plans made under it keep the FMM unless a test
# sets L2C_DIRECT_MAXN itself (tests/test_gpu_small_n.py)
os.environ.setdefault("L2C_DIRECT_MAXN", "0")

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


# ---- Table 1 of the reference's paper (FMM_paper.pdf p.17) ----------------
def table1_published():
    """{N: {(direction, lagrange): published absolute error}} from
    tests/golden/table1_published.json."""
    import json
    with open(os.path.join(ROOT, "tests", "golden", "table1_published.json")) as f:
        t = json.load(f)
    out = {}
    for r in t["rows"]:
        out[int(r[0])] = {(0, 0): r[1], (0, 1): r[3], (1, 0): r[5], (1, 1): r[7]}
    return out


def table1_inputs(N):
    """The harness's inputs u_i = rand()/RAND_MAX after srand(1)
    (src/multiprec/C++/L2Cprec.cpp:78-92, random=1, m=0), drawn from this
    host's C library; the paper's own run used another platform's rand(), so
    the values differ and only the error envelope is comparable."""
    import ctypes
    import numpy as np
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    return np.array([libc.rand() for _ in range(N)], dtype=np.float64) / 2147483647.0


def table1_error(out, ref):
    """(absolute max error, ulp of max(ref)) as the harness measures them
    (src/multiprec/C++/L2Cprec.cpp:123-127, 140-158)."""
    import numpy as np
    mx = float(np.max(ref))
    ulp = float(np.nextafter(mx, 1e8) - mx)
    return float(np.max(np.abs(np.asarray(out) - ref))), ulp


# A transform passes when its absolute error is within the published one plus
# four ulps of ||z*||_inf: the inputs differ from the paper's (see
# table1_inputs) and rounding-order differences move the error by an ulp or two.
TABLE1_SLACK_ULPS = 4.0
