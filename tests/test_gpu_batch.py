"""GPU parity of batches of uneven length (l2c_create_batch): every transform
of the batch against the oracle's single transform of the same length
(src/C/leg2cheb.c:1299-1500, maxs = 64, M = 18).  Tolerance: relative
max-norm error <= 1e-14 * log2(n) per transform."""
import math

import numpy as np
import pytest

from conftest import announce, relerr

pytestmark = pytest.mark.gpu


def tol(n):
    return 1e-14 * math.log2(max(n, 2))


def check(amd, orc, bp, u, g, direction, which=None):
    worst = 0.0
    for t in (range(len(bp.lengths)) if which is None else which):
        n, o = int(bp.lengths[t]), int(bp.offsets[t])
        ref = orc.FMMPlan(n, 64, 18, direction).execute(u[o:o + n], direction)
        e = relerr(g[o:o + n], ref)
        worst = max(worst, e / tol(n))
        assert e <= tol(n), f"transform {t} (n={n}) error {e:.3e}"
    return worst


# the edges of every size class and of the direct/FMM switch
EDGES = [1, 2, 3, 63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 513, 1000, 1023, 1024,
         1025, 2047, 2049, 4095, 4097, 8191, 8192, 300, 64, 5000]


@pytest.mark.parametrize("tune", [False, True])
@pytest.mark.parametrize("direction", [0, 1])
def test_batch_edges(amd, orc, tune, direction):
    announce(f"batch edges tune={tune} dir={direction}")
    bp = amd.BatchPlan(EDGES, 18, direction, tune=tune)
    print("  classes:", bp.classes())
    u = amd.packed_coefficients(bp.lengths, seed=11 + direction, decay=1.0)
    g = bp.execute(u, direction)
    w = check(amd, orc, bp, u, g, direction)
    print(f"  worst error / tol = {w:.3e}")


def test_batch_repeat_and_graph(amd, orc):
    """Eager and graph-replayed executes agree; a replay sees new input data
    in the same buffers; C2L of the L2C output returns the input."""
    announce("batch graph replay")
    lengths = amd.log_uniform_lengths(300, seed=5)
    bp = amd.BatchPlan(lengths, 18, amd.BOTH, tune=False)
    u1 = amd.packed_coefficients(lengths, seed=1)
    u2 = amd.packed_coefficients(lengths, seed=2)
    din, dout = amd.DeviceBuffer(bp.total), amd.DeviceBuffer(bp.total)
    din.upload(u1)
    bp.execute_device(din.ptr, dout.ptr, 0)
    bp.synchronize()
    eager1 = dout.download()
    bp.graph(True)
    for rep in range(3):
        din.upload(u1 if rep % 2 == 0 else u2)
        bp.execute_device(din.ptr, dout.ptr, 0)
        bp.synchronize()
        got = dout.download()
        if rep % 2 == 0:
            assert np.array_equal(got, eager1), f"replay {rep} differs from the eager run"
        else:
            check(amd, orc, bp, u2, got, 0, which=range(0, 300, 7))
    bp.graph(False)
    back = bp.execute(bp.execute(u1, 0), 1)
    assert float(np.max(np.abs(back - u1))) < 1e-10


def test_batch_log_uniform_4096(amd, orc):
    """4096 transforms, n = floor(exp(U(ln 64, ln 8193))), inputs U(-1,1)/(k+1):
    every transform with n <= 256 and a seeded sample of the rest against the
    oracle, and the whole batch against per-transform plans on the GPU."""
    announce("batch log-uniform 4096")
    lengths = amd.log_uniform_lengths(4096, seed=2024)
    bp = amd.BatchPlan(lengths, 18, amd.L2C, tune=True)
    print("  classes:", bp.classes())
    u = amd.packed_coefficients(lengths, seed=77)
    g = bp.execute(u, 0)
    rng = np.random.default_rng(3)
    small = [t for t in range(4096) if lengths[t] <= 256]
    sample = sorted(set(small) | set(rng.choice(4096, 200, replace=False).tolist()))
    w = check(amd, orc, bp, u, g, 0, which=sample)
    print(f"  {len(sample)} transforms checked, worst error / tol = {w:.3e}")


def test_tune_cache_round_trip(amd, orc, tmp_path):
    """The tuner's choices go to the cache file on the first plan and come
    back from it on the second (no timing): same method, leaf size and kernel
    variant per class, and bit-identical results."""
    import time
    path = tmp_path / "l2c_tune.txt"
    amd.tune_cache(path)
    try:
        lengths = amd.log_uniform_lengths(600, seed=31)
        announce(f"tune cache {path}")
        t0 = time.perf_counter()
        bp1 = amd.BatchPlan(lengths, 18, amd.L2C, tune=True)
        t1 = time.perf_counter()
        bp2 = amd.BatchPlan(lengths, 18, amd.L2C, tune=True)
        t2 = time.perf_counter()
        c1, c2 = bp1.classes(), bp2.classes()
        print(f"  plan time cold {1e3 * (t1 - t0):.1f} ms, warm cache {1e3 * (t2 - t1):.1f} ms")
        assert not any(c["cached"] for c in c1) and all(c["cached"] for c in c2)
        key = lambda c: (c["n_max"], c["method"], c["s"], c["variant"])
        assert sorted(map(key, c1)) == sorted(map(key, c2))
        assert len(path.read_text().splitlines()) == len(c1)
        u = amd.packed_coefficients(lengths, seed=8)
        g1, g2 = bp1.execute(u, 0), bp2.execute(u, 0)
        assert np.array_equal(g1, g2)
        o = orc.FMMPlan(int(lengths[0]), 64, 18, 0).execute(u[:lengths[0]], 0)
        assert relerr(g1[:lengths[0]], o) <= 1e-14 * 13
    finally:
        amd.tune_cache(None)
