"""GPU parity of batched strided lines (l2c_execute_lines_device), the path of
a spectral field transformed along each axis (the reference's execute with a
stride, src/C/leg2cheb.c:1299-1500, and its 2-D driver :1267-1297, applied to
every line).  Lines of the direct path (N <= 4 maxs) run as dense MFMA
products; every line is checked against the oracle's single transform.
Tolerance: relative max-norm error <= 1e-14 * log2(N) per line."""
import math

import numpy as np
import pytest

from conftest import announce, relerr

pytestmark = pytest.mark.gpu


def tol(n):
    return 1e-14 * math.log2(max(n, 2))


def along_axis(amd, x, axis, direction, maxs=64):
    """Transform every line of the C-ordered array x along `axis` in one
    batched call on the device."""
    shape = x.shape
    N = shape[axis]
    es = int(np.prod(shape[axis + 1:]))          # element stride along the axis
    inner = es                                   # lines inside one outer slab
    outer = int(np.prod(shape[:axis]))
    nlines = outer * inner
    plan = amd.FMMPlan(N, maxs, 18, amd.BOTH)
    din, dout = amd.DeviceBuffer(x.size), amd.DeviceBuffer(x.size)
    din.upload(x.ravel())
    # line l = (o, i): offset o * N * es + i
    plan.execute_lines_device(din.ptr, dout.ptr, direction, nlines, inner, 1, N * es, es)
    plan.synchronize()
    return dout.download().reshape(shape)


def oracle_along_axis(orc, x, axis, direction, maxs=64):
    N = x.shape[axis]
    op = orc.FMMPlan(N, maxs, 18, orc.BOTH)
    xm = np.moveaxis(x, axis, -1)
    out = np.empty_like(xm)
    for idx in np.ndindex(xm.shape[:-1]):
        out[idx] = op.execute(np.ascontiguousarray(xm[idx]), direction)
    return np.moveaxis(out, -1, axis)


def per_line_error(g, o, axis):
    gm, om = np.moveaxis(g, axis, -1), np.moveaxis(o, axis, -1)
    return max(relerr(gm[idx], om[idx]) for idx in np.ndindex(gm.shape[:-1]))


@pytest.mark.parametrize("shape", [(40, 48, 64), (256, 7, 9), (5, 256, 33), (3, 2, 256)])
@pytest.mark.parametrize("direction", [0, 1])
def test_field_axes(amd, orc, shape, direction):
    announce(f"field {shape} dir={direction}")
    rng = np.random.default_rng(sum(shape) + direction)
    i, j, k = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    x = rng.uniform(-1, 1, shape) / ((i + 1.0) * (j + 1.0) * (k + 1.0))
    for axis in range(3):
        g = along_axis(amd, x, axis, direction)
        o = oracle_along_axis(orc, x, axis, direction)
        e = per_line_error(g, o, axis)
        print(f"  axis {axis}: worst per-line error {e:.3e}")
        assert e <= tol(shape[axis])


@pytest.mark.parametrize("N", [1, 2, 3, 17, 33, 200, 255, 256])
def test_dense_lengths(amd, orc, N):
    """Every padding case of the dense parity blocks (row tiles 1..8, odd and
    even tile counts) on contiguous and on strided lines."""
    announce(f"dense lines N={N}")
    rng = np.random.default_rng(N)
    x = rng.uniform(-1, 1, (70, N))
    for d in (0, 1):
        g1 = along_axis(amd, x, 1, d)
        o1 = oracle_along_axis(orc, x, 1, d)
        assert per_line_error(g1, o1, 1) <= tol(N)
        xt = np.ascontiguousarray(x.T)
        g0 = along_axis(amd, xt, 0, d)
        assert per_line_error(g0, np.ascontiguousarray(o1.T), 0) <= tol(N)


@pytest.mark.parametrize("N", [257, 300, 511, 512, 513, 1000, 1024, 2047, 2048, 4095, 4096])
def test_gemm_lengths(amd, orc, N):
    """Direct-path plans above N = 256 (maxs = N/4 rounded up, so L = 0): the
    tiled dense product, with row tiles partly past NPP (N = 300, 513, 1000,
    2047, 4095), a first chunk starting mid-row-tile, and ragged line tiles
    (70 lines), on contiguous and on strided lines, both directions."""
    maxs = (N + 3) // 4
    announce(f"gemm lines N={N} maxs={maxs}")
    rng = np.random.default_rng(N + 7)
    x = rng.uniform(-1, 1, (70, N)) / np.sqrt(np.arange(1, N + 1))
    for d in (0, 1):
        g1 = along_axis(amd, x, 1, d, maxs=maxs)
        o1 = oracle_along_axis(orc, x, 1, d, maxs=maxs)
        assert per_line_error(g1, o1, 1) <= tol(N), d
        xt = np.ascontiguousarray(x.T)
        g0 = along_axis(amd, xt, 0, d, maxs=maxs)
        assert per_line_error(g0, np.ascontiguousarray(o1.T), 0) <= tol(N), d


@pytest.mark.parametrize("N", [33, 256])
def test_dense_many_lines(amd, orc, N):
    """20,000 lines of N = 33 and 256 in one call: above 16,384 lines the
    one-tile dense kernel takes them (below, the tiled one); contiguous and
    strided, both directions, sampled lines against the oracle."""
    announce(f"dense lines N={N} x 20000")
    x = np.random.default_rng(N + 11).uniform(-1, 1, (20000, N))
    op = orc.FMMPlan(N, 64, 18, orc.BOTH)
    rows = [0, 1, 31, 32, 9999, 19968, 19999]
    for d in (0, 1):
        g1 = along_axis(amd, x, 1, d)
        for r in rows:
            assert relerr(g1[r], op.execute(np.ascontiguousarray(x[r]), d)) <= tol(N), (d, r)
        xt = np.ascontiguousarray(x.T)
        g0 = along_axis(amd, xt, 0, d)
        for r in rows:
            assert relerr(g0[:, r], op.execute(np.ascontiguousarray(x[r]), d)) <= tol(N), (d, r)


def test_field_round_trip(amd):
    """L2C along the three axes, then C2L back: the input returns
    (src/bin/l2c.c:342, 387 thresholds)."""
    announce("field round trip 64^3")
    rng = np.random.default_rng(9)
    n = 64
    i, j, k = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    x = rng.uniform(-1, 1, (n, n, n)) / ((i + 1.0) * (j + 1.0) * (k + 1.0))
    y = x
    for axis in (2, 1, 0):
        y = along_axis(amd, y, axis, 0)
    for axis in (0, 1, 2):
        y = along_axis(amd, y, axis, 1)
    assert float(np.max(np.abs(y - x))) < 1e-10


@pytest.mark.parametrize("N,maxs", [(512, 64), (600, 256), (200, 64)])
def test_more_lines_than_one_launch(amd, orc, N, maxs):
    """65,541 lines along the middle axis of a (9363, N, 7) array in one call:
    the FMM kernels (N = 512) take at most 65,535 lines per launch, so the
    call runs two, the second starting inside a group of n0 = 7 lines
    (65,535 = 7 * 9362 + 1); the dense products of direct plans (N = 600 on
    the tiled kernel, N = 200 on the one-tile kernel) take them all in one.
    The lines either side of the boundary, the first and the last against
    the oracle."""
    P, Q = 9363, 7
    announce(f"lines across launches N={N} maxs={maxs} nlines={P * Q}")
    x = np.random.default_rng(N).random((P, N, Q)) - 0.5
    op = orc.FMMPlan(N, maxs, 18, orc.BOTH)
    for d in (0, 1):
        g = along_axis(amd, x, 1, d, maxs=maxs)
        for line in (0, 65533, 65534, 65535, 65536, P * Q - 1):
            p, q = divmod(line, Q)
            ref = op.execute(np.ascontiguousarray(x[p, :, q]), d)
            assert relerr(g[p, :, q], ref) <= tol(N), (d, line)
        del g


@pytest.mark.parametrize("direction", [0, 1])
def test_direct_rows_lines_over_grid_z(amd, orc, monkeypatch, direction):
    """Direct plans above N = 4096 (no dense parity blocks) run their lines
    on the one-row-per-thread kernel, whose grid holds at most 65,535 lines
    per grid row and takes the rest in grid.z.  L2C_DIRECT_LINES_Y = 5 at
    plan creation cuts the rows to five lines, so 13 lines of N = 4100
    (maxs = 1025: L = 0) span three grid rows; every line against the
    oracle."""
    N, P, Q = 4100, 13, 1
    announce(f"direct rows over grid.z N={N} nlines={P} dir={direction}")
    monkeypatch.setenv("L2C_DIRECT_LINES_Y", "5")
    plan = amd.FMMPlan(N, 1025, 18, amd.BOTH)
    monkeypatch.delenv("L2C_DIRECT_LINES_Y")
    assert plan.info()["L"] == 0
    x = np.random.default_rng(41 + direction).uniform(-1, 1, (P, N))
    din, dout = amd.DeviceBuffer(x.size), amd.DeviceBuffer(x.size)
    din.upload(x.ravel())
    plan.execute_lines_device(din.ptr, dout.ptr, direction, P, P, N, 0, 1)
    plan.synchronize()
    g = dout.download().reshape(P, N)
    op = orc.FMMPlan(N, 1025, 18, orc.BOTH)
    for line in range(P):
        assert relerr(g[line], op.execute(np.ascontiguousarray(x[line]), direction)) <= tol(N), line


def test_field_256_cube_with_vectors(amd, orc):
    """The 256^3 field of the benchmark (c_ijk = U(-1,1)/((i+1)(j+1)(k+1)))
    through L2C along axes 2, 1, 0 (65,536 strided lines per axis, one call
    each) and C2L back along 0, 1, 2, while 64 vectors of N = 2^18 run L2C on
    a batch plan's own streams at the same time: 64 sampled lines of every
    L2C pass against the oracle's single transform, the round trip, and the
    first and last vector against the oracle."""
    n, nv, nvec = 256, 64, 1 << 18
    announce(f"field {n}^3 round trip with {nv} x 2^18 vectors")
    i, j, k = np.meshgrid(*[np.arange(n, dtype=np.float64)] * 3, indexing="ij")
    x = (amd.splitmix_uniform(n ** 3, 3).reshape(n, n, n) / ((i + 1) * (j + 1) * (k + 1)))
    del i, j, k
    fp = amd.FMMPlan(n, 64, 18, amd.BOTH)
    op = orc.FMMPlan(n, 64, 18, orc.BOTH)
    a, b = amd.DeviceBuffer(x.size), amd.DeviceBuffer(x.size)
    a.upload(x.ravel())
    bp = amd.BatchPlan([nvec] * nv, 18, amd.L2C, tune=False)
    uv = amd.packed_coefficients([nvec] * nv, seed=4)
    vin, vout = amd.DeviceBuffer(nv * nvec), amd.DeviceBuffer(nv * nvec)
    vin.upload(uv)
    maps = {2: (n * n, n, 0, 1), 1: (n, 1, n * n, n), 0: (n * n, 1, 0, n * n)}
    bp.execute_device(vin.ptr, vout.ptr, 0)  # the batch streams run beside the field passes
    rng = np.random.default_rng(12)
    src, dst, cur = a, b, x
    for ax in (2, 1, 0):
        fp.execute_lines_device(src.ptr, dst.ptr, 0, n * n, *maps[ax])
        fp.synchronize()
        g = dst.download().reshape(n, n, n)
        gm, cm = np.moveaxis(g, ax, -1), np.moveaxis(cur, ax, -1)
        for _ in range(64):
            idx = tuple(int(v) for v in rng.integers(0, n, 2))
            ref = op.execute(np.ascontiguousarray(cm[idx]), 0)
            assert relerr(gm[idx], ref) <= tol(n), (ax, idx)
        src, dst, cur = dst, src, g
    for ax in (0, 1, 2):
        fp.execute_lines_device(src.ptr, dst.ptr, 1, n * n, *maps[ax])
        src, dst = dst, src
    fp.synchronize()
    bp.synchronize()
    back = src.download().reshape(n, n, n)
    assert float(np.max(np.abs(back - x))) < 1e-10
    gv = vout.download()
    ov = orc.FMMPlan(nvec, 64, 18, 0)
    for t in (0, nv - 1):
        ref = ov.execute(uv[t * nvec:(t + 1) * nvec], 0)
        assert relerr(gv[t * nvec:(t + 1) * nvec], ref) <= tol(nvec), t
