"""GPU parity of the length-N DCT-II / DCT-III (l2c_create_dct,
l2c_execute_dct_device): the FFTW REDFT10 / REDFT01 transforms the reference
times as its yardstick (src/bin/l2c.c:414-463), and the steps between values
at the Chebyshev-Gauss points and Chebyshev coefficients around L2C / C2L.

Every result is checked against the oracle's long-double DCT (a length-4N
FFT of the symmetric extension, pinned to the O(N^2) definition on the CPU);
the 2^25 case, too large for it, against scipy's pocketfft.
Tolerance: relative max-norm error <= 1e-15 * log2(2N) per vector."""
import math

import numpy as np
import pytest

from conftest import announce, relerr

pytestmark = pytest.mark.gpu


def tol(N):
    return 1e-15 * math.log2(2 * N)


def run(amd, x, N, count, kind, inplace=False):
    plan = amd.DCTPlan(N, count)
    din = amd.DeviceBuffer(N * count)
    dout = din if inplace else amd.DeviceBuffer(N * count)
    din.upload(x.ravel())
    plan.execute_device(din.ptr, dout.ptr, kind)
    return dout.download().reshape(count, N)


def check(amd, orc, N, count, kind, seed, inplace=False):
    announce(f"dct kind={kind} N={N} count={count} inplace={inplace}")
    x = np.random.default_rng(seed).uniform(-1, 1, (count, N))
    y = run(amd, x, N, count, kind, inplace)
    ref = orc.redft(x, kind)
    errs = [relerr(y[r], ref[r]) for r in range(count)]
    assert max(errs) < tol(N), (max(errs), tol(N))
    return max(errs)


@pytest.mark.parametrize("kind", [2, 3])
@pytest.mark.parametrize("logN", list(range(0, 14)))
def test_one_pass_sizes(amd, orc, kind, logN):
    """N = 1 ... 4096: whole vectors in one work-group's LDS (several per
    work-group below N = 4096), with a count that leaves the last
    work-group partly empty; N = 8192 is the first two-pass size."""
    N = 1 << logN
    check(amd, orc, N, 1, kind, 10 * logN + kind)
    check(amd, orc, N, 37, kind, 10 * logN + kind + 1)


@pytest.mark.parametrize("kind", [2, 3])
@pytest.mark.parametrize("logN", [14, 15, 16, 17, 18, 19, 20, 21])
def test_multi_pass_sizes(amd, orc, kind, logN):
    """N = 2^14 ... 2^21: two HBM passes of lengths 64 ... 1024."""
    N = 1 << logN
    check(amd, orc, N, 1, kind, 100 + logN + kind)
    if logN <= 17:
        check(amd, orc, N, 3, kind, 200 + logN + kind)


@pytest.mark.parametrize("kind", [2, 3])
@pytest.mark.parametrize("logN,count", [(13, 1024), (16, 64), (17, 32), (20, 4)])
def test_wide_work_groups(amd, orc, kind, logN, count):
    """Jobs of at least 2^21 complex points run the passes with 4096 elements
    per work-group (pass lengths 64 ... 1024)."""
    check(amd, orc, 1 << logN, count, kind, 300 + logN + kind)


@pytest.mark.parametrize("kind", [2, 3])
def test_more_vectors_than_one_launch(amd, orc, kind):
    """65,537 vectors of N = 8192 (4.3 GB each way): the two-pass kernels take
    at most 65,535 vectors per launch, so the call runs two chunks; the
    vectors either side of the boundary and the last one against the oracle."""
    N, count = 8192, 65537
    announce(f"dct kind={kind} N={N} count={count} (two launches per pass)")
    x = np.random.default_rng(500 + kind).random((count, N)) - 0.5
    y = run(amd, x, N, count, kind)
    for r in (0, 65534, 65535, 65536):
        assert relerr(y[r], orc.redft(x[r], kind)) < tol(N), r
    del x, y


@pytest.mark.parametrize("kind", [2, 3])
def test_largest_pass_lengths(amd, kind):
    """N = 2^23 (the reference's largest timed size: two passes of 2048),
    2^24 (three passes of unequal lengths 7/8/8) and 2^25 (three passes of
    256, the largest plan) against scipy's pocketfft."""
    sf = pytest.importorskip("scipy.fft")
    for logN in (23, 24, 25):
        N = 1 << logN
        announce(f"dct kind={kind} N=2^{logN}")
        x = np.random.default_rng(logN + kind).uniform(-1, 1, N)
        y = run(amd, x, N, 1, kind)[0]
        assert relerr(y, sf.dct(x, type=kind)) < tol(N)


@pytest.mark.parametrize("N", [256, 1 << 16])
def test_in_place(amd, orc, N):
    for kind in (2, 3):
        check(amd, orc, N, 2, kind, N + kind, inplace=True)


@pytest.mark.parametrize("N", [64, 1 << 15])
def test_round_trip_and_host_entry(amd, N):
    """REDFT01(REDFT10(x)) = 2N x through the host-array entry point, which
    also equals the device entry point bit for bit."""
    announce(f"dct round trip N={N}")
    x = np.random.default_rng(N).uniform(-1, 1, 2 * N)
    plan = amd.DCTPlan(N, 2)
    y = plan.execute(x, 2)
    assert np.array_equal(y.reshape(2, N), run(amd, x.reshape(2, N), N, 2, 2))
    assert relerr(plan.execute(y, 3) / (2 * N), x) < tol(N)


def test_values_pipeline(amd):
    """Legendre coefficients -> L2C (FMM) -> Chebyshev coefficients -> DCT-III
    -> values at the Chebyshev-Gauss points equal the Legendre series summed
    directly there; values -> DCT-II -> C2L gives the coefficients back."""
    from numpy.polynomial import legendre
    N = 1 << 12
    announce(f"dct values pipeline N={N}")
    rng = np.random.default_rng(5)
    a = rng.uniform(-1, 1, N) / (1 + np.arange(N)) ** 2
    plan = amd.FMMPlan(N, 64, 18, amd.BOTH)
    dct = amd.DCTPlan(N)
    c = plan.execute(a, amd.L2C)
    f = amd.cheb_values(c, dct)
    xj = np.cos(np.pi * (2 * np.arange(N) + 1) / (2 * N))
    assert relerr(f, legendre.legval(xj, a)) < 1e-13
    back = plan.execute(amd.cheb_coefficients(f, dct), amd.C2L)
    assert relerr(back, a) < 1e-13


def test_errors(amd):
    lib = amd.lib()
    assert not lib.l2c_create_dct(12, 1)
    assert "power of two" in amd.last_error()
    plan = amd.DCTPlan(64)
    buf = amd.DeviceBuffer(64)
    assert lib.l2c_execute_dct_device(plan._plan, buf.ptr, buf.ptr, 4, None) == -1
    assert "kind" in amd.last_error()
    assert lib.l2c_execute_dct_device(plan._plan, buf.ptr + 8, buf.ptr, 2, None) == -1
    assert "aligned" in amd.last_error()


# ---- lines: the DCT along one axis of a field (l2c_execute_dct_lines_device)

def axis_map(shape, axis):
    """(nlines, n0, ls0, ls1, es) of the lines of a C-ordered array along axis."""
    es = int(np.prod(shape[axis + 1:]))
    outer = int(np.prod(shape[:axis]))
    return outer * es, es, 1, shape[axis] * es, es


def dct_along(amd, x, axis, kind, inplace=False):
    N = x.shape[axis]
    plan = amd.DCTPlan(N)
    din = amd.DeviceBuffer(x.size)
    dout = din if inplace else amd.DeviceBuffer(x.size)
    din.upload(x.ravel())
    plan.execute_lines_device(din.ptr, dout.ptr, kind, *axis_map(x.shape, axis))
    plan.synchronize()
    return dout.download().reshape(x.shape)


@pytest.mark.parametrize("shape", [(3, 5, 64), (4, 256, 6), (1024, 2, 3), (2, 3, 4096), (7, 1, 2)])
@pytest.mark.parametrize("kind", [2, 3])
def test_lines_along_every_axis(amd, orc, shape, kind):
    x = np.random.default_rng(sum(shape) + kind).uniform(-1, 1, shape)
    for axis in range(3):
        N = shape[axis]
        if N & (N - 1):
            continue
        announce(f"dct lines kind={kind} shape={shape} axis={axis}")
        y = dct_along(amd, x, axis, kind)
        ref = np.moveaxis(orc.redft(np.ascontiguousarray(np.moveaxis(x, axis, -1)), kind), -1, axis)
        err = max(relerr(a, b) for a, b in zip(np.moveaxis(y, axis, -1).reshape(-1, N),
                                                np.moveaxis(ref, axis, -1).reshape(-1, N)))
        assert err < tol(N), (axis, err)


@pytest.mark.parametrize("N", [1, 2, 8, 512])
def test_lines_odd_offsets_and_in_place(amd, orc, N):
    """Unit-stride lines at odd offsets (a padded row stride N + 1) take the
    strided kernels; in place along a strided axis."""
    announce(f"dct lines odd offsets N={N}")
    nl = 37
    x = np.random.default_rng(N).uniform(-1, 1, nl * (N + 1))
    plan = amd.DCTPlan(N)
    din, dout = amd.DeviceBuffer(x.size), amd.DeviceBuffer(x.size)
    din.upload(x)
    dout.upload(np.zeros(x.size))
    plan.execute_lines_device(din.ptr, dout.ptr, 2, nl, nl, N + 1, 0, 1)
    plan.synchronize()
    y = dout.download().reshape(nl, N + 1)
    xr = x.reshape(nl, N + 1)
    for r in (0, 1, nl - 1):
        assert relerr(y[r, :N], orc.redft(xr[r, :N], 2)) < tol(N)
    assert np.all(y[:, N] == 0)  # the padding is not touched
    z = np.random.default_rng(N + 1).uniform(-1, 1, (N, 3))
    got = dct_along(amd, z, 0, 3, inplace=True)
    assert relerr(got[:, 1], orc.redft(z[:, 1], 3)) < tol(N)


def test_field_values_pipeline(amd):
    """A 3-D Legendre series a_ijk -> values at the Chebyshev-Gauss grid
    (L2C, then DCT-III along every axis) equals the series evaluated there
    directly; values -> DCT-II along every axis, scaling, C2L gives a back."""
    from numpy.polynomial import legendre
    n = 32
    announce(f"dct field pipeline n={n}")
    rng = np.random.default_rng(9)
    k = np.arange(n)
    a = rng.uniform(-1, 1, (n, n, n)) / np.multiply.outer(np.multiply.outer(k + 1, k + 1), k + 1) ** 1.5
    fp = amd.FMMPlan(n, 64, 18, amd.BOTH)
    dct = amd.DCTPlan(n)
    bufs = [amd.DeviceBuffer(n ** 3) for _ in range(2)]

    def along(arr, fn):
        bufs[0].upload(arr.ravel())
        for ax in (2, 1, 0):
            fn(bufs[0].ptr, bufs[1].ptr, *axis_map((n, n, n), ax))
            bufs[0], bufs[1] = bufs[1], bufs[0]
        fp.synchronize()
        dct.synchronize()
        return bufs[0].download().reshape(n, n, n)

    c = along(a, lambda i, o, *m: fp.execute_lines_device(i, o, amd.L2C, *m))
    # per axis: values = REDFT01 of (c_0, c_1/2, ...); c = REDFT10(values)/n, c_0 halved
    hv, hc = np.where(k == 0, 1.0, 0.5), np.where(k == 0, 0.5, 1.0)
    c = c * np.multiply.outer(np.multiply.outer(hv, hv), hv)
    f = along(c, lambda i, o, *m: dct.execute_lines_device(i, o, 3, *m))
    x = np.cos(np.pi * (2 * k + 1) / (2 * n))
    V = legendre.legvander(x, n - 1)  # V[j, l] = P_l(x_j)
    f_ref = np.einsum("il,jm,kn,lmn->ijk", V, V, V, a)
    assert relerr(f, f_ref) < 1e-13
    g = along(f, lambda i, o, *m: dct.execute_lines_device(i, o, 2, *m))
    g = g / n ** 3 * np.multiply.outer(np.multiply.outer(hc, hc), hc)
    back = along(g, lambda i, o, *m: fp.execute_lines_device(i, o, amd.C2L, *m))
    assert relerr(back, a) < 1e-13


def test_lines_errors(amd):
    plan = amd.DCTPlan(1 << 13)
    buf = amd.DeviceBuffer(1 << 13)
    lib = amd.lib()
    assert lib.l2c_execute_dct_lines_device(plan._plan, buf.ptr, buf.ptr, 2, 1, 1, 0, 0, 1, None) == -1
    assert "4096" in amd.last_error()
    small = amd.DCTPlan(64)
    assert lib.l2c_execute_dct_lines_device(small._plan, buf.ptr, buf.ptr, 2, 0, 1, 0, 0, 1, None) == -1
    assert lib.l2c_execute_dct_lines_device(small._plan, buf.ptr, buf.ptr, 5, 1, 1, 0, 0, 1, None) == -1
