"""l2c_execute_lines_device_noalloc (the entry point that allocates nothing:
its FMM workspace is a fixed device array) against l2c_execute_lines_device
through the C ABI: identical outputs on two calls in a row, the second with
fewer lines, for FMM plans (several launches when the lines outgrow the
array; few lines of a small plan as direct sums) and a direct-path plan; a
plan whose single line outgrows the array is refused before any launch."""
import math

import numpy as np
import pytest

from conftest import announce, relerr

pytestmark = pytest.mark.gpu


def tol(N):
    return 1e-14 * math.log2(max(N, 2))


def _lines(amd, fn, plan, x, d, nlines, N):
    din, dout = amd.DeviceBuffer(x.size), amd.DeviceBuffer(x.size)
    din.upload(x)
    dout.upload(np.full(x.size, 7.0))
    rc = fn(plan.plan, din.ptr, dout.ptr, d, nlines, nlines, N, 0, 1, None)
    assert rc == 0, amd.last_error()
    plan.synchronize()
    return dout.download()


# (N, maxs, first call's lines, second call's lines): 2^20 lines need 10.5 MB
# of workspace each, so 5 lines take two launches of the 32 MiB array
@pytest.mark.parametrize("N,maxs,n1,n2", [(65536, 64, 40, 7), (1 << 20, 64, 5, 2),
                                          (4097, 32, 300, 33), (200, 64, 100, 10)])
def test_noalloc_matches_lines_entry(amd, orc, N, maxs, n1, n2):
    L = amd.lib()
    plan = amd.FMMPlan(N, maxs, 18, amd.BOTH)
    op = orc.FMMPlan(N, maxs, 18, orc.BOTH)
    for d in (0, 1):
        for nl in (n1, n2):
            announce(f"noalloc N={N} maxs={maxs} lines={nl} dir={d}")
            x = np.random.default_rng(N + nl + d).uniform(-1, 1, nl * N)
            a = _lines(amd, L.l2c_execute_lines_device_noalloc, plan, x, d, nl, N)
            b = _lines(amd, L.l2c_execute_lines_device, plan, x, d, nl, N)
            assert np.array_equal(a, b)
            for line in (0, nl - 1):
                ref = op.execute(np.ascontiguousarray(x[line * N:(line + 1) * N]), d)
                assert relerr(a[line * N:(line + 1) * N], ref) <= tol(N)


def test_noalloc_direct_route(amd, orc, monkeypatch):
    """With the default small-N route a call of at most 16,384 elements of
    an FMM plan runs as direct sums in both entry points (no workspace);
    4 lines of 4096, then 2, identical outputs."""
    monkeypatch.delenv("L2C_DIRECT_MAXN", raising=False)
    N = 4096
    L = amd.lib()
    plan = amd.FMMPlan(N, 64, 18, amd.BOTH)
    op = orc.FMMPlan(N, 64, 18, orc.BOTH)
    for d in (0, 1):
        for nl in (4, 2):
            announce(f"noalloc direct route N={N} lines={nl} dir={d}")
            x = np.random.default_rng(N + nl + d).uniform(-1, 1, nl * N)
            a = _lines(amd, L.l2c_execute_lines_device_noalloc, plan, x, d, nl, N)
            b = _lines(amd, L.l2c_execute_lines_device, plan, x, d, nl, N)
            assert np.array_equal(a, b)
            for line in range(nl):
                ref = op.execute(np.ascontiguousarray(x[line * N:(line + 1) * N]), d)
                assert relerr(a[line * N:(line + 1) * N], ref) <= tol(N)


def test_noalloc_refuses_oversized_line(amd):
    N = 1 << 22
    announce(f"noalloc refuses N={N}")
    L = amd.lib()
    plan = amd.FMMPlan(N, 64, 18, amd.L2C)
    buf = amd.DeviceBuffer(8)
    rc = L.l2c_execute_lines_device_noalloc(plan.plan, buf.ptr, buf.ptr, 0, 1, 1, N, 0, 1, None)
    assert rc == -1 and "scratch" in amd.last_error()
