"""GPU parity: the HIP path against the oracle (CPU restatement of the
reference's execute, src/C/leg2cheb.c:1299-1500) and the extended-precision
direct sums.  Tolerance: relative max-norm error <= 1e-14 * log2(N)."""
import math

import numpy as np
import pytest

from conftest import announce, relerr

pytestmark = pytest.mark.gpu


def tol(N):
    return 1e-14 * math.log2(max(N, 2))


# leaf sizes s: 33 (257), 19 (300), 64 (512), 32 (1000), 33 (1025), 64 (4096), 16 (4096/16),
# 64 (8191, last chunk ragged), 32 (8192), 32 (16001, ragged), 16 (4000, ragged), 49 (12345),
# 64 (65000, ragged), 64 (65536); s = 16, 32, 64 take the k_down8 kernel
CASES = [(257, 64), (300, 32), (512, 64), (1000, 32), (1025, 64), (4096, 64), (4096, 16),
         (8191, 64), (8192, 32), (16001, 32), (4000, 16), (12345, 64), (65000, 64), (65536, 64)]


@pytest.mark.parametrize("N,maxs", CASES)
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("decay", [0.0, 1.5])
def test_fmm_matches_oracle(amd, orc, N, maxs, direction, decay):
    """Both input families of the benchmark at every size: U(-1,1) and
    U(-1,1) (n+1)^-1.5."""
    announce(f"fmm N={N} maxs={maxs} dir={direction} decay={decay}")
    u = amd.coefficients(N, seed=1000 + N, decay=decay)
    g = amd.FMMPlan(N, maxs, 18, direction).execute(u, direction)
    o = orc.FMMPlan(N, maxs, 18, direction).execute(u, direction)
    e = relerr(g, o)
    print(f"  rel err vs oracle FMM = {e:.3e}")
    assert e <= tol(N)
    if N <= 8192:
        ref = orc.leg2cheb_dd(u) if direction == 0 else orc.cheb2leg_dd(u)
        e2 = relerr(g, ref)
        print(f"  rel err vs double-double direct = {e2:.3e}")
        assert e2 <= tol(N)


@pytest.mark.parametrize("N", [1, 2, 3, 20, 64, 100, 255, 256])
@pytest.mark.parametrize("direction", [0, 1])
def test_direct_path(amd, orc, N, direction):
    announce(f"direct-path N={N} dir={direction}")
    u = amd.coefficients(N, seed=7 + N)
    g = amd.FMMPlan(N, 64, 18, direction).execute(u, direction)
    o = orc.FMMPlan(N, 64, 18, direction).execute(u, direction)
    assert relerr(g, o) <= tol(N)
    d = amd.leg2cheb(u, np.zeros(N)) if direction == 0 else amd.cheb2leg(u, np.zeros(N))
    assert relerr(d, o) <= tol(N)


def test_known_answer_N20(amd):
    """src/bin/l2c.c:69-74, 94, 104."""
    announce("KAT N=20")
    a = [1.7913439415860921, 2.3408718378050253, 1.8833606631960720, 1.6587642683880404,
         1.4417173322290182]
    b = [0.52631578947368418, 0.075187969924812026, 0.13268465280849182, 0.17768205680441512,
         0.24367824933176932]
    u = np.ones(20)
    p = amd.FMMPlan(20, 64, 18, 2)
    assert np.linalg.norm(p.execute(u, 0)[:5] - a) < 1e-7
    assert np.linalg.norm(p.execute(u, 1)[:5] - b) < 1e-7


@pytest.mark.parametrize("N,maxs", [(1000, 32), (512, 64), (100, 32), (10000, 36)])
def test_round_trip(amd, N, maxs):
    """src/bin/l2c.c:10-65 (-d2): max error < 1e-10."""
    announce(f"round trip N={N} maxs={maxs}")
    u = amd.coefficients(N, seed=3) / np.arange(1, N + 1)
    p = amd.FMMPlan(N, maxs, 18, 2)
    b = p.execute(u, 0)
    c = p.execute(b, 1)
    assert np.max(np.abs(c - u)) < 1e-10


@pytest.mark.parametrize("N,stride", [(1000, 3), (4097, 2), (200, 5), (8192, 3)])
def test_strided(amd, orc, N, stride):
    announce(f"strided N={N} stride={stride}")
    u = amd.coefficients(N, seed=11)
    big = np.zeros(N * stride)
    big[::stride] = u
    p = amd.FMMPlan(N, 64, 18, 2)
    for d in (0, 1):
        out = np.full(N * stride, 7.0)
        p.execute(big, d, stride=stride, out=out)
        ref = orc.FMMPlan(N, 64, 18, 2).execute(u, d)
        assert relerr(out[::stride], ref) <= tol(N)
        mask = np.ones(N * stride, bool)
        mask[::stride] = False
        assert np.all(out[mask] == 7.0)


@pytest.mark.parametrize("N,maxs", [(3000, 32), (4096, 64)])
@pytest.mark.parametrize("lagrange,precompute", [(1, 0), (0, 1), (1, 1)])
def test_variants(amd, orc, N, maxs, lagrange, precompute):
    announce(f"variant N={N} maxs={maxs} lagrange={lagrange} precompute={precompute}")
    u = amd.coefficients(N, seed=5)
    for d in (0, 1):
        g = amd.FMMPlan(N, maxs, 18, d, lagrange, precompute).execute(u, d)
        o = orc.FMMPlan(N, maxs, 18, d, lagrange, precompute).execute(u, d)
        assert relerr(g, o) <= tol(N)


@pytest.mark.parametrize("N", [5000, 4096])
def test_directM(amd, orc, N):
    announce(f"directM N={N}")
    u = amd.coefficients(N, seed=9)
    g = amd.FMMPlan(N, 64, 18, 0).directM(u)
    o = np.zeros(N)
    op = orc.FMMPlan(N, 64, 18, 0)
    orc.lib().orc_directM(orc._p(u), orc._p(o), op._p, 1)
    assert relerr(g, o) <= tol(N)


@pytest.mark.parametrize("N,maxs,M", [(65536, 64, 18), (8191, 64, 24), (4096, 32, 12)])
def test_downward_kernels_agree(amd, orc, monkeypatch, N, maxs, M):
    """The s = 16/32/64 downward kernel and the general one give the same
    transform (L2C_DOWN8=0 selects the general kernel at plan creation)."""
    announce(f"k_down8 vs k_down N={N} maxs={maxs} M={M}")
    u = amd.coefficients(N, seed=21)
    for d in (0, 1):
        fast = amd.FMMPlan(N, maxs, M, d).execute(u, d)
        monkeypatch.setenv("L2C_DOWN8", "0")
        gen = amd.FMMPlan(N, maxs, M, d).execute(u, d)
        monkeypatch.delenv("L2C_DOWN8")
        o = orc.FMMPlan(N, maxs, M, d).execute(u, d)
        t = tol(N) if M >= 18 else 1e-9
        assert relerr(fast, o) <= t and relerr(gen, o) <= t
        assert relerr(fast, gen) <= tol(N)


@pytest.mark.parametrize("N0,N1", [(400, 402), (512, 600)])
@pytest.mark.parametrize("axis", [0, 1, -1])
def test_2d(amd, orc, axis, N0, N1):
    """src/bin/l2c.c:297-393 (-d4, -d5)."""
    announce(f"2d {N0}x{N1} axis={axis}")
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (N0, N1))
    for d in (0, 1):
        g = amd.FMMPlan2D(N0, N1, axis, 64).execute(x, d)
        o = orc.FMMPlan2D(N0, N1, axis, 64).execute(x, d)
        assert relerr(g, o) <= tol(max(N0, N1))


@pytest.mark.parametrize("N,maxs", [(4096, 64), (65536, 64), (12345, 64), (8192, 32), (3000, 32),
                                    (1 << 18, 64)])
def test_split_form(amd, orc, N, maxs):
    """Precision form 1 (M2L matrices stored per coefficient position as FP64,
    FP32, BF16 or dropped, from their magnitude relative to the block's
    largest): same tolerance against the oracle, and within a few ulps of
    form 0; rows 0-255 and sampled rows against the double-double direct sum."""
    announce(f"split form N={N} maxs={maxs}")
    p = amd.FMMPlan(N, maxs, 18, amd.BOTH)
    for d in (0, 1):
        u = amd.coefficients(N, seed=300 + d, decay=0.0)
        g0 = p.execute(u, d)
        p.set_precision(1)
        g1 = p.execute(u, d)
        p.set_precision(0)
        o = orc.FMMPlan(N, maxs, 18, d).execute(u, d)
        e1, e01 = relerr(g1, o), relerr(g1, g0)
        print(f"  dir {d}: form 1 vs oracle {e1:.3e}, vs form 0 {e01:.3e}")
        assert e1 <= tol(N)
        assert e01 <= 1e-14
        rows = np.unique(np.concatenate([np.arange(256),
                                         np.random.default_rng(N).integers(0, N, 256)]))
        ref = orc.direct_dd_rows(u, rows, d)
        assert float(np.max(np.abs(g1[rows] - ref))) / float(np.max(np.abs(o))) <= tol(N)


def test_split_form_nodal(amd, orc):
    """Nodal M2L matrices (lagrange=1) hold samples that do not decay: the
    split form keeps them in FP64 and the transform is unchanged."""
    announce("split form nodal")
    N = 5000
    u = amd.coefficients(N, seed=17)
    p = amd.FMMPlan(N, 64, 18, 0, 1)
    g0 = p.execute(u, 0)
    p.set_precision(1)
    g1 = p.execute(u, 0)
    assert relerr(g1, g0) <= 1e-15
    assert relerr(g1, orc.FMMPlan(N, 64, 18, 0, 1).execute(u, 0)) <= tol(N)


@pytest.mark.parametrize("N", [1 << 22, 3000001])
def test_large(amd, orc, N):
    """The largest sizes: 2^22 (L = 14, s = 64) and an odd length whose leaf
    size (46) takes the general downward kernel; both directions against the
    oracle, L2C rows 0-63 and sampled rows against the double-double sum."""
    announce(f"large N={N}")
    p = amd.FMMPlan(N, 64, 18, amd.BOTH)
    op = orc.FMMPlan(N, 64, 18, orc.BOTH)
    for d in (0, 1):
        u = amd.coefficients(N, seed=900 + d, decay=1.0 if d else 0.0)
        g = p.execute(u, d)
        o = op.execute(u, d)
        e = relerr(g, o)
        print(f"  dir {d}: rel err vs oracle {e:.3e}")
        assert e <= tol(N)
        if d == 0:
            rows = np.unique(np.concatenate([np.arange(64),
                                             np.random.default_rng(N).integers(0, N, 64)]))
            ref = orc.direct_dd_rows(u, rows, 0)
            assert float(np.max(np.abs(g[rows] - ref))) / float(np.max(np.abs(o))) <= tol(N)


def _fails(rc, amd, text):
    """rc is the error return and the error text names the cause."""
    return rc in (-1, (1 << 64) - 1) and text in amd.last_error()


@pytest.mark.parametrize("N", [200, 4096])
def test_unplanned_direction_rejected(amd, N):
    """Plans made for L2C alone (a direct-path plan at N = 200 and an FMM
    plan at N = 4096) refuse C2L at every entry point before any launch:
    execute, l2c_execute_device, l2c_execute_lines_device, and execute2D
    on 2-D plans along each axis (the reference dereferences NULL here,
    src/C/leg2cheb.c:1309-1312)."""
    announce(f"unplanned direction N={N}")
    L = amd.lib()
    p = amd.FMMPlan(N, 64, 18, amd.L2C)
    d_in, d_out = amd.DeviceBuffer(4 * N), amd.DeviceBuffer(4 * N)
    d_in.upload(np.zeros(4 * N))
    u, out = np.ones(N), np.zeros(N)
    assert _fails(L.execute(u.ctypes.data, out.ctypes.data, p.plan, 1, 1), amd, "not planned")
    assert _fails(L.l2c_execute_device(p.plan, d_in.ptr, d_out.ptr, 1, 1, None), amd, "not planned")
    assert _fails(L.l2c_execute_lines_device(p.plan, d_in.ptr, d_out.ptr, 1, 4, 4, N, 0, 1, None),
                  amd, "not planned")
    # the planned direction still runs
    assert L.l2c_execute_lines_device(p.plan, d_in.ptr, d_out.ptr, 0, 4, 4, N, 0, 1, None) == 0
    p.synchronize()
    for axis in (0, 1, -1):
        q = L.create_fmm_2d(N, 8, axis, 64, 18, amd.L2C, 0, 0, 0)
        assert q
        x, y = np.ones(8 * N), np.zeros(8 * N)
        assert _fails(L.execute2D(x.ctypes.data, y.ctypes.data, q, 1), amd, "not planned")
        assert L.execute2D(x.ctypes.data, y.ctypes.data, q, 0) != (1 << 64) - 1
        L.free_fmm_2d(q)


@pytest.mark.parametrize("N", [200, 4096])
def test_mixed_host_device_pointers_rejected(amd, N):
    """A host input with a device output (or the reverse) is an error at
    execute, direct and directM: the host-staged path would touch device
    memory from the CPU."""
    announce(f"mixed pointers N={N}")
    L = amd.lib()
    p = amd.FMMPlan(N, 64, 18, amd.BOTH)
    dbuf = amd.DeviceBuffer(N)
    dbuf.upload(np.zeros(N))
    h = np.zeros(N)
    for a, b in ((h.ctypes.data, dbuf.ptr), (dbuf.ptr, h.ctypes.data)):
        assert _fails(L.execute(a, b, p.plan, 0, 1), amd, "both be device pointers or both host")
        if N > 256:
            assert _fails(L.directM(a, b, p.plan, 1), amd, "both be device pointers or both host")
    dp = L.create_direct(N, amd.BOTH, 0)
    assert dp
    assert _fails(L.direct(h.ctypes.data, dbuf.ptr, dp, 0, 1), amd, "both be device pointers or both host")
    L.free_direct(dp)
    # matching kinds still work
    u = amd.coefficients(N, seed=3)
    assert relerr(p.execute(u, 0), amd.FMMPlan(N, 64, 18, 0).execute(u, 0)) == 0.0


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("nlines", [1, 8, 9])
def test_upward_level_forms(amd, orc, direction, nlines):
    """k_up's two forms of the subtree levels: on MFMA for subtrees of 2^3 and
    2^4 leaf chunks in calls of up to 8 lines (N = 2^18: one line, 2^3 chunks;
    8 lines, 2^4 chunks), on VALU otherwise (9 lines).  Every line of the call
    is checked against the oracle's single transform (src/C/leg2cheb.c:1385-1399
    for the upward pass)."""
    N = 1 << 18
    announce(f"upward level forms N={N} lines={nlines} dir={direction}")
    plan = amd.FMMPlan(N, 64, 18, direction)
    op = orc.FMMPlan(N, 64, 18, direction)
    u = np.concatenate([amd.coefficients(N, seed=300 + k, decay=1.5 * (k % 2)) for k in range(nlines)])
    din, dout = amd.DeviceBuffer(u.size), amd.DeviceBuffer(u.size)
    din.upload(u)
    if nlines == 1:
        plan.execute_device(din.ptr, dout.ptr, direction)
    else:
        plan.execute_lines_device(din.ptr, dout.ptr, direction, nlines, nlines, N, 0, 1)
    plan.synchronize()
    g = dout.download()
    for k in sorted({0, nlines - 1}):
        o = op.execute(u[k * N:(k + 1) * N], direction)
        e = relerr(g[k * N:(k + 1) * N], o)
        print(f"  line {k}: rel err vs oracle FMM = {e:.3e}")
        assert e <= tol(N)


@pytest.mark.parametrize("N", [1023, 1024, 3001, 3073, 5000, 8192, 12345, 32768])
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("stride", [1, 3])
def test_direct_api_rows(amd, orc, N, direction, stride):
    """The reference's direct() (src/C/leg2cheb.c:152-234) on a direct plan:
    below 1024 one thread per row (k_direct), from 1024 on one wave per row
    (k_direct_rows), from 3073 on the band kernel's register rings
    (k_direct_band; past 8192 its rows' columns in several LDS window
    chunks); strided arrays, both directions, against the oracle's direct
    sums."""
    announce(f"direct() N={N} dir={direction} stride={stride}")
    L = amd.lib()
    u = amd.coefficients(N, seed=41 + N + direction, decay=1.0 * direction)
    x = np.zeros(N * stride)
    x[::stride] = u
    din, dout = amd.DeviceBuffer(N * stride), amd.DeviceBuffer(N * stride)
    din.upload(x)
    dout.upload(np.full(N * stride, 7.0))
    dp = L.create_direct(N, direction, 0)
    assert dp
    try:  # device pointers: the kernels take the stride themselves
        assert L.direct(din.ptr, dout.ptr, dp, direction, stride) != amd._SIZE_MAX
    finally:
        L.free_direct(dp)
    y = dout.download()
    o = orc.DirectPlan(N, direction).execute(u, direction)
    e = relerr(y[::stride], o)
    print(f"  rel err vs oracle direct = {e:.3e}")
    assert e <= tol(N)
    if stride > 1:
        assert np.all(y.reshape(N, stride)[:, 1:] == 7.0)  # elements between the strided ones untouched
