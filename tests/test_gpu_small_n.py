"""Single transforms of small N through an FMM plan: up to N = 8192 (the
plan field direct_max_n, L2C_DIRECT_MAXN) execute() and l2c_execute_device
run the direct sums of src/C/leg2cheb.c:152-234 in one launch
(k_direct_rows, or k_direct_band from N = 3073) instead of the three FMM
launches, and so do l2c_execute_lines_device calls of at most 16,384
elements (L2C_DIRECT_LINES), every line in one launch.  Checked against the oracle's FMM (the reference's execute) and the
double-double direct sums, tolerance 1e-14 log2 N, and against the plan's own
FMM route."""
import math

import numpy as np
import pytest

from conftest import announce, relerr

pytestmark = pytest.mark.gpu


def tol(N):
    return 1e-14 * math.log2(max(N, 2))


SIZES = [1000, 1024, 3001, 3072, 3073, 4096, 6000, 8191, 8192]


def make_plan(amd, monkeypatch, N, maxs, direction, maxn=None, band=None):
    """A plan under the default route (maxn None) or L2C_DIRECT_MAXN=maxn."""
    if maxn is None:
        monkeypatch.delenv("L2C_DIRECT_MAXN", raising=False)
    else:
        monkeypatch.setenv("L2C_DIRECT_MAXN", str(maxn))
    if band is None:
        monkeypatch.delenv("L2C_DIRECT_BAND", raising=False)
    else:
        monkeypatch.setenv("L2C_DIRECT_BAND", str(band))
    return amd.FMMPlan(N, maxs, 18, direction)


def route(plan, run):
    """The profile slots one call of run() fills."""
    plan.profile(True)
    run()
    slots = set(plan.profile_read())
    plan.profile(False)
    return slots


@pytest.mark.parametrize("N", SIZES)
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("decay", [0.0, 1.5])
def test_small_n_default_route(amd, orc, monkeypatch, N, direction, decay):
    announce(f"small-N route N={N} dir={direction} decay={decay}")
    u = amd.coefficients(N, seed=500 + N, decay=decay)
    p = make_plan(amd, monkeypatch, N, 64, direction)
    assert route(p, lambda: p.execute(u, direction)) == {"k_direct"}
    g = p.execute(u, direction)
    o = orc.FMMPlan(N, 64, 18, direction).execute(u, direction)
    e = relerr(g, o)
    ref = orc.leg2cheb_dd(u) if direction == 0 else orc.cheb2leg_dd(u)
    e2 = relerr(g, ref)
    print(f"  rel err vs oracle FMM {e:.3e}, vs double-double direct {e2:.3e}")
    assert e <= tol(N) and e2 <= tol(N)
    # the same plan size on the FMM route agrees within the same bound
    q = make_plan(amd, monkeypatch, N, 64, direction, maxn=0)
    assert "k_up" in route(q, lambda: q.execute(u, direction))
    assert relerr(g, q.execute(u, direction)) <= tol(N)


@pytest.mark.parametrize("N", [3073, 4096, 8192])
@pytest.mark.parametrize("direction", [0, 1])
def test_small_n_kernels_agree(amd, monkeypatch, N, direction):
    """k_direct_band against k_direct_rows (L2C_DIRECT_BAND=0) on the same
    plan size: the same sums in another order."""
    announce(f"direct kernels N={N} dir={direction}")
    u = amd.coefficients(N, seed=77 + N, decay=1.0)
    a = make_plan(amd, monkeypatch, N, 64, direction).execute(u, direction)
    b = make_plan(amd, monkeypatch, N, 64, direction, band=0).execute(u, direction)
    e = relerr(a, b)
    print(f"  band vs rows {e:.3e}")
    assert e <= 2e-15


@pytest.mark.parametrize("N", [4096, 8191])
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("graph", [False, True])
def test_small_n_device_strided(amd, orc, monkeypatch, N, direction, graph):
    """l2c_execute_device with stride 3, plain and replayed from the plan's
    HIP graph (three replays, each checked); elements between the strided
    ones stay untouched."""
    announce(f"small-N device N={N} dir={direction} graph={graph}")
    stride = 3
    u = amd.coefficients(N, seed=9 + N + direction)
    x = np.zeros(N * stride)
    x[::stride] = u
    din, dout = amd.DeviceBuffer(N * stride), amd.DeviceBuffer(N * stride)
    din.upload(x)
    p = make_plan(amd, monkeypatch, N, 64, direction)
    if graph:
        p.graph(True)
    o = orc.FMMPlan(N, 64, 18, direction).execute(u, direction)
    for rep in range(3 if graph else 1):
        dout.upload(np.full(N * stride, 7.0))
        p.execute_device(din.ptr, dout.ptr, direction, stride)
        p.synchronize()
        y = dout.download()
        e = relerr(y[::stride], o)
        print(f"  replay {rep}: rel err vs oracle FMM {e:.3e}")
        assert e <= tol(N)
        assert np.all(y.reshape(N, stride)[:, 1:] == 7.0)


@pytest.mark.parametrize("N", [8193, 16384])
def test_above_threshold_keeps_fmm(amd, orc, monkeypatch, N):
    """Past 8192 the FMM runs as before."""
    announce(f"above threshold N={N}")
    u = amd.coefficients(N, seed=3)
    p = make_plan(amd, monkeypatch, N, 64, 0)
    assert "k_up" in route(p, lambda: p.execute(u, 0))
    assert relerr(p.execute(u, 0), orc.FMMPlan(N, 64, 18, 0).execute(u, 0)) <= tol(N)


@pytest.mark.parametrize("N,B,route_slot", [(4096, 4, "k_direct"), (4096, 5, "k_up"),
                                             (8192, 2, "k_direct"), (2048, 8, "k_direct"),
                                             (8192, 3, "k_up")])
def test_small_n_lines_route(amd, orc, monkeypatch, N, B, route_slot):
    """Line calls of a small plan holding at most 16,384 elements take the
    direct sums in one launch (all lines in grid.z), larger ones the FMM;
    either way every line matches the oracle and the single-line call."""
    announce(f"small-N lines N={N} B={B}")
    p = make_plan(amd, monkeypatch, N, 64, 2)
    u = np.concatenate([amd.coefficients(N, seed=60 + b) for b in range(B)])
    din, dout = amd.DeviceBuffer(B * N), amd.DeviceBuffer(B * N)
    din.upload(u)
    for d in (0, 1):
        slots = route(p, lambda: (p.execute_lines_device(din.ptr, dout.ptr, d, B, B, N, 0, 1),
                                  p.synchronize()))
        assert route_slot in slots
        assert ("k_direct" in slots) == (route_slot == "k_direct")
        y = dout.download().reshape(B, N)
        op = orc.FMMPlan(N, 64, 18, d)
        for b in range(B):
            ub = u[b * N:(b + 1) * N]
            assert relerr(y[b], op.execute(ub, d)) <= tol(N)
            assert relerr(y[b], p.execute(ub, d)) <= tol(N)


def test_small_n_lines_strided(amd, orc, monkeypatch):
    """Two interleaved lines (element stride 2, line offsets 0 and 1) of
    N = 4096 through the direct route."""
    announce("small-N interleaved lines")
    N = 4096
    p = make_plan(amd, monkeypatch, N, 64, 0)
    x = np.stack([amd.coefficients(N, seed=80), amd.coefficients(N, seed=81)], axis=1).ravel()
    din, dout = amd.DeviceBuffer(2 * N), amd.DeviceBuffer(2 * N)
    din.upload(x)
    slots = route(p, lambda: (p.execute_lines_device(din.ptr, dout.ptr, 0, 2, 2, 1, 0, 2),
                              p.synchronize()))
    assert slots == {"k_direct"}
    y = dout.download().reshape(N, 2)
    op = orc.FMMPlan(N, 64, 18, 0)
    for b in range(2):
        assert relerr(y[:, b], op.execute(np.ascontiguousarray(x.reshape(N, 2)[:, b]), 0)) <= tol(N)


@pytest.mark.parametrize("lagrange,precompute", [(1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("maxs", [32, 64])
def test_small_n_plan_variants(amd, orc, monkeypatch, lagrange, precompute, maxs):
    """The nodal (lagrange=1) and precompute plans and another leaf size take
    the same direct route: the transform, not the plan's FMM variant, is
    what they compute."""
    N = 4096
    announce(f"small-N variants lagrange={lagrange} precompute={precompute} maxs={maxs}")
    u = amd.coefficients(N, seed=15)
    monkeypatch.delenv("L2C_DIRECT_MAXN", raising=False)
    for d in (0, 1):
        p = amd.FMMPlan(N, maxs, 18, d, lagrange, precompute)
        assert route(p, lambda: p.execute(u, d)) == {"k_direct"}
        g = p.execute(u, d)
        o = orc.FMMPlan(N, maxs, 18, d, lagrange, precompute).execute(u, d)
        ref = orc.leg2cheb_dd(u) if d == 0 else orc.cheb2leg_dd(u)
        assert relerr(g, o) <= tol(N) and relerr(g, ref) <= tol(N)
