"""GPU parity of the downward pass with wave roles (k_down8r, the default:
the tree and far field on half of a work-group's waves, the band on the
other half; L2C_DOWN_ROLES=0 at plan creation selects the fused k_down8)
against the oracle FMM and against k_down8 (src/C/leg2cheb.c:1437-1483 and
the band of :236-535): leaf sizes 16, 32 and 64, ragged last chunks, both
directions, the nodal and precompute variants, line batches, and the
headline size 2^20.  Tolerance 1e-14 log2 N relative max-norm against the
oracle; the two kernels sum the same terms in different orders, so they
agree to the same bound.  test_roles_shift moves the band split between the
roles both ways (L2C_ROLES_SHIFT).  test_split_pass covers the split
downward pass (k_band, then k_far; L2C_DOWN_SPLIT=1 or 2) the same way.
"""
import math

import numpy as np
import pytest

from conftest import announce, relerr

pytestmark = pytest.mark.gpu


def tol(N):
    return 1e-14 * math.log2(max(N, 2))


def _plans(amd, monkeypatch, *args, **kw):
    """(k_down8 plan, k_down8r plan): the roles kernel is the default."""
    roles = amd.FMMPlan(*args, **kw)
    monkeypatch.setenv("L2C_DOWN_ROLES", "0")
    fused = amd.FMMPlan(*args, **kw)
    monkeypatch.delenv("L2C_DOWN_ROLES")
    return fused, roles


@pytest.mark.parametrize("N,maxs", [(1000, 16), (4096, 64), (16284, 64), (8192, 32), (4000, 32),
                                    (65536, 64), (1 << 20, 64)])
@pytest.mark.parametrize("direction", [0, 1])
def test_roles_match(amd, orc, monkeypatch, N, maxs, direction):
    announce(f"k_down8r N={N} maxs={maxs} dir={direction}")
    fused, roles = _plans(amd, monkeypatch, N, maxs, 18, direction)
    assert roles.info()["s"] in (16, 32, 64)
    for decay in (0.0, 1.5):
        u = amd.coefficients(N, seed=31 + direction, decay=decay)
        g = roles.execute(u, direction)
        f = fused.execute(u, direction)
        o = orc.FMMPlan(N, maxs, 18, direction).execute(u, direction)
        assert relerr(g, o) <= tol(N), (decay, relerr(g, o))
        assert relerr(g, f) <= tol(N), (decay, relerr(g, f))


@pytest.mark.parametrize("lagrange,precompute", [(1, 0), (0, 1), (1, 1)])
def test_roles_variants(amd, orc, monkeypatch, lagrange, precompute):
    N = 16284
    announce(f"k_down8r variants lagrange={lagrange} precompute={precompute}")
    u = amd.coefficients(N, seed=5)
    for d in (0, 1):
        fused, roles = _plans(amd, monkeypatch, N, 64, 18, d, lagrange, precompute)
        o = orc.FMMPlan(N, 64, 18, d, lagrange, precompute).execute(u, d)
        g = roles.execute(u, d)
        assert relerr(g, o) <= tol(N)
        assert relerr(g, fused.execute(u, d)) <= tol(N)


@pytest.mark.parametrize("N,nlines", [(4096, 9), (65536, 4)])
def test_roles_lines(amd, orc, monkeypatch, N, nlines):
    """Strided lines (element stride nlines, line stride 1) through one call."""
    announce(f"k_down8r lines N={N} nlines={nlines}")
    fused, roles = _plans(amd, monkeypatch, N, 64, 18, 0)
    x = np.stack([amd.coefficients(N, seed=60 + l) for l in range(nlines)], axis=1)
    din, dout = amd.DeviceBuffer(N * nlines), amd.DeviceBuffer(N * nlines)
    din.upload(x)
    roles.execute_lines_device(din.ptr, dout.ptr, 0, nlines, nlines, 1, 0, nlines)
    roles.synchronize()
    g = dout.download().reshape(N, nlines)
    op = orc.FMMPlan(N, 64, 18, 0)
    for l in range(nlines):
        assert relerr(g[:, l], op.execute(x[:, l], 0)) <= tol(N)


@pytest.mark.parametrize("shift", [-64, -8, 24])
def test_roles_shift(amd, orc, monkeypatch, shift):
    N = 16284
    announce(f"k_down8r shift={shift}")
    u = amd.coefficients(N, seed=8)
    monkeypatch.setenv("L2C_ROLES_SHIFT", str(shift))
    plans = [amd.FMMPlan(N, 64, 18, d) for d in (0, 1)]
    monkeypatch.delenv("L2C_ROLES_SHIFT")
    for d in (0, 1):
        o = orc.FMMPlan(N, 64, 18, d).execute(u, d)
        assert relerr(plans[d].execute(u, d), o) <= tol(N)


@pytest.mark.parametrize("N,maxs", [(1000, 16), (16284, 64), (4000, 32), (1 << 18, 64)])
@pytest.mark.parametrize("direction", [0, 1])
def test_split_pass(amd, orc, monkeypatch, N, maxs, direction):
    """k_band + k_far (L2C_DOWN_SPLIT=1) for single transforms, and with
    L2C_DOWN_SPLIT=2 for a call of 3 lines, against the oracle."""
    announce(f"split pass N={N} maxs={maxs} dir={direction}")
    monkeypatch.setenv("L2C_DOWN_SPLIT", "1")
    p1 = amd.FMMPlan(N, maxs, 18, direction)
    monkeypatch.setenv("L2C_DOWN_SPLIT", "2")
    p2 = amd.FMMPlan(N, maxs, 18, direction)
    monkeypatch.delenv("L2C_DOWN_SPLIT")
    op = orc.FMMPlan(N, maxs, 18, direction)
    u = amd.coefficients(N, seed=70 + direction)
    assert relerr(p1.execute(u, direction), op.execute(u, direction)) <= tol(N)
    nl = 3
    x = np.concatenate([amd.coefficients(N, seed=80 + l) for l in range(nl)])
    din, dout = amd.DeviceBuffer(N * nl), amd.DeviceBuffer(N * nl)
    din.upload(x)
    p2.execute_lines_device(din.ptr, dout.ptr, direction, nl, nl, N, 0, 1)
    p2.synchronize()
    g = dout.download().reshape(nl, N)
    for l in range(nl):
        assert relerr(g[l], op.execute(x[l * N:(l + 1) * N], direction)) <= tol(N)
