"""GPU parity of the M2L forms (src/C/leg2cheb.c:1419-1435):

- precision form 2: the M2L products of the blocks at levels >= lu as BF16
  slices on MFMA (6 slices of 8 bits per operand, FP32 accumulation of each
  anti-diagonal, FP64 recombination); the error against the oracle FMM and
  against double-double direct-sum rows is stated per N and must be at most
  1e-13 (the precision form's bound);
- the FP64 MFMA form of the same products for line batches (by default for
  calls of 32 lines and more; L2C_M2L_MMA=<min lines> at plan creation, 0
  off): equal to the VALU M2L within 1e-15 relative.
"""
import math

import numpy as np
import pytest

from conftest import announce, relerr

pytestmark = pytest.mark.gpu

FORM2_TOL = 1e-13


def _rows(N):
    """Every row up to N = 2^16; above, rows 0-255 and 4096 random rows (the
    double-double sums of one row cost O(N))."""
    if N <= 1 << 16:
        return np.arange(N)
    return np.unique(np.concatenate([np.arange(256), np.random.default_rng(N).integers(0, N, 4096)]))


@pytest.mark.parametrize("N,maxs", [(1 << 16, 64), (1 << 20, 64), (1 << 16, 32), (12345, 16)])
@pytest.mark.parametrize("direction", [0, 1])
def test_form2_single(amd, orc, N, maxs, direction):
    announce(f"form 2 N={N} maxs={maxs} dir={direction}")
    p = amd.FMMPlan(N, maxs, 18, direction)
    u = amd.coefficients(N, seed=40 + direction)
    g0 = p.execute(u, direction)
    p.set_precision(2)
    g2 = p.execute(u, direction)
    o = orc.FMMPlan(N, maxs, 18, direction).execute(u, direction)
    rows = _rows(N)
    ref = orc.direct_dd_rows(u, rows, direction)
    e_o = relerr(g2, o)
    e_dd = float(np.max(np.abs(g2[rows] - ref))) / float(np.max(np.abs(o)))
    e_0 = relerr(g0, o)
    print(f"  N={N} dir={direction}: form 2 vs oracle {e_o:.3e}, vs double-double rows {e_dd:.3e}; "
          f"form 0 vs oracle {e_0:.3e}")
    assert e_o <= FORM2_TOL and e_dd <= FORM2_TOL


@pytest.mark.parametrize("N,lines", [(1 << 16, 20), (1 << 18, 9)])
def test_form2_lines(amd, orc, N, lines):
    """Form 2 on batched lines (ragged last group of 8) against form 0."""
    announce(f"form 2 lines N={N} lines={lines}")
    p = amd.FMMPlan(N, 64, 18, amd.BOTH)
    x = np.random.default_rng(N).uniform(-1, 1, lines * N)
    din, dout = amd.DeviceBuffer(x.size), amd.DeviceBuffer(x.size)
    din.upload(x)
    for d in (0, 1):
        p.set_precision(0)
        p.execute_lines_device(din.ptr, dout.ptr, d, lines, lines, N, 0, 1)
        p.synchronize()
        g0 = dout.download().reshape(lines, N)
        p.set_precision(2)
        p.execute_lines_device(din.ptr, dout.ptr, d, lines, lines, N, 0, 1)
        p.synchronize()
        g2 = dout.download().reshape(lines, N)
        worst = max(relerr(g2[k], g0[k]) for k in range(lines))
        print(f"  dir {d}: form 2 vs form 0, worst line {worst:.3e}")
        assert worst <= FORM2_TOL


@pytest.mark.parametrize("N,lines", [(1 << 16, 20), (1 << 20, 3), (4096, 33)])
def test_m2l_mma_fp64_lines(amd, monkeypatch, N, lines):
    """The FP64 MFMA M2L for line batches (the default from 32 lines on)
    gives the VALU M2L's result."""
    announce(f"m2l mma fp64 N={N} lines={lines}")
    monkeypatch.setenv("L2C_M2L_MMA", "0")
    p0 = amd.FMMPlan(N, 64, 18, amd.BOTH)
    monkeypatch.setenv("L2C_M2L_MMA", "1")
    p1 = amd.FMMPlan(N, 64, 18, amd.BOTH)
    monkeypatch.delenv("L2C_M2L_MMA")
    x = np.random.default_rng(N + 1).uniform(-1, 1, lines * N)
    din, dout = amd.DeviceBuffer(x.size), amd.DeviceBuffer(x.size)
    din.upload(x)
    for d in (0, 1):
        outs = []
        for p in (p0, p1):
            p.execute_lines_device(din.ptr, dout.ptr, d, lines, lines, N, 0, 1)
            p.synchronize()
            outs.append(dout.download().reshape(lines, N))
        worst = max(relerr(outs[1][k], outs[0][k]) for k in range(lines))
        assert worst <= 1e-15, (d, worst)


@pytest.mark.parametrize("N,lines", [(1 << 20, 1), (1 << 16, 1), (100000, 1), (1 << 18, 8),
                                     (1 << 16, 16), (8191, 3)])
def test_m2l_rows_across_lanes(amd, monkeypatch, N, lines):
    """k_m2l_t (the M2L blocks below the top levels with their rows in
    registers, read from the rows-across-lanes copy of the M2L matrices; the
    default) against k_m2l (blocks staged in LDS, L2C_M2L_T=0): the same
    sums in the same order, so bit-identical outputs."""
    announce(f"m2l rows across lanes N={N} lines={lines}")
    monkeypatch.setenv("L2C_M2L_T", "0")
    p0 = amd.FMMPlan(N, 64, 18, amd.BOTH)
    monkeypatch.delenv("L2C_M2L_T")
    p1 = amd.FMMPlan(N, 64, 18, amd.BOTH)
    x = np.random.default_rng(N + 7).uniform(-1, 1, lines * N)
    din, dout = amd.DeviceBuffer(x.size), amd.DeviceBuffer(x.size)
    din.upload(x)
    for d in (0, 1):
        outs = []
        for p in (p0, p1):
            if lines == 1:
                p.execute_device(din.ptr, dout.ptr, d)
            else:
                p.execute_lines_device(din.ptr, dout.ptr, d, lines, lines, N, 0, 1)
            p.synchronize()
            outs.append(dout.download())
        assert np.array_equal(outs[0], outs[1]), d
