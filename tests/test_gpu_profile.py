"""GPU tests of the per-kernel timing of l2c_profile / l2c_profile_read (the
figures bench.py reports as kernels_us): each timed call adds one launch to
each kernel slot it runs, also where a slot runs two kernels (the M2L of
precision form 2: k_m2l above the upward-pass levels, k_m2l_mma below), the
durations are the kernels' own execution (start/stop events of the dispatch,
so per call they add up to less than the call's wall time), the period n
times every n-th call, and timing leaves the results unchanged.
"""
import time

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


def _run(amd, p, N, calls, direction=0, nlines=1):
    u = np.concatenate([amd.coefficients(N, seed=90 + l) for l in range(nlines)])
    din, dout = amd.DeviceBuffer(N * nlines), amd.DeviceBuffer(N * nlines)
    din.upload(u)
    t0 = time.perf_counter()
    for _ in range(calls):
        if nlines == 1:
            p.execute_device(din.ptr, dout.ptr, direction)
        else:
            p.execute_lines_device(din.ptr, dout.ptr, direction, nlines, nlines, N, 0, 1)
    p.synchronize()
    wall_ms = 1e3 * (time.perf_counter() - t0)
    return dout.download(), wall_ms


@pytest.mark.parametrize("precision", [0, 2])
def test_profile_counts_and_durations(amd, precision):
    N, calls = 1 << 18, 12
    announce(f"profile N={N} precision={precision}")
    p = amd.FMMPlan(N, 64, 18, 0)
    p.set_precision(precision)
    ref, _ = _run(amd, p, N, 3)
    p.profile(True)
    out, wall_ms = _run(amd, p, N, calls)
    prof = p.profile_read()
    p.profile(False)
    assert set(prof) == {"k_up", "k_m2l", "k_down"}, prof
    for k, (ms, c) in prof.items():
        assert c == calls, (k, c)
        assert ms > 0, k
    assert sum(ms for ms, _ in prof.values()) < wall_ms
    assert np.array_equal(out, ref)


def test_profile_period(amd):
    N = 1 << 16
    announce(f"profile period N={N}")
    p = amd.FMMPlan(N, 64, 18, 1)
    p.profile(4)
    _run(amd, p, N, 12, direction=1)
    prof = p.profile_read()
    p.profile(False)
    assert {k: c for k, (_, c) in prof.items()} == {"k_up": 3, "k_m2l": 3, "k_down": 3}


def test_profile_split_lines(amd, monkeypatch):
    """Lines of a deep tree through the split downward pass (L2C_DOWN_SPLIT=2
    at plan creation): k_band and k_far each count one launch per call, and
    no k_down."""
    N, nlines, calls = 1 << 18, 4, 5
    announce(f"profile split lines N={N} nlines={nlines}")
    monkeypatch.setenv("L2C_DOWN_SPLIT", "2")
    p = amd.FMMPlan(N, 64, 18, 0)
    monkeypatch.delenv("L2C_DOWN_SPLIT")
    ref, _ = _run(amd, p, N, 1, nlines=nlines)
    p.profile(True)
    out, wall_ms = _run(amd, p, N, calls, nlines=nlines)
    prof = p.profile_read()
    p.profile(False)
    assert {k: c for k, (_, c) in prof.items()} == {
        "k_up": calls, "k_m2l": calls, "k_band": calls, "k_far": calls}
    assert sum(ms for ms, _ in prof.values()) < wall_ms
    assert np.array_equal(out, ref)


def test_profile_direct_lines(amd):
    N, nlines, calls = 256, 64, 6
    announce(f"profile direct lines N={N} nlines={nlines}")
    p = amd.FMMPlan(N, 64, 18, 0)
    p.profile(True)
    _run(amd, p, N, calls, nlines=nlines)
    prof = p.profile_read()
    p.profile(False)
    assert {k: c for k, (_, c) in prof.items()} == {"k_direct": calls}
