"""The Python routes a reference user already has, pointed at this library:
- the reference's ctypes binding (src/ctypes/l2ctypes.py:1-53) with the
  re-pointing its users make (load this library, plans as opaque pointers,
  restypes declared), then its calling sequence: zeroed output, execute,
  free (l2ctypes.py:95-113);
- this package's mirror of the Cython class and functions
  (src/cython/l2c.pyx:17-154): Leg2Cheb(input, output, M, maxs, direction),
  called per direction, and the O(N^2) leg2cheb / cheb2leg.
Checked against the oracle: 1e-14 log2 N relative max-norm."""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import announce, relerr

pytestmark = pytest.mark.gpu


def ctypes_binding(path):
    lib = C.cdll.LoadLibrary(path)
    lib.create_fmm.restype = C.c_void_p
    lib.create_fmm.argtypes = [C.c_size_t] * 7
    lib.execute.restype = C.c_size_t
    lib.execute.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t]
    lib.free_fmm.argtypes = [C.c_void_p]
    lib.create_direct.restype = C.c_void_p
    lib.create_direct.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t]
    lib.direct.restype = C.c_size_t
    lib.direct.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t]
    lib.free_direct.argtypes = [C.c_void_p]
    lib.Lambda.restype = C.c_double
    lib.Lambda.argtypes = [C.c_double]
    return lib


@pytest.mark.parametrize("N", [700, 1 << 15])
def test_reference_ctypes_sequence(amd, orc, N):
    announce(f"ctypes binding N={N}")
    lib = ctypes_binding(amd.LIB_PATH)
    u = amd.coefficients(N, seed=N)
    out = np.zeros(N)
    plan = lib.create_fmm(N, 36, 18, 2, 0, 0, 0)  # maxs=36, M=18, BOTH: the binding's defaults
    assert plan
    ref = orc.FMMPlan(N, 36, 18, orc.BOTH)
    tol = 1e-14 * math.log2(N)
    for d in (0, 1):
        out[:] = 0
        flops = lib.execute(u.ctypes.data, out.ctypes.data, plan, d, 1)
        assert flops != 2 ** 64 - 1 and flops > 0
        assert relerr(out, ref.execute(u, d)) < tol
    lib.free_fmm(plan)
    dp = lib.create_direct(N, 0, 0)
    out[:] = 0
    lib.direct(u.ctypes.data, out.ctypes.data, dp, 0, 1)
    lib.free_direct(dp)
    assert relerr(out, orc.DirectPlan(N, 0).execute(u, 0)) < tol
    assert lib.Lambda(10.5) == orc.Lambda(10.5)


def test_cython_mirror(amd, orc):
    N = 5000
    announce(f"Leg2Cheb class N={N}")
    u = amd.coefficients(N, seed=3, decay=1.0)
    inp, outp = np.zeros(N), np.zeros(N)
    t = amd.Leg2Cheb(inp, outp, M=18, maxs=36, direction=amd.BOTH, verbose=0)
    ref = orc.FMMPlan(N, 36, 18, orc.BOTH)
    tol = 1e-14 * math.log2(N)
    c = t(u, direction=0).copy()
    assert relerr(c, ref.execute(u, 0)) < tol
    assert t.flops > 0
    back = t(c, direction=1)
    assert relerr(back, u) < 1e-12
    assert t.info()["N"] == N
    # the O(N^2) functions
    o = np.zeros(N)
    assert relerr(amd.leg2cheb(u, o), orc.DirectPlan(N, 0).execute(u, 0)) < tol
    o2 = np.zeros(N)
    assert relerr(amd.cheb2leg(c, o2), orc.DirectPlan(N, 1).execute(c, 1)) < tol
