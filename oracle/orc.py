"""ctypes binding to liborc_l2c.so -- the oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or the timed CPU baseline.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborc_l2c.so")

L2C, C2L, BOTH = 0, 1, 2
_dp = C.POINTER(C.c_double)
_sz = C.c_size_t


class _Direct(C.Structure):
    _fields_ = [("direction", _sz), ("N", _sz), ("a", _dp), ("an", _dp), ("dn", _dp),
                ("lf", _dp), ("lb", _dp)]


class _Plan(C.Structure):
    _fields_ = [("direction", _sz), ("M", _sz), ("N", _sz), ("Nn", _sz), ("L", _sz),
                ("s", _sz), ("lagrange", _sz), ("precompute", _sz),
                ("A", _dp * 2), ("T", _dp), ("B", _dp), ("lf", _dp), ("lb", _dp),
                ("lfoff", C.POINTER(_sz)), ("dplan", C.POINTER(_Direct)),
                ("ia", _dp), ("oa", _dp), ("wk", _dp), ("ck", _dp), ("vin", _dp)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.orc_lambda.restype = C.c_double
        L.orc_lambda.argtypes = [C.c_double]
        L.orc_lambda_q.restype = C.c_double
        L.orc_lambda_q.argtypes = [C.c_double]
        L.orc_lambda_int.restype = C.c_double
        L.orc_lambda_int.argtypes = [_sz]
        L.orc_lambda_series_coef.restype = C.c_double
        L.orc_lambda_series_coef.argtypes = [_sz]
        L.orc_lambdavec.argtypes = [_dp, _dp, _sz]
        for f in ("orc_nblocks", "orc_total_blocks", "orc_total_sub"):
            getattr(L, f).restype = _sz
            getattr(L, f).argtypes = [_sz]
        L.orc_h.restype = _sz
        L.orc_h.argtypes = [_sz, _sz]
        L.orc_binom_B0.argtypes = [_sz, _dp]
        L.orc_dct2.argtypes = [_sz, _dp, _dp]
        L.orc_redft_direct.restype = C.c_int
        L.orc_redft_direct.argtypes = [_sz, C.c_int, _dp, _dp]
        L.orc_redft_fft.restype = C.c_int
        L.orc_redft_fft.argtypes = [_sz, C.c_int, _dp, _dp]
        L.orc_create_direct.restype = C.POINTER(_Direct)
        L.orc_create_direct.argtypes = [_sz, _sz, _sz]
        L.orc_free_direct.argtypes = [C.POINTER(_Direct)]
        L.orc_direct.restype = _sz
        L.orc_direct.argtypes = [_dp, _dp, C.POINTER(_Direct), _sz, _sz]
        L.orc_create_fmm.restype = C.POINTER(_Plan)
        L.orc_create_fmm.argtypes = [_sz] * 7
        L.orc_free_fmm.argtypes = [C.POINTER(_Plan)]
        L.orc_execute.restype = _sz
        L.orc_execute.argtypes = [_dp, _dp, C.POINTER(_Plan), _sz, _sz]
        L.orc_directM.restype = _sz
        L.orc_directM.argtypes = [_dp, _dp, C.POINTER(_Plan), _sz]
        L.orc_create_fmm_2d.restype = C.c_void_p
        L.orc_create_fmm_2d.argtypes = [_sz, _sz, C.c_int, _sz, _sz, _sz, _sz, _sz, _sz]
        L.orc_free_fmm_2d.argtypes = [C.c_void_p]
        L.orc_execute2D.restype = _sz
        L.orc_execute2D.argtypes = [_dp, _dp, C.c_void_p, _sz]
        for f in ("orc_leg2cheb_q", "orc_cheb2leg_q", "orc_leg2cheb_dd", "orc_cheb2leg_dd"):
            getattr(L, f).argtypes = [_dp, _dp, _sz, _sz, _sz]
        for f in ("orc_leg2cheb_dd_rows", "orc_cheb2leg_dd_rows"):
            getattr(L, f).argtypes = [_dp, _dp, _sz, C.POINTER(_sz), _sz]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def Lambda(z):
    return lib().orc_lambda(float(z))


def LambdaI(k):
    return lib().orc_lambda_int(int(k))


def Lambda_q(z):
    return lib().orc_lambda_q(float(z))


def Lambdav(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_lambdavec(_p(x.ravel()), _p(out.ravel()), x.size)
    return out


def binom_B0(M):
    B = np.zeros((M, M))
    lib().orc_binom_B0(M, _p(B))
    return B


def dct2(F):
    F = np.ascontiguousarray(F, dtype=np.float64)
    out = np.empty_like(F)
    lib().orc_dct2(F.shape[0], _p(F), _p(out))
    return out


def redft(x, kind, method="fft"):
    """FFTW REDFT10 (kind 2) / REDFT01 (kind 3) of each row of x, in long
    double: "direct" is the O(N^2) definition, "fft" a length-4N complex FFT
    of the symmetric extension (N a power of two)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    rows = x.reshape(-1, x.shape[-1])
    out = np.empty_like(rows)
    fn = lib().orc_redft_direct if method == "direct" else lib().orc_redft_fft
    for r in range(rows.shape[0]):
        if fn(rows.shape[1], kind, _p(rows[r]), _p(out[r])) != 0:
            raise ValueError("orc_redft: bad N or kind")
    return out.reshape(x.shape)


class FMMPlan:
    """Oracle FMM plan: the reference's create_fmm/execute/free_fmm restated."""

    def __init__(self, N, maxs=64, M=18, direction=BOTH, lagrange=0, precompute=0):
        self.N = N
        self._p = lib().orc_create_fmm(N, maxs, M, direction, lagrange, precompute, 0)
        if not self._p:
            raise ValueError("orc_create_fmm failed")
        c = self._p.contents
        self.L, self.s, self.Nn, self.M = c.L, c.s, c.Nn, c.M

    def execute(self, u, direction, stride=1, out=None):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if out is None:
            out = np.zeros_like(u)
        f = lib().orc_execute(_p(u), _p(out), self._p, direction, stride)
        if f == (1 << 64) - 1:
            raise ValueError("orc_execute failed")
        self.last_flops = f
        return out

    def A(self, direction, live=False):
        """M2L matrices (Nc x M x M).  live=True returns a writable view of the
        plan's own table (tests plant mistakes through it)."""
        Nc = lib().orc_total_sub(self.L)
        v = np.ctypeslib.as_array(self._p.contents.A[direction], shape=(Nc, self.M, self.M))
        return v if live else v.copy()

    def T(self, live=False):
        v = np.ctypeslib.as_array(self._p.contents.T, shape=(2, self.s, self.M))
        return v if live else v.copy()

    def a(self):
        return np.ctypeslib.as_array(self._p.contents.dplan.contents.a, shape=(self.N,)).copy()

    def an(self):
        return np.ctypeslib.as_array(self._p.contents.dplan.contents.an, shape=(self.N,)).copy()

    def dn(self):
        return np.ctypeslib.as_array(self._p.contents.dplan.contents.dn,
                                     shape=((self.N + 1) // 2,)).copy()

    def B(self):
        """Modal plans: B(0) (M x M); nodal: the two dense child matrices
        (2 x M x M)."""
        n = 2 if self._p.contents.lagrange else 1
        return np.ctypeslib.as_array(self._p.contents.B, shape=(n, self.M, self.M)).copy()

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_free_fmm(self._p)
            self._p = None


class DirectPlan:
    def __init__(self, N, direction=BOTH, precompute=0):
        self.N = N
        self._p = lib().orc_create_direct(N, direction, precompute)

    def execute(self, u, direction, stride=1, out=None):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if out is None:
            out = np.zeros_like(u)
        lib().orc_direct(_p(u), _p(out), self._p, direction, stride)
        return out

    def tables(self):
        c = self._p.contents
        a = np.ctypeslib.as_array(c.a, shape=(self.N,)).copy()
        an = np.ctypeslib.as_array(c.an, shape=(self.N,)).copy() if c.an else None
        dn = np.ctypeslib.as_array(c.dn, shape=((self.N + 1) // 2,)).copy() if c.dn else None
        return a, an, dn

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_free_direct(self._p)
            self._p = None


class FMMPlan2D:
    def __init__(self, N0, N1, axis, maxs=64, M=18, direction=BOTH, lagrange=0, precompute=0):
        self.N0, self.N1 = N0, N1
        self._p = lib().orc_create_fmm_2d(N0, N1, axis, maxs, M, direction, lagrange, precompute, 0)

    def execute(self, u, direction):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros_like(u)
        lib().orc_execute2D(_p(u), _p(out), self._p, direction)
        return out

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_free_fmm_2d(self._p)
            self._p = None


def _rows(fn, u, r0, r1):
    u = np.ascontiguousarray(u, dtype=np.float64)
    N = u.size
    r1 = N if r1 is None else r1
    out = np.empty(r1 - r0)
    fn(_p(u), _p(out), N, r0, r1)
    return out


def leg2cheb_q(u, r0=0, r1=None):
    return _rows(lib().orc_leg2cheb_q, u, r0, r1)


def cheb2leg_q(u, r0=0, r1=None):
    return _rows(lib().orc_cheb2leg_q, u, r0, r1)


def leg2cheb_dd(u, r0=0, r1=None):
    return _rows(lib().orc_leg2cheb_dd, u, r0, r1)


def cheb2leg_dd(u, r0=0, r1=None):
    return _rows(lib().orc_cheb2leg_dd, u, r0, r1)


def direct_dd_rows(u, rows, direction):
    """Selected rows of the direct sum (double-double), direction 0 L2C / 1 C2L."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    r = np.ascontiguousarray(rows, dtype=np.uint64)
    out = np.empty(r.size)
    f = lib().orc_leg2cheb_dd_rows if direction == 0 else lib().orc_cheb2leg_dd_rows
    f(_p(u), _p(out), u.size, r.ctypes.data_as(C.POINTER(_sz)), r.size)
    return out


def ofast_rate(N, maxs, M, direction, u, seconds, threads=None):
    """Transforms/s of the same restatement built at -Ofast (liborc_l2c_ofast.so),
    for the CPU baseline only: (rate, repeats, seconds, min_s, avg_s)."""
    import time
    F = C.CDLL(os.path.join(_HERE, "liborc_l2c_ofast.so"))
    F.orc_create_fmm.restype = C.POINTER(_Plan)
    F.orc_create_fmm.argtypes = [_sz] * 7
    F.orc_free_fmm.argtypes = [C.POINTER(_Plan)]
    F.orc_execute.restype = _sz
    F.orc_execute.argtypes = [_dp, _dp, C.POINTER(_Plan), _sz, _sz]
    F.orc_num_threads.restype = C.c_int
    F.orc_set_num_threads.argtypes = [C.c_int]
    if threads:
        F.orc_set_num_threads(threads)
    p = F.orc_create_fmm(N, maxs, M, direction, 0, 0, 0)
    if not p:
        raise ValueError("orc_create_fmm failed")
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros_like(u)
    F.orc_execute(_p(u), _p(out), p, direction, 1)
    reps, per, t0 = 0, [], time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        tc = time.perf_counter()
        F.orc_execute(_p(u), _p(out), p, direction, 1)
        per.append(time.perf_counter() - tc)
        reps += 1
    el = time.perf_counter() - t0
    F.orc_free_fmm(p)
    return reps / el, reps, el, min(per), sum(per) / len(per)


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))
