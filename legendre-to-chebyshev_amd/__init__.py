"""Python host interface of the gfx950 FMM Legendre<->Chebyshev library.

Mirrors the reference's Python bindings (src/cython/l2c.pyx:17-173 and
src/ctypes/l2ctypes.py:53-113): a ``Leg2Cheb`` plan class called with
``direction`` 0 (L2C) or 1 (C2L), the O(N^2) ``leg2cheb``/``cheb2leg`` and
``Lambdav``.  Everything runs through the C ABI of libleg2cheb_amd.so
(include/leg2cheb_amd.h); there is no CPU fallback -- the loader raises if the
library is missing.

The directory name carries hyphens, so load it under a module name of your
choice, e.g.::

    spec = importlib.util.spec_from_file_location(
        "l2c_amd", ".../legendre-to-chebyshev_amd/__init__.py",
        submodule_search_locations=[".../legendre-to-chebyshev_amd"])
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libleg2cheb_amd.so")

L2C, C2L, BOTH = 0, 1, 2
_dp = C.POINTER(C.c_double)
_sz = C.c_size_t
_SIZE_MAX = (1 << 64) - 1


class L2CError(RuntimeError):
    pass


class PlanInfo(C.Structure):
    _fields_ = [(n, _sz) for n in ("N", "Nn", "L", "s", "M", "maxs", "direction", "lagrange",
                                   "Nc", "Nb", "hbm_bytes", "a_bytes")] + [
        ("subtree_up", C.c_int), ("subtree_down", C.c_int)]


class BatchClass(C.Structure):
    _fields_ = [("n_max", _sz), ("count", _sz), ("method", C.c_int), ("s", _sz), ("L", _sz),
                ("ms", C.c_double), ("variant", C.c_int), ("cached", C.c_int)]


_lib = None


def lib():
    """The loaded C library (raises L2CError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise L2CError(f"{LIB_PATH} is missing: build it with `make -C legendre-to-chebyshev_amd`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.create_fmm.restype = vp
        L.create_fmm.argtypes = [_sz] * 7
        L.execute.restype = _sz
        L.execute.argtypes = [vp, vp, vp, _sz, _sz]
        L.free_fmm.argtypes = [vp]
        L.create_direct.restype = vp
        L.create_direct.argtypes = [_sz, _sz, _sz]
        L.direct.restype = _sz
        L.direct.argtypes = [vp, vp, vp, _sz, _sz]
        L.free_direct.argtypes = [vp]
        L.directM.restype = _sz
        L.directM.argtypes = [vp, vp, vp, _sz]
        L.create_fmm_2d.restype = vp
        L.create_fmm_2d.argtypes = [_sz, _sz, C.c_int, _sz, _sz, _sz, _sz, _sz, _sz]
        L.execute2D.restype = _sz
        L.execute2D.argtypes = [vp, vp, vp, _sz]
        L.free_fmm_2d.argtypes = [vp]
        L.Lambda.restype = C.c_double
        L.Lambda.argtypes = [C.c_double]
        L._Lambda0.restype = C.c_double
        L._Lambda0.argtypes = [C.c_double]
        L.LambdaI.restype = C.c_double
        L.LambdaI.argtypes = [_sz]
        L.Lambdavec.argtypes = [vp, vp, _sz]
        L.matvectri.argtypes = [vp, vp, vp, _sz, C.c_bool]
        L.dct2.argtypes = [vp, vp]
        for f in ("get_number_of_blocks", "get_total_number_of_blocks",
                  "get_number_of_submatrices", "get_total_number_of_submatrices"):
            getattr(L, f).restype = _sz
            getattr(L, f).argtypes = [_sz]
        L.get_h.restype = _sz
        L.get_h.argtypes = [_sz, _sz]
        L.get_ij.argtypes = [C.POINTER(_sz), _sz, _sz, _sz, _sz]
        L.l2c_last_error.restype = C.c_char_p
        L.l2c_get_plan_info.restype = C.c_int
        L.l2c_get_plan_info.argtypes = [vp, C.POINTER(PlanInfo)]
        L.l2c_get_plan_table.restype = _sz
        L.l2c_get_plan_table.argtypes = [vp, C.c_int, _sz, vp, _sz]
        L.l2c_execute_device.restype = C.c_int
        L.l2c_execute_device.argtypes = [vp, vp, vp, _sz, _sz, vp]
        L.l2c_execute_lines_device_noalloc.restype = C.c_int
        L.l2c_execute_lines_device_noalloc.argtypes = [vp, vp, vp, _sz, _sz, _sz, _sz, _sz, _sz, vp]
        L.l2c_plan_graph.restype = C.c_int
        L.l2c_plan_graph.argtypes = [vp, C.c_int]
        L.l2c_execute_lines_device.restype = C.c_int
        L.l2c_execute_lines_device.argtypes = [vp, vp, vp, _sz, _sz, _sz, _sz, _sz, _sz, vp]
        L.l2c_set_precision.restype = C.c_int
        L.l2c_set_precision.argtypes = [vp, C.c_int]
        L.l2c_profile.restype = C.c_int
        L.l2c_profile.argtypes = [vp, C.c_int]
        L.l2c_profile_read.restype = C.c_int
        L.l2c_profile_read.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(_sz)]
        L.l2c_synchronize.restype = C.c_int
        L.l2c_synchronize.argtypes = [vp]
        L.l2c_malloc_device.restype = vp
        L.l2c_malloc_device.argtypes = [_sz]
        L.l2c_free_device.restype = C.c_int
        L.l2c_free_device.argtypes = [vp]
        L.l2c_memcpy_h2d.restype = C.c_int
        L.l2c_memcpy_h2d.argtypes = [vp, vp, _sz]
        L.l2c_memcpy_d2h.restype = C.c_int
        L.l2c_memcpy_d2h.argtypes = [vp, vp, _sz]
        L.l2c_memcpy_d2d.restype = C.c_int
        L.l2c_memcpy_d2d.argtypes = [vp, vp, _sz]
        L.l2c_create_batch.restype = vp
        L.l2c_create_batch.argtypes = [C.POINTER(_sz), _sz, _sz, _sz, C.c_int]
        L.l2c_free_batch.argtypes = [vp]
        L.l2c_execute_batch_device.restype = C.c_int
        L.l2c_execute_batch_device.argtypes = [vp, vp, vp, _sz, vp]
        L.l2c_execute_batch.restype = C.c_int
        L.l2c_execute_batch.argtypes = [vp, vp, vp, _sz]
        L.l2c_tune_cache.restype = C.c_int
        L.l2c_tune_cache.argtypes = [C.c_char_p]
        L.l2c_batch_graph.restype = C.c_int
        L.l2c_batch_graph.argtypes = [vp, C.c_int]
        L.l2c_batch_synchronize.restype = C.c_int
        L.l2c_batch_synchronize.argtypes = [vp]
        L.l2c_batch_num_classes.restype = _sz
        L.l2c_batch_num_classes.argtypes = [vp]
        L.l2c_batch_get_class.restype = C.c_int
        L.l2c_batch_get_class.argtypes = [vp, _sz, C.POINTER(BatchClass)]
        L.l2c_create_dct.restype = vp
        L.l2c_create_dct.argtypes = [_sz, _sz]
        L.l2c_free_dct.argtypes = [vp]
        L.l2c_execute_dct_device.restype = C.c_int
        L.l2c_execute_dct_device.argtypes = [vp, vp, vp, C.c_int, vp]
        L.l2c_execute_dct.restype = C.c_int
        L.l2c_execute_dct.argtypes = [vp, vp, vp, C.c_int]
        L.l2c_dct_synchronize.restype = C.c_int
        L.l2c_dct_synchronize.argtypes = [vp]
        L.l2c_execute_dct_lines_device.restype = C.c_int
        L.l2c_execute_dct_lines_device.argtypes = [vp, vp, vp, C.c_int, _sz, _sz, _sz, _sz, _sz, vp]
        _lib = L
    return _lib


def last_error():
    return lib().l2c_last_error().decode()


def _ptr(a):
    if not (a.dtype == np.float64 and a.flags.c_contiguous):
        raise ValueError("arrays must be C-contiguous float64")
    return a.ctypes.data


class Leg2Cheb:
    """Plan for Legendre<->Chebyshev transforms (src/cython/l2c.pyx:17-114).

    Parameters mirror the reference: input_array/output_array (float64, 1-D),
    M (rank), maxs (max leaf size), direction (0 L2C, 1 C2L, 2 both),
    lagrange, precompute, verbose.  Calls overwrite the output array.
    """

    def __init__(self, input_array, output_array, M=18, maxs=36, direction=BOTH, lagrange=0,
                 precompute=0, verbose=1):
        self.N = int(input_array.shape[0])
        self.direction = direction
        self._lib = lib()
        self._plan = self._lib.create_fmm(self.N, maxs, M, direction, lagrange, precompute, verbose)
        if not self._plan:
            raise L2CError("create_fmm: " + last_error())
        self._input_array = input_array
        self._output_array = output_array
        self.flops = 0

    def __call__(self, input_array=None, output_array=None, direction=0):
        assert direction in (0, 1), "Please specify direction of transform!"
        if input_array is not None:
            self._input_array[...] = input_array
        f = lib().execute(_ptr(self._input_array), _ptr(self._output_array), self._plan,
                          direction, 1)
        if f == _SIZE_MAX:
            raise L2CError("execute: " + last_error())
        self.flops = f
        if output_array is not None:
            output_array[...] = self._output_array
            return output_array
        return self._output_array

    @property
    def input_array(self):
        return self._input_array

    @property
    def output_array(self):
        return self._output_array

    @property
    def plan(self):
        return self._plan

    def info(self):
        inf = PlanInfo()
        if lib().l2c_get_plan_info(self._plan, C.byref(inf)) != 0:
            raise L2CError(last_error())
        return {k: getattr(inf, k) for k, _ in PlanInfo._fields_}

    def __del__(self):
        if getattr(self, "_plan", None):
            self._lib.free_fmm(self._plan)
            self._plan = None


class FMMPlan:
    """Thin wrapper over create_fmm/execute/free_fmm with explicit stride."""

    def __init__(self, N, maxs=64, M=18, direction=BOTH, lagrange=0, precompute=0, verbose=0):
        self.N = N
        self._lib = lib()
        self._plan = self._lib.create_fmm(N, maxs, M, direction, lagrange, precompute, verbose)
        if not self._plan:
            raise L2CError("create_fmm: " + last_error())

    def execute(self, u, direction, stride=1, out=None):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if out is None:
            out = np.zeros_like(u)
        f = lib().execute(_ptr(u), _ptr(out), self._plan, direction, stride)
        if f == _SIZE_MAX:
            raise L2CError("execute: " + last_error())
        self.flops = f
        return out

    def directM(self, u, stride=1):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros_like(u)
        f = lib().directM(_ptr(u), _ptr(out), self._plan, stride)
        if f == _SIZE_MAX:
            raise L2CError("directM: " + last_error())
        return out

    def execute_device(self, d_in, d_out, direction, stride=1, stream=None):
        rc = lib().l2c_execute_device(self._plan, d_in, d_out, direction, stride, stream)
        if rc != 0:
            raise L2CError("l2c_execute_device: " + last_error())

    def execute_lines_device(self, d_in, d_out, direction, nlines, n0, ls0, ls1, es, stream=None):
        rc = lib().l2c_execute_lines_device(self._plan, d_in, d_out, direction, nlines, n0, ls0,
                                            ls1, es, stream)
        if rc != 0:
            raise L2CError("l2c_execute_lines_device: " + last_error())

    def graph(self, enable=True):
        """Replay execute_device launches from cached HIP graphs (l2c_plan_graph)."""
        if lib().l2c_plan_graph(self._plan, 1 if enable else 0) != 0:
            raise L2CError(last_error())

    def set_precision(self, form):
        if lib().l2c_set_precision(self._plan, form) != 0:
            raise L2CError(last_error())

    def synchronize(self):
        if lib().l2c_synchronize(self._plan) != 0:
            raise L2CError(last_error())

    # execute kernels by profile slot
    KERNELS = ("k_up", "k_m2l", "k_down", "k_direct", "k_far", "k_band", "reserved6",
               "reserved7")

    def profile(self, enable=True):
        """Time the execute kernels with HIP events: True (every call), False
        (off) or an int n (every n-th call)."""
        n = int(enable) if not isinstance(enable, bool) else (1 if enable else 0)
        if lib().l2c_profile(self._plan, n) != 0:
            raise L2CError(last_error())

    def profile_read(self):
        ms = (C.c_double * 8)()
        cnt = (_sz * 8)()
        if lib().l2c_profile_read(self._plan, ms, cnt) != 0:
            raise L2CError(last_error())
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.KERNELS) if cnt[i]}

    def info(self):
        inf = PlanInfo()
        if lib().l2c_get_plan_info(self._plan, C.byref(inf)) != 0:
            raise L2CError(last_error())
        return {k: getattr(inf, k) for k, _ in PlanInfo._fields_}

    TABLES = ("lambda", "an", "dn", "head", "m2l", "T", "U", "D", "P", "Q")

    def table(self, name, direction=L2C):
        """A copy of one of the plan's device tables (l2c_get_plan_table), as a
        flat float64 array; see include/leg2cheb_amd.h for layouts."""
        t = self.TABLES.index(name)
        n = lib().l2c_get_plan_table(self._plan, t, direction, None, 0)
        if n == _SIZE_MAX:
            raise L2CError("l2c_get_plan_table: " + last_error())
        out = np.empty(n)
        if lib().l2c_get_plan_table(self._plan, t, direction, out.ctypes.data, n) != n:
            raise L2CError("l2c_get_plan_table: " + last_error())
        return out

    @property
    def plan(self):
        return self._plan

    def __del__(self):
        if getattr(self, "_plan", None):
            self._lib.free_fmm(self._plan)
            self._plan = None


class FMMPlan2D:
    """create_fmm_2d / execute2D / free_fmm_2d (src/C/leg2cheb.c:1236-1297)."""

    def __init__(self, N0, N1, axis, maxs=64, M=18, direction=BOTH, lagrange=0, precompute=0):
        self._lib = lib()
        self._plan = self._lib.create_fmm_2d(N0, N1, axis, maxs, M, direction, lagrange, precompute, 0)
        if not self._plan:
            raise L2CError("create_fmm_2d: " + last_error())

    def execute(self, u, direction, out=None):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if out is None:
            out = np.zeros_like(u)
        f = lib().execute2D(_ptr(u), _ptr(out), self._plan, direction)
        if f == _SIZE_MAX:
            raise L2CError("execute2D: " + last_error())
        return out

    def __del__(self):
        if getattr(self, "_plan", None):
            self._lib.free_fmm_2d(self._plan)
            self._plan = None


def _direct(input_array, output_array, direction):
    n = int(input_array.shape[0])
    p = lib().create_direct(n, direction, 0)
    if not p:
        raise L2CError("create_direct: " + last_error())
    try:
        f = lib().direct(_ptr(input_array), _ptr(output_array), p, direction, 1)
        if f == _SIZE_MAX:
            raise L2CError("direct: " + last_error())
    finally:
        lib().free_direct(p)
    return output_array


def leg2cheb(input_array, output_array):
    """O(N^2) Legendre to Chebyshev (src/cython/l2c.pyx:116-135)."""
    return _direct(input_array, output_array, L2C)


def cheb2leg(input_array, output_array):
    """O(N^2) Chebyshev to Legendre (src/cython/l2c.pyx:137-154)."""
    return _direct(input_array, output_array, C2L)


def Lambdav(x):
    """Lambda(x) = Gamma(x+1/2)/Gamma(x+1) elementwise (src/cython/l2c.pyx:156-173)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    a = np.empty(x.size)
    lib().Lambdavec(x.ctypes.data, a.ctypes.data, x.size)
    return a.reshape(np.shape(x))


def Lambda(z):
    return lib().Lambda(float(z))


def LambdaI(k):
    return lib().LambdaI(int(k))


class BatchPlan:
    """Many independent transforms of uneven length (l2c_create_batch).

    Inputs and outputs are packed back to back in the order of `lengths`.
    tune=True times every candidate method per size class at plan time."""

    def __init__(self, lengths, M=18, direction=BOTH, tune=True):
        n = np.ascontiguousarray(lengths, dtype=np.uint64)
        self.lengths = n.astype(np.int64)
        self.total = int(n.sum())
        self.offsets = np.concatenate([[0], np.cumsum(self.lengths)[:-1]]).astype(np.int64)
        self._lib = lib()
        self._plan = self._lib.l2c_create_batch(n.ctypes.data_as(C.POINTER(_sz)), n.size, M,
                                                direction, 1 if tune else 0)
        if not self._plan:
            raise L2CError("l2c_create_batch: " + last_error())

    def execute(self, u, direction, out=None):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.size != self.total:
            raise ValueError(f"packed input must hold {self.total} values")
        if out is None:
            out = np.zeros_like(u)
        if lib().l2c_execute_batch(self._plan, _ptr(u), _ptr(out), direction) != 0:
            raise L2CError("l2c_execute_batch: " + last_error())
        return out

    def execute_device(self, d_in, d_out, direction, stream=None):
        if lib().l2c_execute_batch_device(self._plan, d_in, d_out, direction, stream) != 0:
            raise L2CError("l2c_execute_batch_device: " + last_error())

    def graph(self, enable=True):
        if lib().l2c_batch_graph(self._plan, 1 if enable else 0) != 0:
            raise L2CError(last_error())

    def synchronize(self):
        if lib().l2c_batch_synchronize(self._plan) != 0:
            raise L2CError(last_error())

    def classes(self):
        out = []
        for i in range(lib().l2c_batch_num_classes(self._plan)):
            c = BatchClass()
            if lib().l2c_batch_get_class(self._plan, i, C.byref(c)) != 0:
                raise L2CError(last_error())
            out.append({"n_max": c.n_max, "count": c.count,
                        "method": "direct" if c.method == 0 else "fmm", "s": c.s, "L": c.L,
                        "ms": c.ms, "variant": c.variant, "cached": bool(c.cached)})
        return out

    def split(self, packed):
        """The per-transform views of a packed array."""
        return [packed[o:o + n] for o, n in zip(self.offsets, self.lengths)]

    def __del__(self):
        if getattr(self, "_plan", None):
            self._lib.l2c_free_batch(self._plan)
            self._plan = None


class DCTPlan:
    """Length-N DCT-II (kind 2, FFTW REDFT10) and DCT-III (kind 3, FFTW REDFT01)
    of `count` real vectors stored back to back (l2c_create_dct).  Unnormalised
    as FFTW: dct(dct(x, 2), 3) = 2 N x."""

    def __init__(self, N, count=1):
        self.N, self.count = int(N), int(count)
        self._lib = lib()
        self._plan = self._lib.l2c_create_dct(self.N, self.count)
        if not self._plan:
            raise L2CError("l2c_create_dct: " + last_error())

    def execute(self, u, kind=2, out=None):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.size != self.N * self.count:
            raise ValueError(f"input must hold count * N = {self.N * self.count} values")
        if out is None:
            out = np.empty_like(u)
        if lib().l2c_execute_dct(self._plan, _ptr(u), _ptr(out), kind) != 0:
            raise L2CError("l2c_execute_dct: " + last_error())
        return out

    def execute_device(self, d_in, d_out, kind=2, stream=None):
        if lib().l2c_execute_dct_device(self._plan, d_in, d_out, kind, stream) != 0:
            raise L2CError("l2c_execute_dct_device: " + last_error())

    def synchronize(self):
        if lib().l2c_dct_synchronize(self._plan) != 0:
            raise L2CError(last_error())

    def execute_lines_device(self, d_in, d_out, kind, nlines, n0, ls0, ls1, es, stream=None):
        """Lines of length N (<= 4096) with the line map of
        FMMPlan.execute_lines_device, one launch."""
        if lib().l2c_execute_dct_lines_device(self._plan, d_in, d_out, kind, nlines, n0, ls0, ls1,
                                              es, stream) != 0:
            raise L2CError("l2c_execute_dct_lines_device: " + last_error())

    def __del__(self):
        if getattr(self, "_plan", None):
            self._lib.l2c_free_dct(self._plan)
            self._plan = None


def cheb_values(c, plan=None):
    """Values at the Chebyshev-Gauss points x_j = cos(pi (2j+1) / (2N)) of the
    Chebyshev series with coefficients c: one REDFT01 of (c_0, c_1/2, ...)."""
    c = np.array(c, dtype=np.float64)
    c[1:] *= 0.5
    plan = plan or DCTPlan(c.size)
    return plan.execute(c, 3)


def cheb_coefficients(f, plan=None):
    """Chebyshev coefficients of the interpolant of values f at the N
    Chebyshev-Gauss points: REDFT10(f) / N with c_0 halved."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    plan = plan or DCTPlan(f.size)
    c = plan.execute(f, 2) / f.size
    c[0] *= 0.5
    return c


class DeviceBuffer:
    """A device allocation made by the library (for device-resident runs)."""

    def __init__(self, n):
        self.n = n
        self._lib = lib()
        self.ptr = self._lib.l2c_malloc_device(n * 8)
        if not self.ptr:
            raise L2CError("l2c_malloc_device: " + last_error())

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size <= self.n
        if lib().l2c_memcpy_h2d(self.ptr, a.ctypes.data, a.size * 8) != 0:
            raise L2CError(last_error())

    def download(self, n=None):
        n = self.n if n is None else n
        out = np.empty(n)
        if lib().l2c_memcpy_d2h(out.ctypes.data, self.ptr, n * 8) != 0:
            raise L2CError(last_error())
        return out

    def free(self):
        if self.ptr:
            self._lib.l2c_free_device(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


def tune_cache(path):
    """Set (or with None clear) the batch tuner's cache file (l2c_tune_cache)."""
    lib().l2c_tune_cache(None if path is None else str(path).encode())


def splitmix_uniform(n, seed, lo=-1.0, hi=1.0):
    """n seeded uniforms in [lo, hi) from the SplitMix64 counter generator:
    z_i = mix(seed + (i+1) * 0x9E3779B97F4A7C15), U = (z >> 11) * 2^-53."""
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def coefficients(n, seed, decay=0.0):
    """Seeded Legendre coefficients U(-1,1) * (k+1)^-decay, k < n."""
    u = splitmix_uniform(n, seed)
    if decay:
        u = u / np.power(np.arange(1, n + 1, dtype=np.float64), decay)
    return u


def log_uniform_lengths(count, seed, lo=64, hi=8193):
    """floor(exp(U(ln lo, ln hi))) for `count` transforms (SplitMix64 `seed`)."""
    x = splitmix_uniform(count, seed, 0.0, 1.0)
    return np.floor(np.exp(np.log(lo) + x * (np.log(hi) - np.log(lo)))).astype(np.int64)


def packed_coefficients(lengths, seed, decay=1.0):
    """Packed inputs of a batch: transform t gets U(-1,1) * (k+1)^-decay, k < n_t,
    all drawn from one SplitMix64 stream."""
    lengths = np.asarray(lengths, dtype=np.int64)
    u = splitmix_uniform(int(lengths.sum()), seed)
    if decay:
        local = np.arange(u.size) - np.repeat(np.cumsum(lengths) - lengths, lengths)
        u = u / np.power(local + 1.0, decay)
    return u
