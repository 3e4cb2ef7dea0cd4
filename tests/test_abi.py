"""The C-ABI library: it loads, exports every function include/leg2cheb_amd.h
declares, and its host-side helpers (no GPU needed) agree with the oracle and
the reference's tables.  Error paths that need no device run here too."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "leg2cheb_amd.h")
GOLD = os.path.join(ROOT, "tests", "golden", "reference_tables.json")


def declared_functions():
    """Names of the functions declared in the public header."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"typedef\s+struct\s*\{.*?\}\s*\w+\s*;", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text)
    return sorted({n for n in names if n not in ("if", "while", "for", "sizeof")})


def test_header_parses():
    names = declared_functions()
    for n in ("create_fmm", "execute", "free_fmm", "create_direct", "direct", "free_direct",
              "create_fmm_2d", "execute2D", "free_fmm_2d", "Lambda", "LambdaI", "Lambdavec",
              "l2c_execute_device", "l2c_execute_lines_device", "l2c_last_error"):
        assert n in names


def test_library_exports_every_declared_function(amd):
    lib = amd.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_exports_only_the_interface(amd):
    """Internal C++ symbols stay hidden: every exported function is one the
    header declares."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    extra = sorted(exported - set(declared_functions()))
    assert not extra, extra


def test_lambda_host_functions(amd, orc):
    lib = amd.lib()
    ref = json.load(open(GOLD))
    assert [lib.LambdaI(k) for k in range(256)] == ref["LamInt"]["values"]
    z = np.concatenate([np.linspace(0, 30, 241), [140.5, 141.5, 2013.5, 2014.5, 1e6]])
    ours = np.array([lib.Lambda(float(v)) for v in z])
    assert np.array_equal(ours, np.array([orc.Lambda(v) for v in z]))
    w = amd.Lambdav(z)
    assert np.array_equal(w, ours)
    assert abs(lib._Lambda0(24.5) - ref["lambda_kat"]["z245_exact"]) < 1e-14
    for zz, v in zip(ref["lambda_kat"]["z"], ref["lambda_kat"]["values"]):
        assert abs(lib.Lambda(zz) - v) < ref["lambda_kat"]["tol"]


def test_geometry_helpers(amd, orc):
    lib = amd.lib()
    for L in range(1, 14):
        assert lib.get_total_number_of_blocks(L) == orc.lib().orc_total_blocks(L)
        assert lib.get_total_number_of_submatrices(L) == orc.lib().orc_total_sub(L)
    for lev in range(12):
        assert lib.get_number_of_blocks(lev) == orc.lib().orc_nblocks(lev)
        assert lib.get_number_of_submatrices(lev) == 3 * orc.lib().orc_nblocks(lev)
        assert lib.get_h(lev, 12) == orc.lib().orc_h(lev, 12)


def test_matvectri_uses_reference_B(amd):
    """matvectri with the compact triangles (src/C/leg2cheb.c:537-619): lower
    gives B(0) x0 + B(1) x1 for two children back to back; upper adds
    [B(0)^T x; B(1)^T x] into b.  B(1) is B(0) with its odd diagonals negated."""
    lib = amd.lib()
    ref = json.load(open(GOLD))
    m = 18
    B0 = np.zeros((m, m))
    B0[np.tril_indices(m)] = ref["BMe"]["values"]
    i, j = np.indices((m, m))
    B1 = np.where((i + j) % 2 == 1, -B0, B0)
    rng = np.random.default_rng(0)
    x2 = rng.standard_normal(2 * m)
    lo = np.ascontiguousarray(ref["BMe"]["values"], dtype=np.float64)
    out = np.zeros(m)
    lib.matvectri(lo.ctypes.data, x2.ctypes.data, out.ctypes.data, m, False)
    assert np.max(np.abs(out - (B0 @ x2[:m] + B1 @ x2[m:]))) < 1e-14
    x = rng.standard_normal(m)
    up = np.ascontiguousarray(ref["BMeT"]["values"], dtype=np.float64)
    b = rng.standard_normal(2 * m)
    want = b + np.concatenate([B0.T @ x, B1.T @ x])
    lib.matvectri(up.ctypes.data, x.ctypes.data, b.ctypes.data, m, True)
    assert np.max(np.abs(b - want)) < 1e-14


def test_dct2_host(amd, orc):
    lib = amd.lib()
    F = np.random.default_rng(3).standard_normal((18, 18))
    out = np.zeros((18, 18))
    lib.dct2(F.ctypes.data, out.ctypes.data)
    assert np.max(np.abs(out - orc.dct2(F))) < 1e-14


def test_invalid_arguments_fail_cleanly(amd):
    """Bad plan parameters return NULL with a last-error text, never abort
    (the reference asserts or dereferences NULL: src/C/leg2cheb.c:1309-1312)."""
    lib = amd.lib()
    for args in ((1000, 64, 1, 0, 0, 0, 0),      # M below the supported range
                 (1000, 64, 18, 3, 0, 0, 0),     # unknown direction
                 (0, 64, 18, 0, 0, 0, 0)):       # empty transform
        assert not lib.create_fmm(*args)
        assert lib.l2c_last_error()
    assert lib.execute(None, None, None, 0, 1) == (1 << 64) - 1
    assert lib.l2c_last_error()


def test_batch_and_precision_arguments_fail_cleanly(amd):
    """The extension entry points check their arguments before touching the
    device: NULL or -1 / 0 with a last-error text."""
    import ctypes as C
    lib = amd.lib()
    n = (C.c_size_t * 3)(100, 0, 5000)
    for args in ((None, 3, 18, 0, 1),   # no lengths
                 (n, 0, 18, 0, 1),      # empty batch
                 (n, 3, 18, 0, 1),      # a zero length
                 ((C.c_size_t * 1)(100), 1, 1, 0, 1),    # rank below the supported range
                 ((C.c_size_t * 1)(100), 1, 18, 3, 1)):  # unknown direction
        assert not lib.l2c_create_batch(*args)
        assert lib.l2c_last_error()
    assert lib.l2c_execute_batch_device(None, None, None, 0, None) == -1
    assert lib.l2c_execute_batch(None, None, None, 0) == -1
    assert lib.l2c_batch_num_classes(None) == 0
    assert lib.l2c_batch_get_class(None, 0, None) == -1
    assert lib.l2c_batch_graph(None, 1) == -1
    assert lib.l2c_set_precision(None, 1) == -1
    assert lib.l2c_last_error()
    for N, count in ((0, 1), (12, 1), (64, 0), (1 << 26, 1)):
        assert not lib.l2c_create_dct(N, count)
        assert lib.l2c_last_error()
    assert lib.l2c_execute_dct_device(None, None, None, 2, None) == -1
    assert lib.l2c_execute_dct(None, None, None, 2) == -1
    assert lib.l2c_dct_synchronize(None) == -1


def test_plan_wrapper_reports_errors(amd):
    with pytest.raises(amd.L2CError):
        amd.FMMPlan(1000, 64, 1, 0)
