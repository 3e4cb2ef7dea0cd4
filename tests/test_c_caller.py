"""A plain C99 program (tests/c_caller/l2c_caller.c) that calls the library
through include/leg2cheb_amd.h exactly as the reference's driver calls the
reference library (src/bin/l2c.c:181-212, :297-351).  On the CPU it must
compile warning-free as C99 and link; on the GPU its outputs must equal the
oracle's (1e-14 log2 N relative max-norm; the DCT 1e-15 log2 2N)."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR, ROOT, announce, relerr

SRC = os.path.join(ROOT, "tests", "c_caller", "l2c_caller.c")


def build(tmp_path):
    exe = str(tmp_path / "l2c_caller")
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           SRC, "-o", exe, "-L", PKG_DIR, "-lleg2cheb_amd", f"-Wl,-rpath,{PKG_DIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_header_and_caller_compile_as_c99(tmp_path, amd):
    if not os.path.exists(amd.LIB_PATH):
        pytest.skip("library not built")
    exe = build(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.gpu
@pytest.mark.parametrize("N", [1000, 4096])
def test_c_caller_matches_oracle(tmp_path, orc, N):
    exe = build(tmp_path)
    out = str(tmp_path / "out.bin")
    announce(f"c caller N={N}")
    r = subprocess.run([exe, str(N), "7", out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr
    v = np.fromfile(out)
    u, b, c, d = (v[k * N:(k + 1) * N] for k in range(4))
    o2 = v[4 * N:12 * N].reshape(8, N)
    o3 = v[12 * N:20 * N].reshape(N, 8)
    tol = 1e-14 * math.log2(N)
    op = orc.FMMPlan(N, 64, 18, orc.BOTH)
    ref = op.execute(u, 0)
    assert relerr(b, ref) < tol
    assert relerr(c, u) < 1e-12
    assert relerr(d, orc.DirectPlan(N, 0).execute(u, 0)) < tol
    line = op.execute(1.0 + np.arange(N), 0)
    for i in range(8):
        assert relerr(o2[i], line) < tol
        assert relerr(o3[:, i], line) < tol
    if N & (N - 1) == 0:
        y = v[20 * N:21 * N]
        assert relerr(y, orc.redft(u, 2)) < 1e-15 * math.log2(2 * N)
