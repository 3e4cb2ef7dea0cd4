"""The oracle pinned against the reference's own known answers and against
independent mathematics (CPU only).

Fixtures: tests/golden/reference_tables.json (LamInt, BMe/BMeT, the N=20 and
Lambda known answers, lifted as data from src/C/lambda.c:102-188,
src/C/leg2cheb.c:6-123 and src/bin/l2c.c:69-74,695-707) and
tests/golden/mp_fixtures.json (40-digit mpmath Lambda values and L2C/C2L of
seeded vectors computed from the definition, not from the Lambda closed form).
Both are written by tests/golden/make_golden.py.
"""
import json
import math
import os

import mpmath as mp
import numpy as np
import pytest

from conftest import relerr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


REF = _load("reference_tables.json")
MPF = _load("mp_fixtures.json")


def tol(N):
    return 1e-14 * max(math.log2(N), 1.0)


# ---------------------------------------------------------------- Lambda
def test_lamint_table_exact(orc):
    """LambdaI(k) reproduces the reference's exact table bit for bit."""
    t = REF["LamInt"]["values"]
    assert [orc.LambdaI(k) for k in range(256)] == t


def test_lamint_table_is_correctly_rounded(orc):
    """The table itself equals Gamma(k+1/2)/Gamma(k+1) rounded to double
    (__float128 lgamma), so the oracle's table is the mathematics, not a copy."""
    t = REF["LamInt"]["values"]
    bad = [k for k in range(256) if abs(orc.Lambda_q(k) - t[k]) > 1 * np.spacing(t[k])]
    assert not bad, bad


def test_lambda_kat(orc):
    k = REF["lambda_kat"]
    for z, v in zip(k["z"], k["values"]):
        assert abs(orc.Lambda(z) - v) < k["tol"]
    assert abs(orc.Lambda(24.5) - k["z245_exact"]) <= 2 * np.spacing(k["z245_exact"])


@pytest.mark.parametrize("i", range(len(MPF["lambda"]["z"])))
def test_lambda_vs_mpmath(orc, i):
    """Lambda(z) within 4 ulp of the 40-digit value on both sides of every
    series cut-off (14, 20, 40, 141, 2014) and in the recursion range (<10)."""
    z = MPF["lambda"]["z"][i]
    v = float(MPF["lambda"]["values"][i])
    assert abs(orc.Lambda(z) - v) <= 4 * np.spacing(v), (z, orc.Lambda(z), v)
    assert abs(orc.Lambda_q(z) - v) <= np.spacing(v)


def test_lambda_series_coefficients(orc):
    """The series coefficients c_k make Lambda(z) sqrt(z+1/4) - sum c_k w^k,
    w = (z+1/4)^-2, an O(w^7) remainder (checked at 80 digits).  A wrong or
    dropped coefficient c_k leaves an O(w^k) remainder instead."""
    c = [orc.lib().orc_lambda_series_coef(k) for k in range(7)]
    assert c[0] == 1.0
    with mp.workdps(80):
        for z in (200, 1000, 5000):
            zz = mp.mpf(z) + mp.mpf(1) / 4
            w = 1 / zz ** 2
            lam = mp.gamma(mp.mpf(z) + mp.mpf(1) / 2) / mp.gamma(mp.mpf(z) + 1)
            r = lam * mp.sqrt(zz) - mp.fsum(mp.mpf(ck) * w ** k for k, ck in enumerate(c))
            assert abs(r) < 10 * w ** 7, (z, r, w ** 7)


# ------------------------------------------------------------ B, T, DCT
def test_B0_equals_reference_tables(orc):
    B = orc.binom_B0(18)
    lo = [B[i, j] for i in range(18) for j in range(i + 1)]
    up = [B[j, i] for i in range(18) for j in range(i, 18)]
    assert lo == REF["BMe"]["values"]
    assert up == REF["BMeT"]["values"]
    assert np.all(np.triu(B, 1) == 0)


@pytest.mark.parametrize("M", [5, 18, 30, 48])
def test_B0_is_chebyshev_shift(orc, M):
    """Row k of B(0) holds the Chebyshev coefficients of T_k((x-1)/2):
    checked by evaluating both sides at random x in 50-digit arithmetic."""
    B = orc.binom_B0(M)
    rng = np.random.default_rng(M)
    with mp.workdps(50):
        for x in rng.uniform(-1, 1, 7):
            x = mp.mpf(float(x))
            for k in range(M):
                lhs = mp.chebyt(k, (x - 1) / 2)
                rhs = mp.fsum(mp.mpf(B[k, j]) * mp.chebyt(j, x) for j in range(k + 1))
                assert abs(lhs - rhs) < mp.mpf(10) ** -40 * (1 + abs(lhs)) * 2 ** M


@pytest.mark.parametrize("M", [4, 7, 18, 24])
def test_dct2_recovers_chebyshev_coefficients(orc, M):
    """dct2 of f sampled on the M x M Chebyshev-Gauss grid returns the
    coefficients of f = sum_kl c_kl T_k(x) T_l(y) for degree < M."""
    rng = np.random.default_rng(M)
    c = rng.standard_normal((M, M))
    x = np.cos(np.pi * (np.arange(M) + 0.5) / M)
    Tx = np.cos(np.outer(np.arange(M), np.arccos(x)))  # Tx[k, m] = T_k(x_m)
    F = Tx.T @ c @ Tx
    assert np.max(np.abs(orc.dct2(F) - c)) < 1e-13 * M


def test_vandermonde_is_chebyshev(orc):
    """T[p][m][k] = T_k(-1 + (2m + p)/s) (src/C/leg2cheb.c:621-648)."""
    p = orc.FMMPlan(4096, 64, 18, 0)
    T = p.T()
    s = p.s
    for par in (0, 1):
        x = -1 + (2 * np.arange(s) + par) / s
        ref = np.cos(np.outer(np.arccos(x), np.arange(18)))
        assert np.max(np.abs(T[par] - ref)) < 1e-13


# ------------------------------------------------------------- geometry
# (N, maxs) -> (L, s, Nn, Nb, Nc) from src/C/leg2cheb.c:860-871, 125-143
GEOM = [(256, 32, 1, 32, 256, 1, 3), (512, 64, 1, 64, 512, 1, 3), (4096, 64, 4, 64, 4096, 26, 78),
        (8192, 64, 5, 64, 8192, 57, 171), (65536, 64, 8, 64, 65536, 502, 1506),
        (1000, 32, 3, 32, 1024, 11, 33), (400, 64, 1, 50, 400, 1, 3)]


@pytest.mark.parametrize("N,maxs,L,s,Nn,Nb,Nc", GEOM)
def test_geometry(orc, N, maxs, L, s, Nn, Nb, Nc):
    p = orc.FMMPlan(N, maxs, 18, 0)
    assert (p.L, p.s, p.Nn) == (L, s, Nn)
    assert orc.lib().orc_total_blocks(L) == Nb
    assert orc.lib().orc_total_sub(L) == Nc


def test_direct_fallback_geometry(orc):
    assert orc.FMMPlan(256, 64, 18, 0).L == 0
    assert orc.FMMPlan(100, 32, 18, 0).L == 0


# ------------------------------------------------------- known answers
@pytest.mark.parametrize("maxs", [64, 4])
def test_kat_n20(orc, maxs):
    """src/bin/l2c.c:69-104: N=20 all ones, first five outputs."""
    k = REF["kat_n20_ones"]
    p = orc.FMMPlan(20, maxs, 18, 2)
    u = np.ones(20)
    b = p.execute(u, 0)
    c = p.execute(u, 1)
    assert np.sqrt(np.sum((b[:5] - k["l2c_first5"]) ** 2)) < k["tol_l2"]
    assert np.sqrt(np.sum((c[:5] - k["c2l_first5"]) ** 2)) < k["tol_l2"]
    # the known answers are the exact values to within rounding (direct path)
    # or to within the FMM's parity tolerance (maxs 4: L = 1, s = 4)
    t = 1e-14 if p.L == 0 else tol(20)
    assert np.max(np.abs(b[:5] - k["l2c_first5"])) < t * np.max(np.abs(b))
    assert np.max(np.abs(c[:5] - k["c2l_first5"])) < t * np.max(np.abs(c))


def test_kat_n20_direct_plan(orc):
    k = REF["kat_n20_ones"]
    d = orc.DirectPlan(20, 2)
    u = np.ones(20)
    assert np.max(np.abs(d.execute(u, 0)[:5] - k["l2c_first5"])) < 1e-14
    assert np.max(np.abs(d.execute(u, 1)[:5] - k["c2l_first5"])) < 1e-14


# ---------------------------------------------- transforms vs mpmath
CASES = list(range(len(MPF["transforms"])))


def _case(i):
    c = MPF["transforms"][i]
    return (c["N"], np.array(c["input"]), np.array([float(v) for v in c["l2c"]]),
            np.array([float(v) for v in c["c2l"]]))


@pytest.mark.parametrize("i", CASES)
def test_quad_direct_vs_mpmath(orc, i):
    N, u, b, x = _case(i)
    assert relerr(orc.leg2cheb_q(u), b) < 1e-16
    assert relerr(orc.cheb2leg_q(u), x) < 1e-16
    assert relerr(orc.leg2cheb_dd(u), b) < 4e-16
    assert relerr(orc.cheb2leg_dd(u), x) < 4e-16


@pytest.mark.parametrize("i", CASES)
def test_fp64_direct_vs_mpmath(orc, i):
    N, u, b, x = _case(i)
    d = orc.DirectPlan(N, 2)
    assert relerr(d.execute(u, 0), b) < tol(N)
    assert relerr(d.execute(u, 1), x) < tol(N)
    d1 = orc.DirectPlan(N, 2, precompute=1)
    assert relerr(d1.execute(u, 0), b) < tol(N)
    assert relerr(d1.execute(u, 1), x) < tol(N)


VARIANTS = [(8, 18, 0, 0), (16, 18, 0, 0), (4, 18, 0, 0), (8, 24, 0, 0), (8, 48, 0, 0),
            (8, 18, 1, 0), (8, 18, 0, 1), (16, 18, 1, 1), (8, 12, 0, 0), (4, 10, 1, 0)]
# The rank-M Chebyshev expansion of the kernel converges geometrically: full
# FP64 accuracy needs M >= 18; below that the error is the truncation error.
RANK_TOL = {10: 1e-8, 12: 1e-9}


@pytest.mark.parametrize("i", CASES)
@pytest.mark.parametrize("maxs,M,lag,pre", VARIANTS)
def test_fmm_vs_mpmath(orc, i, maxs, M, lag, pre):
    N, u, b, x = _case(i)
    p = orc.FMMPlan(N, maxs, M, 2, lagrange=lag, precompute=pre)
    assert p.L >= 1 or N <= 4 * maxs
    t = RANK_TOL.get(M, tol(N))
    assert relerr(p.execute(u, 0), b) < t
    assert relerr(p.execute(u, 1), x) < t
    if M < 18:  # and it is the truncation error, not a defect: it shrinks with M
        p2 = orc.FMMPlan(N, maxs, M + 4, 2, lagrange=lag, precompute=pre)
        assert relerr(p2.execute(u, 1), x) < 0.1 * max(relerr(p.execute(u, 1), x), 1e-14)


@pytest.mark.parametrize("N,maxs", [(1000, 32), (4096, 64), (12345, 64), (8191, 16)])
@pytest.mark.parametrize("direction", [0, 1])
def test_fmm_vs_double_double(orc, N, maxs, direction):
    u =np.random.default_rng(N).uniform(-1, 1, N)
    p = orc.FMMPlan(N, maxs, 18, direction)
    ref = orc.leg2cheb_dd(u) if direction == 0 else orc.cheb2leg_dd(u)
    assert relerr(p.execute(u, direction), ref) < tol(N)


# --------------------------------------------- the reference's own tests
def test_roundtrip_n1000_constant(orc):
    """src/bin/l2c.c:27-60 (-N1000 -d2): input 0.5, round trip < 1e-10."""
    p = orc.FMMPlan(1000, 32, 18, 2)
    u = np.full(1000, 0.5)
    assert np.max(np.abs(p.execute(p.execute(u, 0), 1) - u)) < 1e-10


def test_roundtrip_n100_direct(orc):
    """-N100 -d2 -v2: the direct path (maxs 32 gives L = 0)."""
    p = orc.FMMPlan(100, 32, 18, 2)
    assert p.L == 0
    u = np.full(100, 0.5)
    assert np.max(np.abs(p.execute(p.execute(u, 0), 1) - u)) < 1e-10


def test_direct_roundtrip_n100(orc):
    """src/bin/l2c.c:243-266 (-d3): direct L2C then C2L, L2 error < 1e-10."""
    d = orc.DirectPlan(100, 2)
    u = np.random.default_rng(1).uniform(-1, 1, 100)
    assert np.linalg.norm(d.execute(d.execute(u, 0), 1) - u) < 1e-10


@pytest.mark.parametrize("axis", [0, 1])
def test_2d_equals_lines(orc, axis):
    """src/bin/l2c.c:297-351 (-d4): execute2D on 400 x 402 equals the 1-D
    transform of each line (input 1 + i)."""
    N0, N1 = 400, 402
    u = (1.0 + np.arange(N0 * N1, dtype=np.float64)).reshape(N0, N1) / (N0 * N1)
    p2 = orc.FMMPlan2D(N0, N1, axis, 64, 18, 2)
    out = p2.execute(u, 0)
    n = N0 if axis == 0 else N1
    p1 = orc.FMMPlan(n, 64, 18, 2)
    lines = u.T if axis == 0 else u
    ref = np.array([p1.execute(np.ascontiguousarray(l), 0) for l in lines])
    ref = ref.T if axis == 0 else ref
    assert np.max(np.abs(out - ref)) < 1e-10 * np.max(np.abs(ref))


def test_2d_roundtrip_both_axes(orc):
    """-d5: axis -1 (both), round trip L2 < 1e-8."""
    u = np.random.default_rng(5).uniform(-1, 1, (400, 402))
    p2 = orc.FMMPlan2D(400, 402, -1, 64, 18, 2)
    assert np.linalg.norm(p2.execute(p2.execute(u, 0), 1) - u) < 1e-8


@pytest.mark.parametrize("direction", [0, 1])
def test_strided_execute(orc, direction):
    N, st = 3000, 3
    u = np.random.default_rng(2).uniform(-1, 1, N)
    p = orc.FMMPlan(N, 64, 18, direction)
    ref = p.execute(u, direction)
    big = np.zeros(N * st)
    big[::st] = u
    out = p.execute(big, direction, stride=st)
    assert np.array_equal(out[::st], ref)
    assert np.all(np.delete(out, np.arange(0, N * st, st)) == 0)


# --------------------------------------- planted mistakes are caught
def _fmm_err(orc, plan, i, direction):
    N, u, b, x = _case(i)
    return relerr(plan.execute(u, direction), b if direction == 0 else x)


@pytest.mark.parametrize("mistake", ["transpose_A_block", "swap_A_r", "T_sign", "T_index"])
def test_planted_mistakes_fail(orc, mistake):
    """Each plausible slip, planted in the oracle's tables, pushes the error
    far above the parity tolerance, so the fixtures above discriminate."""
    i = 2  # N = 200
    N = MPF["transforms"][i]["N"]
    p = orc.FMMPlan(N, 8, 18, 2)
    assert _fmm_err(orc, p, i, 0) < tol(N)
    if mistake == "transpose_A_block":
        A = p.A(0, live=True)
        A[1] = A[1].T.copy()
    elif mistake == "swap_A_r":
        A = p.A(0, live=True)
        a0 = A[0].copy()
        A[0] = A[1]
        A[1] = a0
    elif mistake == "T_sign":
        T = p.T(live=True)
        T[1, :, 3] = -T[1, :, 3]
    elif mistake == "T_index":
        T = p.T(live=True)
        T[0] = np.roll(T[0], 1, axis=0)
    assert _fmm_err(orc, p, i, 0) > 1e3 * tol(N)


def test_ofast_build_times_the_same_transform(orc):
    """The -Ofast build of the oracle's sources (timed beside the -O3 build
    for the CPU baseline, never a checker) loads, runs the same plan and
    reports a positive rate with per-call min <= average."""
    N = 4096
    u = np.random.default_rng(3).uniform(-1, 1, N)
    nt = orc.num_threads()  # both builds share one OpenMP runtime
    try:
        rate, reps, el, mn, av = orc.ofast_rate(N, 64, 18, 0, u, 0.2, threads=1)
    finally:
        orc.set_num_threads(nt)
    assert rate > 0 and reps >= 1 and el > 0
    assert 0 < mn <= av


# ------------------------------------------------------------ Table 1
@pytest.mark.parametrize("N", [256, 512, 1024, 2048, 4096, 8192])
def test_oracle_table1_envelope(orc, N):
    """The oracle reproduces the reference's published accuracy (Table 1,
    FMM_paper.pdf p.17; tests/golden/table1_published.json): maxs = 32,
    M = 18, modal and nodal, L2C and C2L, U(0,1) inputs, absolute error
    against the __float128 direct sum within the published error plus
    four ulps of ||z*||_inf."""
    from conftest import TABLE1_SLACK_ULPS, table1_error, table1_inputs, table1_published
    pub = table1_published()[N]
    u = table1_inputs(N)
    for d in (0, 1):
        ref = orc.leg2cheb_q(u) if d == 0 else orc.cheb2leg_q(u)
        for lag in (0, 1):
            g = orc.FMMPlan(N, 32, 18, d, lag).execute(u, d)
            e, ulp = table1_error(g, ref)
            assert e <= pub[(d, lag)] + TABLE1_SLACK_ULPS * ulp, (d, lag, e, e / ulp)
