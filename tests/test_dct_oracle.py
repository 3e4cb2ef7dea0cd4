"""The oracle's length-N DCT (FFTW REDFT10 / REDFT01 conventions, the
transform the reference times as its yardstick, src/bin/l2c.c:414-463).

FFTW itself is absent here, so the oracle is pinned by the definitions: its
two long-double routes (the O(N^2) sums and a length-4N FFT of the symmetric
extension) agree with each other, with closed forms that tell REDFT10 from
its transpose, with the inverse relation REDFT01(REDFT10(x)) = 2N x, and with
scipy's independent pocketfft implementation of the same definitions."""
import numpy as np
import pytest

from conftest import relerr


def cheb_basis(N, j):
    """x_n = cos(pi j (2n+1) / (2N)), the j-th Chebyshev polynomial at the
    Chebyshev-Gauss points."""
    u = (j * (2 * np.arange(N) + 1)) % (4 * N)   # exact angle index, 2 pi u / (4N)
    return np.cos(np.pi * u / (2 * N))


@pytest.mark.parametrize("N", [1, 2, 4, 8, 16, 64, 256, 1024])
@pytest.mark.parametrize("kind", [2, 3])
def test_routes_agree(orc, N, kind):
    x = np.random.default_rng(N + kind).uniform(-1, 1, N)
    assert relerr(orc.redft(x, kind, "fft"), orc.redft(x, kind, "direct")) < 2e-16


@pytest.mark.parametrize("N", [8, 64, 512])
def test_closed_forms(orc, N):
    """REDFT10 of the j-th Chebyshev basis vector is N e_j (2N e_0 for j = 0);
    REDFT01 of e_j is 2 cos(pi j (2k+1) / (2N)) (all ones for j = 0)."""
    for j in (0, 1, N // 3, N - 1):
        y = orc.redft(cheb_basis(N, j), 2, "direct")
        want = np.zeros(N)
        want[j] = 2 * N if j == 0 else N
        assert np.max(np.abs(y - want)) < 1e-13 * N
        e = np.zeros(N)
        e[j] = 1.0
        want = np.ones(N) if j == 0 else 2 * cheb_basis(N, j)
        assert np.max(np.abs(orc.redft(e, 3, "fft") - want)) < 1e-15 * np.log2(N)


def test_closed_forms_tell_the_kinds_apart(orc):
    """The closed-form check fails for the transposed operator (kind 3 where
    kind 2 is meant), so it pins which of the two the oracle computes."""
    N, j = 64, 5
    y = orc.redft(cheb_basis(N, j), 3, "direct")
    want = np.zeros(N)
    want[j] = N
    assert np.max(np.abs(y - want)) > 1.0


@pytest.mark.parametrize("N", [2, 16, 4096])
def test_inverse_relation(orc, N):
    x = np.random.default_rng(N).uniform(-1, 1, N)
    assert relerr(orc.redft(orc.redft(x, 2), 3) / (2 * N), x) < 1e-15


@pytest.mark.parametrize("N", [4, 128, 1 << 12, 1 << 16])
@pytest.mark.parametrize("kind", [2, 3])
def test_against_scipy(orc, N, kind):
    sf = pytest.importorskip("scipy.fft")
    x = np.random.default_rng(3 * N + kind).uniform(-1, 1, N)
    assert relerr(sf.dct(x, type=kind), orc.redft(x, kind)) < 1e-15 * max(1, np.log2(N))


def test_rows(orc):
    x = np.random.default_rng(7).uniform(-1, 1, (3, 32))
    y = orc.redft(x, 2)
    for r in range(3):
        assert np.array_equal(y[r], orc.redft(x[r], 2))


def test_bad_arguments(orc):
    with pytest.raises(ValueError):
        orc.redft(np.zeros(12), 2)          # the FFT route needs a power of two
    with pytest.raises(ValueError):
        orc.redft(np.zeros(8), 4)
