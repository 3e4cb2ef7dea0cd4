/* orc_dct.c -- oracle (test infrastructure only): the length-N DCT-II and
 * DCT-III with FFTW's REDFT10 / REDFT01 conventions, the transform the
 * reference times as its yardstick (src/bin/l2c.c:414-463, FFTW_REDFT10; FFTW
 * itself is a third-party library, version unpinned in the reference's
 * l2copenblas.yml, absent here).  Restated from the published definitions:
 *   REDFT10: y_k = 2 sum_{n<N} x_n cos(pi k (2n+1) / (2N))
 *   REDFT01: y_k = x_0 + 2 sum_{1<=n<N} x_n cos(pi n (2k+1) / (2N))
 * Two independent routes, both in long double:
 *   orc_redft_direct -- the O(N^2) sums with every cosine index reduced
 *                       exactly modulo 4N;
 *   orc_redft_fft    -- a length-4N complex FFT of the even/odd symmetric
 *                       extension of x (any N that is a power of two).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef long double ld;

static const ld kPi = 3.141592653589793238462643383279502884L;

/* cos and sin of 2 pi u / R for R a multiple of 4, reduced to [0, pi/4]. */
static void root_ld(size_t u, size_t R, ld *c, ld *s) {
  int neg_s = 0, neg_c = 0, swap = 0;
  u %= R;
  if (2 * u > R) { u = R - u; neg_s = 1; }
  if (4 * u > R) { u = R / 2 - u; neg_c = 1; }
  if (8 * u > R) { u = R / 4 - u; swap = 1; }
  ld th = 2 * kPi * (ld)u / (ld)R, cc = cosl(th), ss = sinl(th);
  if (swap) { ld t = cc; cc = ss; ss = t; }
  if (neg_c) cc = -cc;
  if (neg_s) ss = -ss;
  *c = cc;
  *s = ss;
}

int orc_redft_direct(size_t N, int kind, const double *in, double *out) {
  if (N == 0 || (kind != 2 && kind != 3)) return -1;
  const size_t R = 4 * N;
  ld *cs = malloc(R * sizeof(ld));
  if (!cs) return -1;
  for (size_t u = 0; u < R; ++u) {
    ld c, s;
    root_ld(u, R, &c, &s);
    cs[u] = c;
  }
  for (size_t k = 0; k < N; ++k) {
    ld acc = 0;
    if (kind == 2) {
      for (size_t n = 0; n < N; ++n) acc += (ld)in[n] * cs[(k * (2 * n + 1)) % R];
      out[k] = (double)(2 * acc);
    } else {
      for (size_t n = 1; n < N; ++n) acc += (ld)in[n] * cs[(n * (2 * k + 1)) % R];
      out[k] = (double)((ld)in[0] + 2 * acc);
    }
  }
  free(cs);
  return 0;
}

/* In-place iterative radix-2 FFT, X_k = sum_u x_u e^{-2 pi i u k / n}. */
static void fft_ld(long double complex *a, size_t n) {
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      long double complex t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const size_t h = len / 2;
    for (size_t k = 0; k < h; ++k) {
      ld c, s;
      root_ld(k, len, &c, &s);
      const long double complex w = c - s * I;
      for (size_t i = k; i < n; i += len) {
        const long double complex t = w * a[i + h];
        a[i + h] = a[i] - t;
        a[i] += t;
      }
    }
  }
}

int orc_redft_fft(size_t N, int kind, const double *in, double *out) {
  if (N == 0 || (N & (N - 1)) || (kind != 2 && kind != 3)) return -1;
  const size_t R = 4 * N;
  long double complex *g = calloc(R, sizeof(long double complex));
  if (!g) return -1;
  if (kind == 2) {
    /* x_n at the odd positions 2n+1 and 4N-(2n+1): y_k = Re G_k */
    for (size_t n = 0; n < N; ++n) {
      g[2 * n + 1] = in[n];
      g[R - 2 * n - 1] = in[n];
    }
    fft_ld(g, R);
    for (size_t k = 0; k < N; ++k) out[k] = (double)creall(g[k]);
  } else {
    /* x_0 at 0, x_n at n and 4N-n: y_k = Re G_{2k+1} */
    g[0] = in[0];
    for (size_t n = 1; n < N; ++n) {
      g[n] = in[n];
      g[R - n] = in[n];
    }
    fft_ld(g, R);
    for (size_t k = 0; k < N; ++k) out[k] = (double)creall(g[2 * k + 1]);
  }
  free(g);
  return 0;
}
