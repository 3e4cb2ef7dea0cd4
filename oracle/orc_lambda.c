/*
 * orc_lambda.c -- oracle: Lambda(z) = Gamma(z+1/2)/Gamma(z+1) in double, and
 * the integer tables a[k], an[k], dn[k] of the reference's direct plan.
 * TEST INFRASTRUCTURE ONLY (see orc_l2c.h).  Compiled with -ffp-contract=off
 * so that the tables are reproducible bit for bit by the device code, which
 * evaluates the same expressions with contraction switched off.
 */
#include <math.h>
#include <pthread.h>
#include <quadmath.h>
#include <stdlib.h>

#include "orc_l2c.h"

/* Coefficients of Lambda(z) = (z+1/4)^(-1/2) sum_k c_k (z+1/4)^(-2k), the
 * Stirling-series expansion of Gamma(t+1/4)/Gamma(t+3/4), t = z+1/4
 * (FMM_paper.pdf eq. 4.8; the same constants as src/C/lambda.c:77-93).
 * tests/test_oracle.py::test_lambda_series_coefficients checks them against
 * the Gamma ratio at 80 digits (an O(w^7) remainder). */
static const double LCOEF[7] = {1.0,
                                -1.0 / 64.0,
                                21.0 / 8192.0,
                                -671.0 / 524288.0,
                                180323.0 / 134217728.0,
                                -20898423.0 / 8589934592.0,
                                7426362705.0 / 1099511627776.0};

double orc_lambda_series_coef(size_t k) { return k < 7 ? LCOEF[k] : 0.0; }

/* Series branch, z >= 10, with the reference's term cut-offs
 * (src/C/lambda.c:76-97: one correction above 2014, two above 141, three
 * above 40, four above 20, five above 14, six otherwise). */
static double lam_series(double z) {
  const double zp = z + 0.25;
  const double sc = 1 / sqrt(zp);
  const double w = 1 / (zp * zp);
  double y = 1 + LCOEF[1] * w;
  double p;
  if (z > 2014) return y * sc;
  p = w * w;
  y += LCOEF[2] * p;
  if (z > 141) return y * sc;
  p *= w;
  y += LCOEF[3] * p;
  if (z > 40) return y * sc;
  p *= w;
  y += LCOEF[4] * p;
  if (z > 20) return y * sc;
  p *= w;
  y += LCOEF[5] * p;
  if (z > 14) return y * sc;
  p *= w;
  y += LCOEF[6] * p;
  return y * sc;
}

/* src/C/lambda.c:65-98.  For z < 10 the reference recurses
 * Lambda(z) = Lambda(z+1)(z+1)/(z+1/2); here the same chain is unrolled
 * (arguments formed by repeated +1, products applied innermost first). */
double orc_lambda(double z) {
  double zs[64];
  int n = 0;
  while (z < 10 && n < 64) {
    zs[n++] = z;
    z = z + 1;
  }
  double v = lam_series(z);
  for (int k = n - 1; k >= 0; k--) v = v * (zs[k] + 1) / (zs[k] + 0.5);
  return v;
}

void orc_lambdavec(const double *z, double *w, size_t n) {
  for (size_t i = 0; i < n; i++) w[i] = orc_lambda(z[i]);
}

/* Exact Lambda(k), k < 256, correctly rounded from __float128:
 * Lambda(0) = sqrt(pi), Lambda(k) = Lambda(k-1)(2k-1)/(2k).  Plays the role
 * of the reference's LamInt table (src/C/lambda.c:102-188); the test suite
 * checks every entry against that table. */
static double LAMINT[256];
static pthread_once_t lamint_once = PTHREAD_ONCE_INIT;
static void lamint_init(void) {
  __float128 v = sqrtq(M_PIq);
  LAMINT[0] = (double)v;
  for (int k = 1; k < 256; k++) {
    v = v * (__float128)(2 * k - 1) / (__float128)(2 * k);
    LAMINT[k] = (double)v;
  }
}

double orc_lambda_int(size_t k) {
  pthread_once(&lamint_once, lamint_init);
  if (k > 255) return orc_lambda((double)k);
  return LAMINT[k];
}

/* a[k] = Lambda(k), an[k], dn[k]: src/C/leg2cheb.c:758-795.  Below 256 the
 * exact table; beyond, every 8th value from the series and the ones between
 * by a[j] = a[j-1] - a[j-1]/2/j. */
void orc_fill_a(double *a, size_t N) {
  size_t N0 = N < 256 ? N : 256;
  for (size_t i = 0; i < N0; i++) a[i] = orc_lambda_int(i);
  for (size_t i = N0; i < N; i += 8) {
    a[i] = orc_lambda_int(i);
    size_t J0 = i + 8 < N ? i + 8 : N;
    for (size_t j = i + 1; j < J0; j++) a[j] = a[j - 1] - 0.5 * a[j - 1] / j;
  }
}

void orc_fill_an_dn(const double *a, double *an, double *dn, size_t N) {
  dn[0] = 0;
  an[0] = M_2_SQRTPI;
  for (size_t i = 1; i < N; i++) an[i] = 1 / (a[i] * (double)(2 * i * i + i));
  for (size_t i = 1; i < (N + 1) / 2; i++) dn[i] = a[i - 1] / (double)(2 * i);
}

/* Lambda(x) from lgammaq in __float128, rounded to double. */
double orc_lambda_q(double x) {
  __float128 z = x;
  return (double)expq(lgammaq(z + 0.5Q) - lgammaq(z + 1));
}
