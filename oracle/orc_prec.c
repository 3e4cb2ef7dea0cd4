/*
 * orc_prec.c -- oracle: extended-precision direct O(N^2) sums, the role of the
 * reference's multiprecision oracle (src/multiprec/C++/L2Cprec.cpp:19-69 with
 * Boost cpp_dec_float<32>; src/multiprec/python/L2Cprecision.py:10-31 with
 * mpmath).  Boost and mpmath-in-C are absent, so:
 *   _q   : __float128 (113-bit significand, ~34 digits) everywhere;
 *   _dd  : Lambda tables from __float128, products and sums in double-double
 *          (~106 bits), fast enough for whole vectors at N = 2^16 and for
 *          sampled rows at N = 2^20.
 * TEST INFRASTRUCTURE ONLY (see orc_l2c.h).
 */
#include <math.h>
#include <quadmath.h>
#include <stdlib.h>

#include "orc_l2c.h"

/* Lambda(k), k < n, by Lambda(k) = Lambda(k-1)(2k-1)/(2k) from sqrt(pi). */
static __float128 *lam_table_q(size_t n) {
  __float128 *t = (__float128 *)malloc((n ? n : 1) * sizeof(__float128));
  __float128 v = sqrtq(M_PIq);
  for (size_t k = 0; k < n; k++) {
    if (k > 0) v = v * (__float128)(2 * k - 1) / (__float128)(2 * k);
    t[k] = v;
  }
  return t;
}

/* L2C: b_i = (2/pi) c_i sum_{n even, i+n<N} Lambda(n/2) Lambda(n/2+i) u_{n+i},
 * c_0 = 1/2 (L2Cprec.cpp:19-36). */
void orc_leg2cheb_q(const double *u, double *b, size_t N, size_t r0, size_t r1) {
  __float128 *a = lam_table_q(N);
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t i = r0; i < r1; i++) {
    __float128 acc = 0;
    for (size_t n = 0; i + n < N; n += 2) acc += a[n / 2] * a[n / 2 + i] * (__float128)u[n + i];
    acc *= 2 / M_PIq;
    if (i == 0) acc /= 2;
    b[i - r0] = (double)acc;
  }
  free(a);
}

/* C2L: v_0 = u_0, v_n = n u_n;
 * b_n = (n+1/2)[sqrt(pi) an_n v_n - sum_{m even >= 2} dn_{m/2} an_{n+m/2} v_{n+m}],
 * an_0 = 2/sqrt(pi), an_n = 1/(2 Lambda(n) n (n+1/2)), dn_j = Lambda(j-1)/(2j)
 * (L2Cprec.cpp:38-69). */
void orc_cheb2leg_q(const double *u, double *b, size_t N, size_t r0, size_t r1) {
  __float128 *a = lam_table_q(N);
  __float128 *an = (__float128 *)malloc(N * sizeof(__float128));
  __float128 *dn = (__float128 *)malloc((N / 2 + 2) * sizeof(__float128));
  an[0] = 2 / sqrtq(M_PIq);
  for (size_t i = 1; i < N; i++) an[i] = 1 / (2 * a[i] * (__float128)i * ((__float128)i + 0.5Q));
  dn[0] = 0;
  for (size_t j = 1; j < N / 2 + 2; j++) dn[j] = a[j - 1 < N ? j - 1 : 0] / (2 * (__float128)j);
  const __float128 sp = sqrtq(M_PIq);
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t n = r0; n < r1; n++) {
    __float128 vn = n == 0 ? (__float128)u[0] : (__float128)u[n] * (__float128)n;
    __float128 acc = sp * an[n] * vn;
    for (size_t m = 2; n + m < N; m += 2)
      acc -= dn[m / 2] * an[m / 2 + n] * ((__float128)u[n + m] * (__float128)(n + m));
    acc *= ((__float128)n + 0.5Q);
    b[n - r0] = (double)acc;
  }
  free(a);
  free(an);
  free(dn);
}

/* ---- double-double ------------------------------------------------------- */
typedef struct {
  double hi, lo;
} dd;
static inline dd dd_from_q(__float128 q) {
  dd r;
  r.hi = (double)q;
  r.lo = (double)(q - (__float128)r.hi);
  return r;
}
static inline dd fast2(double a, double b) {
  dd r;
  r.hi = a + b;
  r.lo = b - (r.hi - a);
  return r;
}
static inline dd dd_add(dd a, dd b) {
  double s = a.hi + b.hi;
  double bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (b.hi - bb);
  e += a.lo + b.lo;
  return fast2(s, e);
}
static inline dd dd_mul(dd a, dd b) {
  double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  e += a.hi * b.lo + a.lo * b.hi;
  return fast2(p, e);
}
static inline dd dd_mul_d(dd a, double b) {
  double p = a.hi * b;
  double e = fma(a.hi, b, -p);
  e += a.lo * b;
  return fast2(p, e);
}
static inline double dd_to_d(dd a) { return a.hi + a.lo; }

/* rows[k] (k < nr) of the L2C / C2L direct sums in double-double. */
static void l2c_dd_rows(const double *u, double *b, size_t N, const size_t *rows, size_t nr) {
  __float128 *aq = lam_table_q(N);
  dd *a = (dd *)malloc((N ? N : 1) * sizeof(dd));
  for (size_t k = 0; k < N; k++) a[k] = dd_from_q(aq[k]);
  const dd twopi = dd_from_q(2 / M_PIq);
#pragma omp parallel for schedule(dynamic, 4)
  for (size_t k = 0; k < nr; k++) {
    const size_t i = rows[k];
    dd acc = {0, 0};
    for (size_t n = 0; i + n < N; n += 2) acc = dd_add(acc, dd_mul_d(dd_mul(a[n / 2], a[n / 2 + i]), u[n + i]));
    acc = dd_mul(acc, twopi);
    if (i == 0) acc = dd_mul_d(acc, 0.5);
    b[k] = dd_to_d(acc);
  }
  free(aq);
  free(a);
}

static void c2l_dd_rows(const double *u, double *b, size_t N, const size_t *rows, size_t nr);

void orc_leg2cheb_dd(const double *u, double *b, size_t N, size_t r0, size_t r1) {
  size_t *rows = (size_t *)malloc((r1 > r0 ? r1 - r0 : 1) * sizeof(size_t));
  for (size_t i = r0; i < r1; i++) rows[i - r0] = i;
  l2c_dd_rows(u, b, N, rows, r1 - r0);
  free(rows);
}

void orc_leg2cheb_dd_rows(const double *u, double *b, size_t N, const size_t *rows, size_t nr) {
  l2c_dd_rows(u, b, N, rows, nr);
}

void orc_cheb2leg_dd_rows(const double *u, double *b, size_t N, const size_t *rows, size_t nr) {
  c2l_dd_rows(u, b, N, rows, nr);
}

void orc_cheb2leg_dd(const double *u, double *b, size_t N, size_t r0, size_t r1) {
  size_t *rows = (size_t *)malloc((r1 > r0 ? r1 - r0 : 1) * sizeof(size_t));
  for (size_t i = r0; i < r1; i++) rows[i - r0] = i;
  c2l_dd_rows(u, b, N, rows, r1 - r0);
  free(rows);
}

static void c2l_dd_rows(const double *u, double *b, size_t N, const size_t *rows, size_t nr) {
  __float128 *aq = lam_table_q(N);
  dd *an = (dd *)malloc((N ? N : 1) * sizeof(dd));
  dd *dn = (dd *)malloc((N / 2 + 2) * sizeof(dd));
  an[0] = dd_from_q(2 / sqrtq(M_PIq));
  for (size_t i = 1; i < N; i++) an[i] = dd_from_q(1 / (2 * aq[i] * (__float128)i * ((__float128)i + 0.5Q)));
  dn[0] = dd_from_q(0);
  for (size_t j = 1; j < N / 2 + 2; j++) dn[j] = dd_from_q(aq[j - 1 < N ? j - 1 : 0] / (2 * (__float128)j));
  const dd sp = dd_from_q(sqrtq(M_PIq));
#pragma omp parallel for schedule(dynamic, 4)
  for (size_t k = 0; k < nr; k++) {
    const size_t n = rows[k];
    /* v_n = n u_n is exact in double-double */
    dd vn = n == 0 ? (dd){u[0], 0} : dd_mul_d((dd){u[n], 0}, (double)n);
    dd acc = dd_mul(dd_mul(sp, an[n]), vn);
    for (size_t m = 2; n + m < N; m += 2) {
      dd v = dd_mul_d((dd){u[n + m], 0}, (double)(n + m));
      dd t = dd_mul(dd_mul(dn[m / 2], an[m / 2 + n]), v);
      t.hi = -t.hi;
      t.lo = -t.lo;
      acc = dd_add(acc, t);
    }
    acc = dd_mul_d(acc, (double)n + 0.5);
    b[k] = dd_to_d(acc);
  }
  free(aq);
  free(an);
  free(dn);
}
