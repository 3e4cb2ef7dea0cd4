/*
 * orc_l2c.c -- the oracle: CPU restatement of the reference's FMM plan and
 * execute (src/C/leg2cheb.c), with its own dense loops in place of CBLAS and
 * FFTW.  TEST INFRASTRUCTURE ONLY (see orc_l2c.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "orc_l2c.h"

void orc_fill_a(double *a, size_t N);
void orc_fill_an_dn(const double *a, double *an, double *dn, size_t N);

static const double SQRT_PI = 1.77245385090551602729816748334e0;

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

static double *dalloc(size_t n) {
  double *p = (double *)aligned_alloc(64, ((n ? n : 1) * sizeof(double) + 63) / 64 * 64);
  return p;
}
static double *dzalloc(size_t n) {
  double *p = dalloc(n);
  memset(p, 0, (n ? n : 1) * sizeof(double));
  return p;
}

/* ---- geometry: src/C/leg2cheb.c:125-150 --------------------------------- */
size_t orc_nblocks(size_t level) { return ((size_t)2 << level) - 1; }
size_t orc_h(size_t level, size_t L) { return (size_t)1 << (L - level - 1); }
size_t orc_total_blocks(size_t L) { return ((size_t)2 << L) - (L + 2); }
size_t orc_total_sub(size_t L) { return 3 * orc_total_blocks(L); }

/* ---- exact B(0): Chebyshev coefficients of T_k((x-1)/2) -----------------
 * Row k holds the expansion of T_k((x-1)/2) in T_0..T_k.  Built with the
 * three-term recurrence P_{k+1} = (x-1) P_k - P_{k-1} and
 * x T_0 = T_1, x T_l = (T_{l+1} + T_{l-1})/2.  Every value is a dyadic
 * rational that double holds exactly for the M used here (checked against
 * exact fractions in the tests).  Same matrix as W^-1 A(1/2,-1/2) W in
 * src/python/binommat.py:33-38 and BMe in src/C/leg2cheb.c:6-63. */
void orc_binom_B0(size_t M, double *B0) {
  memset(B0, 0, M * M * sizeof(double));
  if (M == 0) return;
  B0[0] = 1.0;
  if (M == 1) return;
  B0[M + 0] = -0.5;
  B0[M + 1] = 0.5;
  double *xp = dzalloc(M + 1);
  for (size_t k = 1; k + 1 < M; k++) {
    const double *Pk = &B0[k * M];
    const double *Pm = &B0[(k - 1) * M];
    double *Pn = &B0[(k + 1) * M];
    memset(xp, 0, (M + 1) * sizeof(double));
    for (size_t l = 0; l <= k; l++) {
      if (Pk[l] == 0.0) continue;
      if (l == 0) {
        xp[1] += Pk[0];
      } else {
        xp[l + 1] += 0.5 * Pk[l];
        xp[l - 1] += 0.5 * Pk[l];
      }
    }
    for (size_t l = 0; l <= k + 1; l++) Pn[l] = xp[l] - (l <= k ? Pk[l] : 0.0) - (l < k ? Pm[l] : 0.0);
  }
  free(xp);
}

/* ---- 2-D DCT-II: src/C/dct.c:582-598 and the FFTW REDFT10 branch
 * src/C/leg2cheb.c:974-985.  out = (4/M^2) C in C^T, C[k][i] = cos(pi k (i+1/2)/M),
 * then row 0 and column 0 halved. */
void orc_dct2(size_t M, const double *in, double *out) {
  double *C = dalloc(M * M), *tmp = dalloc(M * M);
  for (size_t k = 0; k < M; k++)
    for (size_t i = 0; i < M; i++) C[k * M + i] = cos(M_PI * (double)k * ((double)i + 0.5) / (double)M);
  for (size_t i = 0; i < M; i++)
    for (size_t l = 0; l < M; l++) {
      double acc = 0;
      for (size_t j = 0; j < M; j++) acc += in[i * M + j] * C[l * M + j];
      tmp[i * M + l] = acc;
    }
  const double sc = 4.0 / ((double)M * (double)M);
  for (size_t k = 0; k < M; k++)
    for (size_t l = 0; l < M; l++) {
      double acc = 0;
      for (size_t i = 0; i < M; i++) acc += C[k * M + i] * tmp[i * M + l];
      out[k * M + l] = acc * sc;
    }
  for (size_t i = 0; i < M; i++) {
    out[i] *= 0.5;
    out[i * M] *= 0.5;
  }
  free(C);
  free(tmp);
}

/* ---- direct plan: src/C/leg2cheb.c:758-818 -------------------------------- */
orc_direct_plan *orc_create_direct(size_t N, size_t direction, size_t precompute) {
  orc_direct_plan *p = (orc_direct_plan *)calloc(1, sizeof(orc_direct_plan));
  p->N = N;
  p->direction = direction;
  p->a = dalloc(N);
  orc_fill_a(p->a, N);
  if (direction == ORC_C2L || direction == ORC_BOTH || precompute) {
    p->an = dalloc(N);
    p->dn = dalloc((N + 1) / 2 + 1);
    orc_fill_an_dn(p->a, p->an, p->dn, N);
  }
  if (precompute == 1) {
    p->lf = dalloc((N * N + N) / 2);
    p->lb = dalloc((N * N + N) / 2);
    double *q = p->lf;
    for (size_t n = 0; n < N; n += 2) {
      const double a0 = p->a[n / 2] * M_2_PI;
      for (size_t i = 0; i < N - n; i++) *q++ = a0 * p->a[n / 2 + i];
    }
    q = p->lb;
    for (size_t n = 2; n < N; n += 2)
      for (size_t i = 0; i < N - n; i++) *q++ = p->dn[n / 2] * p->an[n / 2 + i];
  }
  return p;
}

void orc_free_direct(orc_direct_plan *p) {
  if (!p) return;
  free(p->a);
  free(p->an);
  free(p->dn);
  free(p->lf);
  free(p->lb);
  free(p);
}

/* O(N^2) sums: src/C/leg2cheb.c:152-234 (stride honoured for both
 * directions and the output overwritten). */
size_t orc_direct(const double *u, double *b, orc_direct_plan *p, size_t direction,
                  size_t st) {
  const size_t N = p->N;
  size_t flops = 0;
  if (direction > 1 || (direction == ORC_C2L && p->an == NULL)) return (size_t)-1;
  for (size_t i = 0; i < N; i++) b[i * st] = 0.0;
  flops += N;
  if (direction == ORC_L2C) {
    const double *a = p->a;
    const double *lf = p->lf;
    for (size_t n = 0; n < N; n += 2) {
      const double a0 = a[n / 2] * M_2_PI;
      for (size_t i = 0; i < N - n; i++) {
        double c = lf ? *lf++ : a0 * a[n / 2 + i];
        b[i * st] += c * u[(n + i) * st];
      }
      flops += (lf ? 2 : 3) * (N - n);
    }
    b[0] /= 2;
    flops += N;
  } else {
    const double *an = p->an, *dn = p->dn;
    const double *lb = p->lb;
    double *vn = dalloc(N);
    vn[0] = u[0];
    for (size_t i = 1; i < N; i++) vn[i] = u[i * st] * (double)i;
    for (size_t n = 0; n < N; n++) b[n * st] = SQRT_PI * vn[n] * an[n];
    for (size_t n = 2; n < N; n += 2) {
      for (size_t i = 0; i < N - n; i++) {
        if (lb)
          b[i * st] -= (*lb++) * vn[n + i];
        else
          b[i * st] -= dn[n / 2] * an[n / 2 + i] * vn[n + i];
      }
      flops += (lb ? 2 : 3) * (N - n);
    }
    for (size_t i = 0; i < N; i++) b[i * st] *= ((double)i + 0.5);
    flops += N;
    free(vn);
  }
  return flops;
}

/* ---- T: src/C/leg2cheb.c:621-648 (modal), :1009-1031 (nodal) ------------ */
static void vandermonde(double *T, size_t s, size_t M) {
  double *Tm = dalloc(2 * s * M);
  for (size_t i = 0; i < 2 * s; i++) {
    const double x = -1 + (double)i / (double)s;
    const double x2 = 2 * x;
    Tm[i * M] = 1;
    if (M > 1) Tm[i * M + 1] = x;
    for (size_t j = 2; j < M; j++) Tm[i * M + j] = Tm[i * M + j - 1] * x2 - Tm[i * M + j - 2];
  }
  for (size_t i = 0; i < s; i++)
    for (size_t j = 0; j < M; j++) {
      T[i * M + j] = Tm[2 * i * M + j];
      T[(s + i) * M + j] = Tm[(2 * i + 1) * M + j];
    }
  free(Tm);
}

static void lagrange_T(double *T, size_t s, size_t M, const double *xj, const double *wj) {
  for (size_t i = 0; i < s; i++) {
    const double xe = -1 + (double)(2 * i) / (double)s;
    const double xo = -1 + (double)(2 * i + 1) / (double)s;
    double se = 0, so = 0;
    for (size_t j = 0; j < M; j++) {
      const double e = wj[j] / (xe - xj[j]);
      const double o = wj[j] / (xo - xj[j]);
      T[i * M + j] = e;
      T[(s + i) * M + j] = o;
      se += e;
      so += o;
    }
    for (size_t j = 0; j < M; j++) {
      T[i * M + j] /= se;
      T[(s + i) * M + j] /= so;
    }
  }
}

/* ---- FMM plan: src/C/leg2cheb.c:820-1234 --------------------------------- */
orc_fmm_plan *orc_create_fmm(size_t N, size_t maxs, size_t M, size_t direction,
                             size_t lagrange, size_t precompute, size_t v) {
  (void)v;
  if (N == 0 || maxs == 0 || direction > 2) return NULL;
  orc_fmm_plan *P = (orc_fmm_plan *)calloc(1, sizeof(orc_fmm_plan));
  P->direction = direction;
  P->lagrange = lagrange;
  P->precompute = precompute;
  P->N = N;
  P->M = M;
  int L = (int)ceil(log2((double)N / (double)maxs)) - 2;
  if (L < 1) {
    P->dplan = orc_create_direct(N, direction, precompute);
    P->Nn = N;
    P->L = 0;
    return P;
  }
  if (M == 0) {
    free(P);
    return NULL;
  }
  const size_t s = (size_t)ceil((double)N / (double)((size_t)1 << (L + 2)));
  const size_t Nn = s << (L + 2);
  P->dplan = orc_create_direct(N, direction == ORC_L2C ? ORC_L2C : ORC_BOTH, 0);
  P->L = (size_t)L;
  P->s = s;
  P->Nn = Nn;
  const size_t MM = M * M;
  const size_t Nc = orc_total_sub((size_t)L);
  size_t dirs[2], nd = 0;
  if (direction == ORC_L2C || direction == ORC_BOTH) dirs[nd++] = ORC_L2C;
  if (direction == ORC_C2L || direction == ORC_BOTH) dirs[nd++] = ORC_C2L;
  for (size_t d = 0; d < nd; d++) P->A[dirs[d]] = dalloc(Nc * MM);

  double *xj = dalloc(M), *xjh = dalloc(M), *fx0 = dalloc(2 * MM), *fx1 = dalloc(MM),
         *fun = dalloc(MM);
  for (size_t i = 0; i < M; i++) xj[i] = cos(((double)i + 0.5) * M_PI / (double)M);

  size_t kk = 0;
  for (size_t level = 0; level < (size_t)L; level++) {
    const size_t h = s * orc_h(level, (size_t)L);
    for (size_t k = 0; k < M; k++) xjh[k] = xj[k] * (double)h;
    for (size_t block = 0; block < orc_nblocks(level); block++) {
      const size_t ij0 = 2 * block * h, ij1 = ij0 + 2 * h;
      for (size_t q = 0; q < 2; q++) {
        const size_t y0 = 2 * (ij1 + q * h) + h;
        for (size_t p = 0; p < q + 1; p++) {
          const size_t x0 = 2 * (ij0 + p * h) + h;
          for (size_t di = 0; di < nd; di++) {
            const size_t dir = dirs[di];
            double *ff = lagrange ? &P->A[dir][kk * MM] : fun;
            double *f00 = &fx0[q * MM];
            const double *fpq = &fx0[(q - p) * MM];
            for (size_t i = 0; i < M; i++) {
              const double x = (double)x0 + xjh[i];
              for (size_t j = 0; j < M; j++) {
                const double y = (double)y0 + xjh[j];
                const size_t ix = i * M + j, xi = j * M + i;
                /* Lambda((y-x)/2) depends on y-x only: computed for block 0
                 * of each level and reused (persymmetric), :948-955. */
                if (di == 0 && block == 0 && p == 0 && j < M - i) {
                  const double m0 = orc_lambda((y - x) / 2);
                  f00[ix] = m0;
                  f00[MM - xi - 1] = m0;
                }
                /* Lambda((x+y)/2) is symmetric in the block, :956-963. */
                if (di == 0 && j >= i) {
                  const double m1 = orc_lambda((x + y) / 2);
                  fx1[ix] = m1;
                  fx1[xi] = m1;
                }
                if (dir == ORC_L2C)
                  *ff++ = fpq[ix] * fx1[ix];
                else
                  *ff++ = 2 * (fpq[ix] / (fx1[ix] * (x + y) * (x + y + 1) * (x - y + 1)));
              }
            }
            if (!lagrange) orc_dct2(M, fun, &P->A[dir][kk * MM]);
          }
          kk++;
        }
      }
    }
  }

  P->T = dalloc(2 * s * M);
  if (!lagrange) {
    vandermonde(P->T, s, M);
    P->B = dalloc(MM);
    orc_binom_B0(M, P->B);
  } else {
    double *wj = dalloc(M);
    for (size_t j = 0; j < M; j++) wj[j] = (j % 2 == 0 ? 1.0 : -1.0) * sin(((double)j + 0.5) * M_PI / (double)M);
    lagrange_T(P->T, s, M, xj, wj);
    /* dense nodal B: src/C/leg2cheb.c:1081-1105 */
    P->B = dalloc(2 * MM);
    for (size_t i = 0; i < M; i++) {
      const double x0 = (xj[i] - 1) / 2, x1 = (xj[i] + 1) / 2;
      double s0 = 0, s1 = 0;
      for (size_t j = 0; j < M; j++) {
        const double e0 = wj[j] / (x0 - xj[j]);
        const double e1 = wj[j] / (x1 - xj[j]);
        P->B[i * M + j] = e0;
        P->B[MM + i * M + j] = e1;
        s0 += e0;
        s1 += e1;
      }
      for (size_t j = 0; j < M; j++) {
        P->B[i * M + j] /= s0;
        P->B[MM + i * M + j] /= s1;
      }
    }
    free(wj);
  }

  /* precomputed band coefficients: src/C/leg2cheb.c:1109-1192 */
  const size_t nL = N / (2 * s);
  if (precompute == 1) {
    const size_t hh = 2 * s;
    P->lfoff = (size_t *)calloc(nL + 1, sizeof(size_t));
    for (size_t b = 0; b + 1 < nL; b++) P->lfoff[b + 1] = P->lfoff[b] + 3 * s * s + s;
    const size_t i0 = (nL - 1) * hh;
    size_t last = 0;
    for (size_t n = 0; 2 * n + i0 < N; n++) last += N - (2 * n + i0);
    P->lfoff[nL] = P->lfoff[nL - 1] + last;
    const double *a = P->dplan->a;
    if (direction == ORC_L2C || direction == ORC_BOTH) {
      P->lf = dalloc(P->lfoff[nL]);
      for (size_t b = 0; b + 1 < nL; b++) {
        double *l0 = &P->lf[P->lfoff[b]];
        const size_t ib = b * hh;
        for (size_t n = 0; n < 4 * s; n += 2) {
          const size_t cnt = n < hh ? hh : 2 * hh - n;
          for (size_t i = 0; i < cnt; i++) *l0++ = a[n / 2] * a[ib + n / 2 + i];
        }
      }
      double *l0 = &P->lf[P->lfoff[nL - 1]];
      for (size_t n = 0; 2 * n + i0 < N; n++)
        for (size_t i = 0; i < N - (2 * n + i0); i++) *l0++ = a[n] * a[n + i0 + i];
    }
    if (direction == ORC_C2L || direction == ORC_BOTH) {
      const double *an = P->dplan->an, *dn = P->dplan->dn;
      P->lb = dalloc(P->lfoff[nL]);
      for (size_t b = 0; b + 1 < nL; b++) {
        double *l1 = &P->lb[P->lfoff[b]];
        const size_t ib = b * hh;
        for (size_t n = 2; n < 4 * s; n += 2) {
          const size_t cnt = n < hh ? hh : 2 * hh - n;
          for (size_t i = 0; i < cnt; i++) *l1++ = dn[n / 2] * an[ib + n / 2 + i];
        }
      }
      double *l1 = &P->lb[P->lfoff[nL - 1]];
      for (size_t n = 1; 2 * n + i0 < N; n++)
        for (size_t i = i0; i < N - 2 * n; i++) *l1++ = dn[n] * an[n + i];
    }
  }

  const size_t Nb = orc_total_blocks((size_t)L);
  P->ia = dalloc(Nn / 2);
  P->oa = dalloc(Nn / 2);
  P->wk = dalloc(Nb * 2 * M);
  P->ck = dalloc(Nb * 2 * M);
  P->vin = dalloc(N);
  free(xj);
  free(xjh);
  free(fx0);
  free(fx1);
  free(fun);
  return P;
}

void orc_free_fmm(orc_fmm_plan *P) {
  if (!P) return;
  free(P->A[0]);
  free(P->A[1]);
  free(P->T);
  free(P->B);
  free(P->lf);
  free(P->lb);
  free(P->lfoff);
  free(P->ia);
  free(P->oa);
  free(P->wk);
  free(P->ck);
  free(P->vin);
  orc_free_direct(P->dplan);
  free(P);
}

/* ---- near field, L2C: src/C/leg2cheb.c:236-376 ----------------------------
 * Row block beta (2s rows): columns [i, 2s(beta+2)) of the same parity;
 * the last block (beta = nL-1, rows to N) takes every column to N. */
static size_t band_l2c(const double *in, double *out, const orc_fmm_plan *P, size_t st) {
  const size_t s = P->s, N = P->N, hh = 2 * s, nL = N / hh;
  const double *a = P->dplan->a;
  const double *lf = P->lf;
  size_t flops = 0;
  const int pre = lf != NULL;
#pragma omp parallel for reduction(+ : flops)
  for (size_t b = 0; b < nL - 1; b++) {
    const size_t i0 = b * hh;
    const double *lp = pre ? &lf[P->lfoff[b]] : NULL;
    for (size_t n = 0; n < 4 * s; n += 2) {
      const size_t cnt = n < hh ? hh : 2 * hh - n;
      const double a0 = a[n / 2];
      const double *ap = &a[i0 + n / 2];
      for (size_t i = 0; i < cnt; i++) {
        const double c = pre ? *lp++ : a0 * ap[i];
        out[(i0 + i) * st] += c * in[(i0 + n + i) * st];
      }
      flops += cnt * (pre ? 2 : 3);
    }
  }
  const size_t i0 = (nL - 1) * hh;
  const double *lp = pre ? &lf[P->lfoff[nL - 1]] : NULL;
  for (size_t n = 0; 2 * n + i0 < N; n++) {
    const size_t cnt = N - (2 * n + i0);
    for (size_t i = 0; i < cnt; i++) {
      const double c = pre ? *lp++ : a[n] * a[n + i0 + i];
      out[(i0 + i) * st] += c * in[(i0 + 2 * n + i) * st];
    }
    flops += cnt * (pre ? 2 : 3);
  }
  for (size_t i = 0; i < N; i++) out[i * st] *= M_2_PI;
  flops += N;
  out[0] *= 0.5;
  return flops;
}

/* ---- near field, C2L: src/C/leg2cheb.c:378-535 (vin = pre-scaled input,
 * contiguous) ---------------------------------------------------------------- */
static size_t band_c2l(const double *vin, double *out, const orc_fmm_plan *P, size_t st) {
  const size_t s = P->s, N = P->N, hh = 2 * s, nL = N / hh;
  const double *an = P->dplan->an, *dn = P->dplan->dn;
  const double *lb = P->lb;
  const int pre = lb != NULL;
  size_t flops = 0;
  for (size_t i = 0; i < N; i++) out[i * st] += SQRT_PI * vin[i] * an[i];
  flops += 3 * N;
#pragma omp parallel for reduction(+ : flops)
  for (size_t b = 0; b < nL - 1; b++) {
    const size_t i0 = b * hh;
    const double *lp = pre ? &lb[P->lfoff[b]] : NULL;
    for (size_t n = 2; n < 4 * s; n += 2) {
      const size_t cnt = n < hh ? hh : 2 * hh - n;
      const double d0 = dn[n / 2];
      for (size_t i = 0; i < cnt; i++) {
        if (pre)
          out[(i0 + i) * st] -= (*lp++) * vin[i0 + n + i];
        else
          out[(i0 + i) * st] -= d0 * vin[i0 + n + i] * an[i0 + n / 2 + i];
      }
      flops += cnt * (pre ? 2 : 3);
    }
  }
  const size_t i0 = (nL - 1) * hh;
  const double *lp = pre ? &lb[P->lfoff[nL - 1]] : NULL;
  for (size_t n = 1; 2 * n + i0 < N; n++) {
    for (size_t i = i0; i < N - 2 * n; i++) {
      if (pre)
        out[i * st] -= (*lp++) * vin[i + 2 * n];
      else
        out[i * st] -= dn[n] * an[n + i] * vin[i + 2 * n];
    }
  }
  /* the reference adds (N-(2n+i0))*3 in size_t for every n < N-i0: the terms
   * cancel modulo 2^64 (src/C/leg2cheb.c:455-459) */
  for (size_t i = 0; i < N; i++) out[i * st] *= ((double)i + 0.5);
  flops += 2 * N;
  return flops;
}

size_t orc_directM(const double *in, double *out, orc_fmm_plan *P, size_t st) {
  return band_l2c(in, out, P, st);
}

/* ---- execute: src/C/leg2cheb.c:1299-1500 --------------------------------- */
size_t orc_execute(const double *in, double *out, orc_fmm_plan *P, size_t dir, size_t st) {
  if (P == NULL || dir > 1 || st == 0) return (size_t)-1;
  if (P->L == 0) return orc_direct(in, out, P->dplan, dir, st);
  if (P->A[dir] == NULL) return (size_t)-1;
  const size_t N = P->N, Nn = P->Nn, L = P->L, s = P->s, M = P->M, MM = M * M;
  const size_t Nb = orc_total_blocks(L);
  const size_t K = 2 * orc_nblocks(L - 1);
  const double *A = P->A[dir];
  const double *B = P->B;
  const int nodal = P->lagrange != 0;
  double *ia = P->ia, *oa = P->oa, *wk = P->wk, *ck = P->ck;
  size_t flops = 0;

  for (size_t i = 0; i < N; i++) out[i * st] = 0.0;
  if (dir == ORC_C2L) {
    P->vin[0] = in[0];
    if (N > 1) P->vin[1] = in[st];
    for (size_t i = 2; i < N; i++) P->vin[i] = in[i * st] * (double)i;
  }
  const size_t rest = N % 2;
  for (size_t odd = 0; odd < 2; odd++) {
    memset(oa, 0, Nn / 2 * sizeof(double));
    memset(wk, 0, Nb * 2 * M * sizeof(double));
    memset(ck, 0, Nb * 2 * M * sizeof(double));
    const size_t cnt = N / 2 + rest * (1 - odd);
    for (size_t i = 0; i < cnt; i++)
      ia[i] = dir == ORC_C2L ? P->vin[2 * i + odd] : in[(2 * i + odd) * st];
    for (size_t i = cnt; i < Nn / 2; i++) ia[i] = 0;
    const double *T = &P->T[odd * s * M];

    /* step 1: wk[L-1] = reshape(ia[2s:], K x s) . T_odd  (:1379-1383) */
    double *wl = &wk[orc_total_blocks(L - 1) * 2 * M];
#pragma omp parallel for
    for (size_t r = 0; r < K; r++) {
      const double *x = &ia[2 * s + r * s];
      double *w = &wl[r * M];
      for (size_t k = 0; k < M; k++) w[k] = 0;
      for (size_t m = 0; m < s; m++) {
        const double xm = x[m];
        const double *t = &T[m * M];
        for (size_t k = 0; k < M; k++) w[k] += xm * t[k];
      }
    }
    flops += 2 * K * s * M;

    /* step 2: upward, child pairs to parents (:1385-1418, matvectri :537-602) */
    for (size_t level = L; level-- > 1;) {
      double *w1 = &wk[orc_total_blocks(level - 1) * 2 * M];
      const double *w2 = &wk[orc_total_blocks(level) * 2 * M];
#pragma omp parallel for
      for (size_t block = 1; block < orc_nblocks(level); block++) {
        const double *x0 = &w2[block * 2 * M], *x1 = x0 + M;
        double *y = &w1[(((block - 1) / 2) * 2 + (block - 1) % 2) * M];
        if (!nodal) {
          for (size_t i = 0; i < M; i++) {
            double acc = 0;
            for (size_t j = 0; j <= i; j++) {
              const double z = ((i + j) & 1) ? x0[j] - x1[j] : x0[j] + x1[j];
              acc += B[i * M + j] * z;
            }
            y[i] = acc;
          }
        } else {
          for (size_t j = 0; j < M; j++) {
            double acc = 0;
            for (size_t i = 0; i < M; i++) acc += B[i * M + j] * x0[i];
            for (size_t i = 0; i < M; i++) acc += B[MM + i * M + j] * x1[i];
            y[j] = acc;
          }
        }
      }
      flops += (orc_nblocks(level) - 1) * MM * (nodal ? 4 : 1);
    }

    /* step 3: M2L (:1419-1435) */
    for (size_t level = 0; level < L; level++) {
      const double *ap = &A[orc_total_sub(level) * MM];
      double *cp = &ck[orc_total_blocks(level) * 2 * M];
      const double *wq = &wk[orc_total_blocks(level) * 2 * M];
#pragma omp parallel for
      for (size_t block = 0; block < orc_nblocks(level); block++) {
        const double *A0 = &ap[3 * block * MM], *A1 = A0 + MM, *A2 = A1 + MM;
        const double *w0 = &wq[block * 2 * M], *w1 = w0 + M;
        double *c0 = &cp[block * 2 * M], *c1 = c0 + M;
        for (size_t k = 0; k < M; k++) {
          double s0 = 0, s1 = 0, s2 = 0;
          for (size_t l = 0; l < M; l++) {
            s0 += A0[k * M + l] * w0[l];
            s1 += A1[k * M + l] * w1[l];
            s2 += A2[k * M + l] * w1[l];
          }
          c0[k] = s0 + s1;
          c1[k] = s2;
        }
      }
    }
    flops += orc_total_sub(L) * 2 * MM;

    /* step 4: downward (:1437-1465, matvectri upper :603-618) */
    for (size_t level = 0; level + 1 < L; level++) {
      const double *c0 = &ck[orc_total_blocks(level) * 2 * M];
      double *c1 = &ck[orc_total_blocks(level + 1) * 2 * M];
#pragma omp parallel for
      for (size_t t = 0; t < orc_nblocks(level + 1) - 1; t++) {
        const double *x = &c0[t * M];
        double *y0 = &c1[t * 2 * M], *y1 = y0 + M;
        if (!nodal) {
          for (size_t i = 0; i < M; i++) {
            double se = 0, so = 0;
            for (size_t j = i; j < M; j += 2) se += B[j * M + i] * x[j];
            for (size_t j = i + 1; j < M; j += 2) so += B[j * M + i] * x[j];
            y0[i] += se + so;
            y1[i] += se - so;
          }
        } else {
          for (size_t i = 0; i < M; i++) {
            double a0 = 0, a1 = 0;
            for (size_t j = 0; j < M; j++) {
              a0 += B[i * M + j] * x[j];
              a1 += B[MM + i * M + j] * x[j];
            }
            y0[i] += a0;
            y1[i] += a1;
          }
        }
      }
      flops += (orc_nblocks(level + 1) - 1) * MM * (nodal ? 4 : 1);
    }

    /* step 5: oa = ck[L-1] . T_odd^T, scatter (:1467-1483) */
    const double *cl = &ck[orc_total_blocks(L - 1) * 2 * M];
#pragma omp parallel for
    for (size_t r = 0; r < K; r++) {
      const double *c = &cl[r * M];
      for (size_t m = 0; m < s; m++) {
        const double *t = &T[m * M];
        double acc = 0;
        for (size_t k = 0; k < M; k++) acc += c[k] * t[k];
        oa[r * s + m] = acc;
      }
    }
    flops += 2 * K * s * M;
    for (size_t i = 0; i < cnt; i++) out[(2 * i + odd) * st] += oa[i];
  }
  flops += dir == ORC_L2C ? band_l2c(in, out, P, st) : band_c2l(P->vin, out, P, st);
  return flops;
}

/* ---- 2-D driver: src/C/leg2cheb.c:1236-1297 ------------------------------- */
orc_fmm_plan_2d *orc_create_fmm_2d(size_t N0, size_t N1, int axis, size_t maxs, size_t M,
                                   size_t direction, size_t lagrange, size_t precompute,
                                   size_t v) {
  orc_fmm_plan_2d *p = (orc_fmm_plan_2d *)calloc(1, sizeof(orc_fmm_plan_2d));
  p->N0 = N0;
  p->N1 = N1;
  p->axis = axis;
  if (axis == 0 || axis == -1) p->p0 = orc_create_fmm(N0, maxs, M, direction, lagrange, precompute, v);
  if (axis == 1) p->p1 = orc_create_fmm(N1, maxs, M, direction, lagrange, precompute, v);
  if (axis == -1)
    p->p1 = N0 == N1 ? p->p0 : orc_create_fmm(N1, maxs, M, direction, lagrange, precompute, v);
  return p;
}

void orc_free_fmm_2d(orc_fmm_plan_2d *p) {
  if (!p) return;
  if (p->p1 != p->p0) orc_free_fmm(p->p1);
  orc_free_fmm(p->p0);
  free(p);
}

size_t orc_execute2D(const double *in, double *out, orc_fmm_plan_2d *p, size_t dir) {
  size_t flops = 0;
  const size_t N0 = p->N0, N1 = p->N1;
  if (p->axis == 0) {
    for (size_t i = 0; i < N1; i++) flops += orc_execute(&in[i], &out[i], p->p0, dir, N1);
  } else if (p->axis == 1) {
    for (size_t i = 0; i < N0; i++) flops += orc_execute(&in[i * N1], &out[i * N1], p->p1, dir, 1);
  } else {
    double *tmp = dzalloc(N0 * N1);
    for (size_t i = 0; i < N1; i++) flops += orc_execute(&in[i], &tmp[i], p->p0, dir, N1);
    for (size_t i = 0; i < N0; i++) flops += orc_execute(&tmp[i * N1], &out[i * N1], p->p1, dir, 1);
    free(tmp);
  }
  return flops;
}
