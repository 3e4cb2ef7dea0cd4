/*
 * orc_l2c.h -- the oracle: a CPU restatement of the reference's fast multipole
 * Legendre<->Chebyshev transform (spectralDNS/Legendre-to-Chebyshev), plus
 * extended-precision direct sums.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline -- never as the product path.
 *
 * Every function names the reference lines it restates (paths relative to the
 * reference repository root).  Deliberate differences from the reference,
 * all of them bug fixes that do not change results on valid inputs:
 *   - outputs are overwritten, not accumulated (src/C/leg2cheb.c:1475,1480
 *     accumulate while direct() at :157 overwrites);
 *   - the odd-N parity-1 scatter writes N/2 elements, never out[N]
 *     (src/C/leg2cheb.c:1473-1476 writes one past the end);
 *   - direct() honours the stride for both directions (src/C/leg2cheb.c:157,
 *     :165-168 read and zero contiguously);
 *   - an unplanned direction returns (size_t)-1 instead of dereferencing NULL
 *     (src/C/leg2cheb.c:1309-1312);
 *   - the precomputed last-block band table is sized for the last block
 *     (src/C/leg2cheb.c:1118-1119 allocates 3s^2+s, which the last block
 *     can exceed);
 *   - B(0) is the exact dyadic-rational matrix for every M (the reference
 *     uses exact tables at M=18, src/C/leg2cheb.c:6-123, and an FFTW DCT
 *     otherwise, :1042-1080).
 */
#ifndef ORC_L2C_H
#define ORC_L2C_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_L2C = 0, ORC_C2L = 1, ORC_BOTH = 2 };

typedef struct {
  size_t direction;
  size_t N;
  double *a;  /* Lambda(k), k < N */
  double *an; /* C2L diagonal scale */
  double *dn; /* C2L band factor */
  double *lf; /* precomputed L2C coefficients (precompute=1), (N^2+N)/2 */
  double *lb; /* precomputed C2L coefficients (precompute=1) */
} orc_direct_plan;

typedef struct {
  size_t direction, M, N, Nn, L, s, lagrange, precompute;
  double *A[2];   /* Nc x M x M per direction, level -> block -> r */
  double *T;      /* 2 x s x M: even rows, then odd rows */
  double *B;      /* modal: dense M x M B(0); nodal: 2 x M x M */
  double *lf;     /* precomputed band (L2C), nL blocks */
  double *lb;     /* precomputed band (C2L), nL blocks */
  size_t *lfoff;  /* block offsets into lf/lb (nL+1 entries) */
  orc_direct_plan *dplan;
  /* workspaces (execute is not re-entrant per plan, like the reference) */
  double *ia, *oa, *wk, *ck, *vin;
} orc_fmm_plan;

typedef struct {
  orc_fmm_plan *p0, *p1;
  size_t N0, N1;
  int axis;
} orc_fmm_plan_2d;

/* Lambda(z) = Gamma(z+1/2)/Gamma(z+1).  src/C/lambda.c:65-98 */
double orc_lambda(double z);
/* Lambda at integer k; exact table for k < 256.  src/C/lambda.c:101-192 */
double orc_lambda_int(size_t k);
void orc_lambdavec(const double *z, double *w, size_t n);
/* k-th asymptotic-series coefficient used by orc_lambda (k = 0..6). */
double orc_lambda_series_coef(size_t k);

/* Plan geometry.  src/C/leg2cheb.c:125-150 */
size_t orc_nblocks(size_t level);
size_t orc_h(size_t level, size_t L);
size_t orc_total_blocks(size_t L);
size_t orc_total_sub(size_t L);

/* Exact B(0) (dense M x M, lower triangular): Chebyshev coefficients of
 * T_k((x-1)/2).  Matches BMe at M=18 (src/C/leg2cheb.c:6-63). */
void orc_binom_B0(size_t M, double *B0);
/* 2-D DCT-II scaled by 4/M^2, row 0 and column 0 halved.  src/C/dct.c:582-598 */
void orc_dct2(size_t M, const double *in, double *out);

orc_direct_plan *orc_create_direct(size_t N, size_t direction, size_t precompute);
void orc_free_direct(orc_direct_plan *p);
size_t orc_direct(const double *u, double *b, orc_direct_plan *p,
                  size_t direction, size_t stride);

orc_fmm_plan *orc_create_fmm(size_t N, size_t maxs, size_t M, size_t direction,
                             size_t lagrange, size_t precompute, size_t v);
void orc_free_fmm(orc_fmm_plan *p);
size_t orc_execute(const double *in, double *out, orc_fmm_plan *p,
                   size_t direction, size_t stride);
/* Band (near field) alone, with the final scaling.  src/C/leg2cheb.c:236-535 */
size_t orc_directM(const double *in, double *out, orc_fmm_plan *p, size_t stride);

orc_fmm_plan_2d *orc_create_fmm_2d(size_t N0, size_t N1, int axis, size_t maxs,
                                   size_t M, size_t direction, size_t lagrange,
                                   size_t precompute, size_t v);
void orc_free_fmm_2d(orc_fmm_plan_2d *p);
size_t orc_execute2D(const double *in, double *out, orc_fmm_plan_2d *p,
                     size_t direction);

/* Extended precision direct sums (src/multiprec/C++/L2Cprec.cpp:19-69,
 * src/multiprec/python/L2Cprecision.py:10-31), rows [r0, r1) of the output.
 * _q: __float128 throughout; _dd: double-double accumulation (Lambda tables
 * built in __float128).  Inputs are doubles, outputs rounded to double. */
void orc_leg2cheb_q(const double *u, double *b, size_t N, size_t r0, size_t r1);
void orc_cheb2leg_q(const double *u, double *b, size_t N, size_t r0, size_t r1);
void orc_leg2cheb_dd(const double *u, double *b, size_t N, size_t r0, size_t r1);
void orc_cheb2leg_dd(const double *u, double *b, size_t N, size_t r0, size_t r1);
void orc_leg2cheb_dd_rows(const double *u, double *b, size_t N, const size_t *rows, size_t nr);
void orc_cheb2leg_dd_rows(const double *u, double *b, size_t N, const size_t *rows, size_t nr);
/* Lambda(x) in __float128 via lgammaq, rounded to double (pins orc_lambda). */
double orc_lambda_q(double x);

int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
