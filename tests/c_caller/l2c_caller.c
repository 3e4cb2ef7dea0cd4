/* A plain C99 caller of the drop-in library, written the way the reference's
 * own driver calls its library (src/bin/l2c.c:181-212 plan / execute with
 * host pointers on a zeroed output, :297-351 create_fmm_2d / execute2D with
 * input 1 + index): L2C then C2L, the O(N^2) direct sum, the 2-D driver along
 * both axes, an error path, and the DCT-II extension.
 *
 *   l2c_caller N seed out.bin
 *
 * N >= 8.  Writes raw doubles: u (N, SplitMix64 U(-1,1)), L2C(u), C2L(L2C(u)),
 * direct L2C(u), the 2-D L2C of an 8 x N array along axis 1 and of an
 * N x 8 array along axis 0 (each 8 N values), and, when N is a power of two,
 * REDFT10(u) (N values; else nothing).  Prints "ok" on success. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "leg2cheb_amd.h"

static double uniform(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return -1.0 + 2.0 * (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static int fail(const char *what) {
  fprintf(stderr, "%s: %s\n", what, l2c_last_error());
  return 1;
}

int main(int argc, char **argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s N seed out.bin\n", argv[0]);
    return 2;
  }
  const size_t N = (size_t)strtoull(argv[1], NULL, 10);
  uint64_t state = strtoull(argv[2], NULL, 10);
  if (N < 8) return 2;
  double *u = malloc(N * sizeof(double)), *b = calloc(N, sizeof(double));
  double *c = calloc(N, sizeof(double)), *d = calloc(N, sizeof(double));
  double *a2 = malloc(8 * N * sizeof(double)), *o2 = calloc(8 * N, sizeof(double));
  double *o3 = calloc(8 * N, sizeof(double)), *y = calloc(N, sizeof(double));
  for (size_t i = 0; i < N; i++) u[i] = uniform(&state);

  fmm_plan *plan = create_fmm(N, 64, 18, BOTH, 0, 0, 0);
  if (!plan) return fail("create_fmm");
  if (execute(u, b, plan, L2C, 1) == (size_t)-1) return fail("execute L2C");
  if (execute(b, c, plan, C2L, 1) == (size_t)-1) return fail("execute C2L");
  /* an unplanned direction fails cleanly instead of asserting */
  if (execute(u, d, plan, 7, 1) != (size_t)-1 || !l2c_last_error()[0]) {
    fprintf(stderr, "error path not taken\n");
    return 1;
  }
  free_fmm(plan);

  direct_plan *dp = create_direct(N, L2C, 0);
  if (!dp) return fail("create_direct");
  if (direct(u, d, dp, L2C, 1) == (size_t)-1) return fail("direct");
  free_direct(dp);

  /* 8 x N, transformed along axis 1 (rows); N x 8 along axis 0 (columns) */
  for (size_t i = 0; i < 8; i++)
    for (size_t j = 0; j < N; j++) a2[i * N + j] = 1.0 + j;
  fmm_plan_2d *p2 = create_fmm_2d(8, N, 1, 64, 18, BOTH, 0, 0, 0);
  if (!p2) return fail("create_fmm_2d axis 1");
  if (execute2D(a2, o2, p2, L2C) == (size_t)-1) return fail("execute2D axis 1");
  free_fmm_2d(p2);
  for (size_t i = 0; i < N; i++)
    for (size_t j = 0; j < 8; j++) a2[i * 8 + j] = 1.0 + i;
  p2 = create_fmm_2d(N, 8, 0, 64, 18, BOTH, 0, 0, 0);
  if (!p2) return fail("create_fmm_2d axis 0");
  if (execute2D(a2, o3, p2, L2C) == (size_t)-1) return fail("execute2D axis 0");
  free_fmm_2d(p2);

  const int pow2 = (N & (N - 1)) == 0;
  if (pow2) {
    l2c_dct_plan *dct = l2c_create_dct(N, 1);
    if (!dct) return fail("l2c_create_dct");
    if (l2c_execute_dct(dct, u, y, 2)) return fail("l2c_execute_dct");
    l2c_free_dct(dct);
  }

  FILE *f = fopen(argv[3], "wb");
  if (!f) return 1;
  fwrite(u, sizeof(double), N, f);
  fwrite(b, sizeof(double), N, f);
  fwrite(c, sizeof(double), N, f);
  fwrite(d, sizeof(double), N, f);
  fwrite(o2, sizeof(double), 8 * N, f);
  fwrite(o3, sizeof(double), 8 * N, f);
  if (pow2) fwrite(y, sizeof(double), N, f);
  fclose(f);
  free(u); free(b); free(c); free(d); free(a2); free(o2); free(o3); free(y);
  printf("ok\n");
  return 0;
}
