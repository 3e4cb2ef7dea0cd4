// l2c_exec.hip -- execute path of the gfx950 FMM Legendre<->Chebyshev
// transform (src/C/leg2cheb.c:1299-1500 and the near field :236-535).
//
// One transform (or one line of a batch) runs as three launches:
//   k_up_subtree   leaf moments (step 1) and the upward pass (step 2) inside
//                  a subtree of 2^gu leaf chunks, both parities; writes the
//                  multipole store w for every level >= lu.
//   k_up_top       multipoles of the levels above lu, each from its level-lu
//                  descendants through composite operators P (no level loop).
//   k_down_band    per subtree of 2^gd leaf chunks: M2L (step 3) and L2L
//                  (step 4) along its ancestor chain and inside the subtree,
//                  leaf evaluation (step 5), the near-field band, the final
//                  scaling, one coalesced store of the output.
// Both parities share each M2L matrix read.  The output is overwritten.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "l2c_internal.h"

namespace l2c {

struct LineMap {
  size_t n0, ls0, ls1, es;  // line l starts at (l/n0)*ls1 + (l%n0)*ls0; element i at +i*es
};
__device__ inline size_t line_off(const LineMap &m, size_t l) {
  return (l / m.n0) * m.ls1 + (l % m.n0) * m.ls0;
}

constexpr double kSqrtPi = 1.77245385090551602729816748334;
constexpr double k2Pi = 0.63661977236758134307553505349;  // 2/pi

// ---------------------------------------------------------------------------
// k_up_subtree: grid (2^(lu+2), lines), 256 threads.
__global__ void __launch_bounds__(256) k_up_subtree(const double *__restrict__ in, LineMap lm,
                                                    Geo g, int dir,
                                                    const double *__restrict__ T,
                                                    const double *__restrict__ U,
                                                    double *__restrict__ wall, size_t wstride) {
  extern __shared__ double sm[];
  const int s = g.s, M = g.M, L = g.L;
  const int G = 1 << g.gu;
  const int j = blockIdx.x;
  const int tid = threadIdx.x, bd = blockDim.x;
  const double *src = in + line_off(lm, blockIdx.y);
  double *w = wall + blockIdx.y * wstride;
  double *xs = sm;               // [2][G*s]
  double *w0 = xs + 2 * G * s;   // [G][2][M]
  double *w1 = w0 + G * 2 * M;   // [G/2][2][M]
  const double *Ts = T, *Us = U;  // small, cache resident
  const size_t base = (size_t)j * G * 2 * s;
  for (int e = tid; e < G * 2 * s; e += bd) {
    const size_t i = base + e;
    double v = 0;
    if (i < (size_t)g.N) {
      v = src[i * lm.es];
      if (dir == C2L && i > 0) v *= (double)i;  // v_i = i u_i (src/C/leg2cheb.c:1325-1335)
    }
    xs[(e & 1) * G * s + (e >> 1)] = v;
  }
  __syncthreads();
  // step 1: moments of each leaf chunk, w[k] = sum_m T_par[m][k] x_par[m]
  for (int o = tid; o < G * 2 * M; o += bd) {
    const int c = o / (2 * M), sg = (o / M) & 1, k = o % M;
    const double *x = xs + sg * G * s + c * s;
    const double *t = Ts + sg * s * M + k;
    double acc = 0;
    for (int m = 0; m < s; m++) acc += t[m * M] * x[m];
    w0[o] = acc;
    w[(slot_base(L - 1) + (size_t)j * G + c) * 2 * M + sg * M + k] = acc;
  }
  __syncthreads();
  // step 2: parent = U0 child0 + U1 child1, level by level inside the subtree
  int n = G;
  for (int lev = L - 2; lev >= g.lu; lev--) {
    n >>= 1;
    for (int o = tid; o < n * 2 * M; o += bd) {
      const int c = o / (2 * M), sg = (o / M) & 1, k = o % M;
      const double *x0 = w0 + (2 * c * 2 + sg) * M;
      const double *x1 = w0 + ((2 * c + 1) * 2 + sg) * M;
      const double *u0 = Us + k * M, *u1 = Us + M * M + k * M;
      double acc = 0;
      for (int l = 0; l < M; l++) acc += u0[l] * x0[l] + u1[l] * x1[l];
      w1[o] = acc;
      w[(slot_base(lev) + (size_t)j * n + c) * 2 * M + sg * M + k] = acc;
    }
    __syncthreads();
    double *t = w0;
    w0 = w1;
    w1 = t;
  }
}

// ---------------------------------------------------------------------------
// k_up_top: grid (slot_base(lu), lines), 256 threads.  Multipole of chunk c
// at level l < lu = sum_d P(lu-l, d) w[lu][c 2^(lu-l) + d].
__global__ void __launch_bounds__(256) k_up_top(Geo g, const double *__restrict__ P,
                                                double *__restrict__ wall, size_t wstride) {
  __shared__ double red[256];
  const int M = g.M, MM = M * M, M2 = 2 * M;
  double *w = wall + blockIdx.y * wstride;
  const int q = blockIdx.x;
  int lev = 0;
  while ((int)slot_base(lev + 1) <= q) lev++;
  const int c = q - (int)slot_base(lev);
  const int k = g.lu - lev, nd = 1 << k;
  const double *Pk = P + (size_t)((1 << k) - 2) * MM;
  const double *child = w + (slot_base(g.lu) + (size_t)c * nd) * M2;
  const int tid = threadIdx.x, ngrp = blockDim.x / M2;
  const int o = tid % M2, grp = tid / M2;
  double acc = 0;
  if (grp < ngrp) {
    const int sg = o / M, i = o % M;
    for (int d = grp; d < nd; d += ngrp) {
      const double *pr = Pk + (size_t)d * MM + i * M;
      const double *x = child + (size_t)d * M2 + sg * M;
      for (int l = 0; l < M; l++) acc += pr[l] * x[l];
    }
  }
  red[tid] = acc;
  __syncthreads();
  if (tid < M2) {
    double sum = 0;
    for (int gi = 0; gi < ngrp; gi++) sum += red[gi * M2 + tid];
    w[(slot_base(lev) + c) * M2 + tid] = sum;
  }
}

// M2L for row chunk r at `level` (src/C/leg2cheb.c:1419-1435): block b = r/2;
// r even: A(b,0) w[r+2] + A(b,1) w[r+3]; r odd: A(b,2) w[r+2].
__device__ inline double m2l_row(const double *__restrict__ A, const double *__restrict__ w,
                                 int level, size_t r, int sg, int k, int M) {
  const size_t MM = (size_t)M * M;
  const double *Ab = A + (3 * total_blocks(level) + 3 * (r >> 1)) * MM;
  const double *wb = w + (slot_base(level) + r + 2) * 2 * M + sg * M;
  double acc = 0;
  if ((r & 1) == 0) {
    const double *a0 = Ab + k * M, *a1 = Ab + MM + k * M;
    double acc1 = 0;
    for (int l = 0; l < M; l++) acc += a0[l] * wb[l];
    for (int l = 0; l < M; l++) acc1 += a1[l] * wb[2 * M + l];
    acc += acc1;
  } else {
    const double *a2 = Ab + 2 * MM + k * M;
    for (int l = 0; l < M; l++) acc += a2[l] * wb[l];
  }
  return acc;
}

// ---------------------------------------------------------------------------
// k_down_band: grid (2^(ld+2), lines), 256 threads.
__global__ void __launch_bounds__(256) k_down_band(
    const double *__restrict__ in, double *__restrict__ out, LineMap lm, Geo g, int dir,
    int band_only, int precompute, const double *__restrict__ A,
    const double *__restrict__ wall, size_t wstride, const double *__restrict__ D,
    const double *__restrict__ T, const double *__restrict__ a, const double *__restrict__ an,
    const double *__restrict__ dn) {
  extern __shared__ double sm[];
  const int s = g.s, M = g.M, L = g.L, N = g.N, M2 = 2 * M;
  const int G = 1 << g.gd;
  const int R = blockIdx.x;
  const int tid = threadIdx.x, bd = blockDim.x;
  const size_t row0 = (size_t)R * G * 2 * s;
  if (row0 >= (size_t)N) return;
  const size_t row1 = std::min((size_t)N, row0 + (size_t)G * 2 * s);
  const size_t off = line_off(lm, blockIdx.y);
  const double *src = in + off;
  double *dst = out + off;
  const double *w = wall + blockIdx.y * wstride;

  // LDS carve-up
  double *cA = sm;                     // [G][2][M]
  double *cB = cA + G * M2;            // [G][2][M]
  double *chain = cB + G * M2;         // [ld+1][2][M]
  double *xw = chain + (g.ld + 1) * M2;  // input window, G*2s + 2s
  const double *Ds = D, *Tl = T;       // small, cache resident
  double *aw = xw + G * 2 * s + 2 * s; // Lambda window (a or an), G*2s + 2s
  double *hd = aw + G * 2 * s + 2 * s; // a[0..2s) or dn[0..2s)
  double *cur = cA;

  // input window [row0, min(N, row1 + 2s)) with the C2L pre-scale
  const size_t wend = std::min((size_t)N, row1 + 2 * (size_t)s);
  for (size_t e = row0 + tid; e < wend; e += bd) {
    double v = src[e * lm.es];
    if (dir == C2L && e > 0) v *= (double)e;
    xw[e - row0] = v;
  }
  const double *tab = dir == L2C ? a : an;
  for (size_t e = row0 + tid; e < wend; e += bd) aw[e - row0] = tab[e];
  const double *htab = dir == L2C ? a : dn;
  const int nh = std::min(2 * s, dir == L2C ? N : (N + 1) / 2);
  for (int e = tid; e < 2 * s; e += bd) hd[e] = e < nh ? htab[e] : 0.0;

  if (!band_only) {
    // M2L along the ancestor chain, levels 0..ld (level ld is chunk R itself)
    for (int o = tid; o < (g.ld + 1) * M2; o += bd) {
      const int lev = o / M2, sg = (o % M2) / M, k = o % M;
      const size_t r = (size_t)R >> (g.ld - lev);
      chain[o] = r < ((size_t)4 << lev) - 2 ? m2l_row(A, w, lev, r, sg, k, M) : 0.0;
    }
    __syncthreads();
    // L2L down the chain: c_l = D(p) c_{l-1} + m2l_l (src/C/leg2cheb.c:1437-1449)
    for (int lev = 1; lev <= g.ld; lev++) {
      const size_t r = (size_t)R >> (g.ld - lev);
      const bool valid = r < ((size_t)4 << lev) - 2;
      const int p = (int)(r & 1);
      double v = 0;
      if (tid < M2) {
        const int sg = tid / M, i = tid % M;
        const double *par = chain + (lev - 1) * M2 + sg * M;
        const double *dr = Ds + p * M * M + i * M;
        double acc = 0;
        for (int jj = 0; jj < M; jj++) acc += dr[jj] * par[jj];
        v = valid ? acc + chain[lev * M2 + tid] : 0.0;
      }
      __syncthreads();
      if (tid < M2) chain[lev * M2 + tid] = v;
      __syncthreads();
    }
    for (int e = tid; e < M2; e += bd) cA[e] = chain[g.ld * M2 + e];
    __syncthreads();
    // inside the subtree: levels ld+1 .. L-1
    double *nxt = cB;
    for (int lev = g.ld + 1; lev < L; lev++) {
      const int n = 1 << (lev - g.ld);
      for (int o = tid; o < n * M2; o += bd) {
        const int gi = o / M2, sg = (o % M2) / M, k = o % M;
        const size_t r = (size_t)R * n + gi;
        double v = 0;
        if (r < ((size_t)4 << lev) - 2) {
          const double *par = cur + (gi >> 1) * M2 + sg * M;
          const double *dr = Ds + (int)(r & 1) * M * M + k * M;
          double acc = 0;
          for (int jj = 0; jj < M; jj++) acc += dr[jj] * par[jj];
          v = acc + m2l_row(A, w, lev, r, sg, k, M);
        }
        nxt[o] = v;
      }
      __syncthreads();
      double *t = cur;
      cur = nxt;
      nxt = t;
    }
  }
  __syncthreads();

  // near field + leaf evaluation + final scaling, one row per thread
  const size_t nLm1 = (size_t)g.nL - 1;
  const size_t nvalid = ((size_t)2 << L) - 2;  // leaf chunks with a local expansion
  for (size_t i = row0 + tid; i < row1; i += bd) {
    const size_t beta = i / (2 * s);
    const size_t jend = beta < nLm1 ? 2 * (size_t)s * (beta + 2) : (size_t)N;
    double acc = 0;
    if (dir == L2C) {
      for (size_t jx = i; jx < jend; jx += 2) {
        const size_t h = (jx - i) >> 1, pp = (jx + i) >> 1;
        acc += hd[h] * aw[pp - row0] * xw[jx - row0];
      }
    } else {
      acc = kSqrtPi * xw[i - row0] * aw[i - row0];
      for (size_t jx = i + 2; jx < jend; jx += 2) {
        const size_t h = (jx - i) >> 1, pp = (jx + i) >> 1;
        if (precompute)
          acc -= hd[h] * aw[pp - row0] * xw[jx - row0];
        else
          acc -= hd[h] * xw[jx - row0] * aw[pp - row0];
      }
    }
    double far = 0;
    if (!band_only) {
      const size_t r = i >> 1;
      const int sg = (int)(i & 1);
      const size_t t = r / s;
      if (t < nvalid) {
        const int m = (int)(r - t * s);
        const int gi = (int)(t - (size_t)R * G);
        const double *c = cur + gi * M2 + sg * M;
        const double *tr = Tl + (sg * s + m) * M;
        for (int k = 0; k < M; k++) far += tr[k] * c[k];
      }
    }
    double v;
    if (dir == L2C) {
      v = (far + acc) * k2Pi;
      if (i == 0) v *= 0.5;
    } else {
      v = (far + acc) * ((double)i + 0.5);
    }
    dst[i * lm.es] = v;
  }
}

size_t down_shared_bytes(const Geo &g) {
  const int G = 1 << g.gd;
  return sizeof(double) * ((size_t)2 * G * 2 * g.M + (g.ld + 1) * 2 * g.M +
                           2 * ((size_t)G * 2 * g.s + 2 * g.s) + 2 * g.s);
}
size_t up_shared_bytes(const Geo &g) {
  const int G = 1 << g.gu;
  return sizeof(double) * ((size_t)2 * G * g.s + G * 2 * g.M + (G / 2 + 1) * 2 * g.M);
}

// ---------------------------------------------------------------------------
// direct O(N^2) sums, one row per thread: src/C/leg2cheb.c:152-234
__global__ void __launch_bounds__(256) k_direct(const double *__restrict__ in,
                                                double *__restrict__ out, LineMap lm, int N,
                                                int dir, const double *__restrict__ a,
                                                const double *__restrict__ an,
                                                const double *__restrict__ dn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const size_t off = line_off(lm, blockIdx.y);
  const double *u = in + off;
  const size_t es = lm.es;
  double acc = 0;
  if (dir == L2C) {
    for (int n = 0; n + i < N; n += 2) acc += a[n / 2] * k2Pi * a[n / 2 + i] * u[(size_t)(n + i) * es];
    if (i == 0) acc /= 2;
  } else {
    const double vi = i == 0 ? u[0] : u[(size_t)i * es] * (double)i;
    acc = kSqrtPi * vi * an[i];
    for (int n = 2; n + i < N; n += 2)
      acc -= dn[n / 2] * an[n / 2 + i] * (u[(size_t)(n + i) * es] * (double)(n + i));
    acc *= ((double)i + 0.5);
  }
  out[off + (size_t)i * es] = acc;
}

// ---------------------------------------------------------------------------
// host orchestration

// Multipole store for a batch of lines: grown on demand, owned by the plan.
static int ensure_lines_ws(fmm_plan_s *p, size_t nlines, double **w) {
  if (nlines <= 1) {
    *w = p->d_w;
    return 0;
  }
  const size_t need = nlines * slot_base(p->geo.L) * 2 * p->M;
  if (p->wl_cap < need) {
    L2C_HIP(hipStreamSynchronize(p->stream), -1);
    if (p->d_wl) (void)hipFree(p->d_wl);
    p->d_wl = nullptr;
    p->wl_cap = 0;
    L2C_HIP(hipMalloc(&p->d_wl, need * sizeof(double)), -1);
    p->wl_cap = need;
  }
  *w = p->d_wl;
  return 0;
}

static int launch_fmm(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir,
                      const LineMap &lm, size_t nlines, hipStream_t st, int band_only,
                      double *w) {
  const Geo &g = p->geo;
  const size_t wstride = slot_base(g.L) * 2 * g.M;
  if (!band_only) {
    dim3 gu((unsigned)((size_t)4 << g.lu), (unsigned)nlines);
    k_up_subtree<<<gu, 256, up_shared_bytes(g), st>>>(d_in, lm, g, (int)dir, p->t.T, p->t.U, w,
                                                     wstride);
    L2C_HIP(hipGetLastError(), -1);
    if (g.lu > 0) {
      dim3 gt((unsigned)slot_base(g.lu), (unsigned)nlines);
      k_up_top<<<gt, 256, 0, st>>>(g, p->t.P, w, wstride);
      L2C_HIP(hipGetLastError(), -1);
    }
  }
  dim3 gd((unsigned)((size_t)4 << g.ld), (unsigned)nlines);
  k_down_band<<<gd, 256, down_shared_bytes(g), st>>>(
      d_in, d_out, lm, g, (int)dir, band_only, (int)p->precompute, p->t.A[dir], w, wstride,
      p->t.D, p->t.T, p->t.a, p->t.an, p->t.dn);
  L2C_HIP(hipGetLastError(), -1);
  return 0;
}

int run_fmm(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir, size_t stride,
            hipStream_t st, int band_only) {
  LineMap lm{1, 0, 0, stride};
  return launch_fmm(p, d_in, d_out, dir, lm, 1, st, band_only, p->d_w);
}

int run_direct(const direct_plan_s *dp, const double *d_in, double *d_out, size_t dir,
               size_t stride, hipStream_t st, size_t precompute) {
  (void)precompute;
  LineMap lm{1, 0, 0, stride};
  const int N = (int)dp->N;
  k_direct<<<dim3((N + 255) / 256, 1), 256, 0, st>>>(d_in, d_out, lm, N, (int)dir, dp->d_a,
                                                     dp->d_an, dp->d_dn);
  L2C_HIP(hipGetLastError(), -1);
  return 0;
}

int run_lines(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir, size_t nlines,
              size_t n0, size_t ls0, size_t ls1, size_t es, hipStream_t st) {
  LineMap lm{n0 ? n0 : 1, ls0, ls1, es};
  if (p->L == 0) {
    const direct_plan_s *dp = p->dplan;
    const int N = (int)dp->N;
    for (size_t l0 = 0; l0 < nlines; l0 += 65535) {
      const size_t nl = std::min<size_t>(65535, nlines - l0);
      // line offsets are computed from blockIdx.y: shift the map by l0 lines
      LineMap lml = lm;
      const size_t base = (l0 / lm.n0) * lm.ls1 + (l0 % lm.n0) * lm.ls0;
      if (l0 % lm.n0 != 0) {
        set_error("run_lines: batches above 65535 lines need n0 to divide 65535-line chunks");
        return -1;
      }
      k_direct<<<dim3((N + 255) / 256, (unsigned)nl), 256, 0, st>>>(
          d_in + base, d_out + base, lml, N, (int)dir, dp->d_a, dp->d_an, dp->d_dn);
      L2C_HIP(hipGetLastError(), -1);
    }
    return 0;
  }
  if (nlines > 65535) {
    set_error("run_lines: at most 65535 lines per call for FMM plans");
    return -1;
  }
  double *w = nullptr;
  if (ensure_lines_ws(p, nlines, &w)) return -1;
  return launch_fmm(p, d_in, d_out, dir, lm, nlines, st, 0, w);
}

// ---------------------------------------------------------------------------
// The reference's flop count (the return value of execute), same formulas:
// src/C/leg2cheb.c:1383, 1398, 1415, 1435, 1447, 1463, 1469 and the band
// counters of directM/directL (:236-535) and direct (:152-234).
size_t flops_direct(size_t N, size_t dir, size_t pre) {
  size_t f = N;
  const size_t per = pre ? 2 : 3;
  for (size_t n = dir == L2C ? 0 : 2; n < N; n += 2) f += per * (N - n);
  f += N;
  return f;
}

size_t flops_fmm(const fmm_plan_s *p, size_t dir) {
  if (p->L == 0) return flops_direct(p->N, dir, p->precompute);
  const size_t L = p->L, s = p->s, M = p->M, MM = M * M, N = p->N;
  const size_t K = 2 * nblocks(L - 1);
  const size_t mul = p->lagrange ? 4 : 1;
  size_t f = 0;
  for (int odd = 0; odd < 2; odd++) {
    f += 2 * K * s * M;
    for (size_t level = L; level-- > 1;) f += (nblocks(level) - 1) * MM * mul;
    f += total_sub(L) * 2 * MM;
    for (size_t level = 0; level + 1 < L; level++) f += (nblocks(level + 1) - 1) * MM * mul;
    f += 2 * K * s * M;
  }
  const size_t h = 2 * s, nL = N / h;
  const size_t per = p->precompute ? 2 : 3;
  if (dir == L2C) {
    for (size_t b = 0; b + 1 < nL; b++)
      for (size_t n = 0; n < 4 * s; n += 2) f += per * (n < h ? h : 2 * h - n);
    const size_t i0 = (nL - 1) * h;
    for (size_t n = 0; 2 * n + i0 < N; n++) f += per * (N - (2 * n + i0));
    f += N;
  } else {
    f += 3 * N;
    for (size_t b = 0; b + 1 < nL; b++)
      for (size_t n = 2; n < 4 * s; n += 2) f += per * (n < h ? h : 2 * h - n);
    // last block: the reference's size_t terms sum to zero modulo 2^64
    f += 2 * N;
  }
  return f;
}

// ---------------------------------------------------------------------------
// pointer classification

static bool is_device_ptr(const void *ptr) {
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, ptr);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
#if HIP_VERSION_MAJOR >= 6
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
#else
  return attr.memoryType == hipMemoryTypeDevice;
#endif
}

}  // namespace l2c

using namespace l2c;

extern "C" {

size_t execute(const double *input_array, double *output_array, fmm_plan *p, size_t direction,
               const size_t stride) {
  clear_error();
  if (!p || !input_array || !output_array) {
    set_error("execute: NULL argument");
    return (size_t)-1;
  }
  if (direction > 1) {
    set_error("execute: direction must be L2C (0) or C2L (1)");
    return (size_t)-1;
  }
  if (stride == 0) {
    set_error("execute: stride must be >= 1");
    return (size_t)-1;
  }
  const bool planned = p->L == 0 ? (p->dplan->direction == BOTH || p->dplan->direction == direction)
                                 : p->t.A[direction] != nullptr;
  if (!planned) {
    set_error("execute: direction was not planned");
    return (size_t)-1;
  }
  const size_t N = p->N;
  hipStream_t st = p->stream;
  const bool dev = is_device_ptr(input_array) && is_device_ptr(output_array);
  int rc;
  if (dev) {
    rc = p->L == 0 ? run_direct(p->dplan, input_array, output_array, direction, stride, st,
                                p->precompute)
                   : run_fmm(p, input_array, output_array, direction, stride, st, 0);
    if (rc) return (size_t)-1;
    L2C_HIP(hipStreamSynchronize(st), (size_t)-1);
  } else {
    double *hin = p->L == 0 ? p->dplan->h_in : p->h_in;
    double *hout = p->L == 0 ? p->dplan->h_out : p->h_out;
    double *din = p->L == 0 ? p->dplan->d_in : p->d_in;
    double *dout = p->L == 0 ? p->dplan->d_out : p->d_out;
    for (size_t i = 0; i < N; i++) hin[i] = input_array[i * stride];
    L2C_HIP(hipMemcpyAsync(din, hin, N * sizeof(double), hipMemcpyHostToDevice, st), (size_t)-1);
    rc = p->L == 0 ? run_direct(p->dplan, din, dout, direction, 1, st, p->precompute)
                   : run_fmm(p, din, dout, direction, 1, st, 0);
    if (rc) return (size_t)-1;
    L2C_HIP(hipMemcpyAsync(hout, dout, N * sizeof(double), hipMemcpyDeviceToHost, st), (size_t)-1);
    L2C_HIP(hipStreamSynchronize(st), (size_t)-1);
    for (size_t i = 0; i < N; i++) output_array[i * stride] = hout[i];
  }
  return flops_fmm(p, direction);
}

size_t directM(const double *input_array, double *output_array, fmm_plan *p, const size_t strides) {
  clear_error();
  if (!p || !input_array || !output_array || strides == 0 || p->L == 0 || !p->t.a) {
    set_error("directM: needs an FMM plan (L >= 1), non-NULL arrays and stride >= 1");
    return (size_t)-1;
  }
  const size_t N = p->N;
  hipStream_t st = p->stream;
  if (is_device_ptr(input_array) && is_device_ptr(output_array)) {
    if (run_fmm(p, input_array, output_array, L2C, strides, st, 1)) return (size_t)-1;
    L2C_HIP(hipStreamSynchronize(st), (size_t)-1);
  } else {
    for (size_t i = 0; i < N; i++) p->h_in[i] = input_array[i * strides];
    L2C_HIP(hipMemcpyAsync(p->d_in, p->h_in, N * 8, hipMemcpyHostToDevice, st), (size_t)-1);
    if (run_fmm(p, p->d_in, p->d_out, L2C, 1, st, 1)) return (size_t)-1;
    L2C_HIP(hipMemcpyAsync(p->h_out, p->d_out, N * 8, hipMemcpyDeviceToHost, st), (size_t)-1);
    L2C_HIP(hipStreamSynchronize(st), (size_t)-1);
    for (size_t i = 0; i < N; i++) output_array[i * strides] = p->h_out[i];
  }
  // band flops only (src/C/leg2cheb.c:236-376)
  const size_t s = p->s, h = 2 * s, nL = N / h, per = p->precompute ? 2 : 3;
  size_t f = 0;
  for (size_t b = 0; b + 1 < nL; b++)
    for (size_t n = 0; n < 4 * s; n += 2) f += per * (n < h ? h : 2 * h - n);
  const size_t i0 = (nL - 1) * h;
  for (size_t n = 0; 2 * n + i0 < N; n++) f += per * (N - (2 * n + i0));
  return f + N;
}

size_t direct(const double *input_array, double *output_array, direct_plan *dp, size_t direction,
              size_t stride) {
  clear_error();
  if (!dp || !input_array || !output_array || direction > 1 || stride == 0) {
    set_error("direct: invalid argument");
    return (size_t)-1;
  }
  if (direction == C2L && !dp->d_an) {
    set_error("direct: C2L was not planned");
    return (size_t)-1;
  }
  const size_t N = dp->N;
  hipStream_t st = dp->stream;
  if (is_device_ptr(input_array) && is_device_ptr(output_array)) {
    if (run_direct(dp, input_array, output_array, direction, stride, st, dp->precompute))
      return (size_t)-1;
    L2C_HIP(hipStreamSynchronize(st), (size_t)-1);
  } else {
    for (size_t i = 0; i < N; i++) dp->h_in[i] = input_array[i * stride];
    L2C_HIP(hipMemcpyAsync(dp->d_in, dp->h_in, N * 8, hipMemcpyHostToDevice, st), (size_t)-1);
    if (run_direct(dp, dp->d_in, dp->d_out, direction, 1, st, dp->precompute)) return (size_t)-1;
    L2C_HIP(hipMemcpyAsync(dp->h_out, dp->d_out, N * 8, hipMemcpyDeviceToHost, st), (size_t)-1);
    L2C_HIP(hipStreamSynchronize(st), (size_t)-1);
    for (size_t i = 0; i < N; i++) output_array[i * stride] = dp->h_out[i];
  }
  return flops_direct(N, direction, dp->precompute);
}

int l2c_execute_device(fmm_plan *p, const double *d_in, double *d_out, size_t direction,
                       size_t stride, void *stream) {
  clear_error();
  if (!p || !d_in || !d_out || direction > 1 || stride == 0) {
    set_error("l2c_execute_device: invalid argument");
    return -1;
  }
  hipStream_t st = stream ? (hipStream_t)stream : p->stream;
  if (p->L == 0) return run_direct(p->dplan, d_in, d_out, direction, stride, st, p->precompute);
  if (!p->t.A[direction]) {
    set_error("l2c_execute_device: direction was not planned");
    return -1;
  }
  return run_fmm(p, d_in, d_out, direction, stride, st, 0);
}

int l2c_execute_lines_device(fmm_plan *p, const double *d_in, double *d_out, size_t direction,
                             size_t nlines, size_t n0, size_t ls0, size_t ls1, size_t es,
                             void *stream) {
  clear_error();
  if (!p || !d_in || !d_out || direction > 1 || es == 0 || nlines == 0) {
    set_error("l2c_execute_lines_device: invalid argument");
    return -1;
  }
  if (p->L > 0 && !p->t.A[direction]) {
    set_error("l2c_execute_lines_device: direction was not planned");
    return -1;
  }
  hipStream_t st = stream ? (hipStream_t)stream : p->stream;
  return run_lines(p, d_in, d_out, direction, nlines, n0, ls0, ls1, es, st);
}

// 2-D driver: src/C/leg2cheb.c:1236-1297, every line of an axis in one launch.
fmm_plan_2d *create_fmm_2d(size_t N0, size_t N1, int axis, size_t maxs, size_t M,
                           size_t direction, const size_t lagrange, const size_t precompute,
                           size_t v) {
  clear_error();
  if (axis < -1 || axis > 1 || N0 == 0 || N1 == 0) {
    set_error("create_fmm_2d: axis in {-1,0,1}, N0, N1 > 0");
    return nullptr;
  }
  fmm_plan_2d_s *q = new fmm_plan_2d_s();
  q->N0 = N0;
  q->N1 = N1;
  q->axis = axis;
  if (axis == 0 || axis == -1) q->p0 = create_fmm(N0, maxs, M, direction, lagrange, precompute, v);
  if (axis == 1) q->p1 = create_fmm(N1, maxs, M, direction, lagrange, precompute, v);
  if (axis == -1)
    q->p1 = N0 == N1 ? q->p0 : create_fmm(N1, maxs, M, direction, lagrange, precompute, v);
  if ((axis != 1 && !q->p0) || (axis != 0 && !q->p1)) {
    free_fmm_2d(q);
    return nullptr;
  }
  const size_t n = N0 * N1;
  if (hipMalloc((void **)&q->d_in, n * 8) != hipSuccess ||
      hipMalloc((void **)&q->d_out, n * 8) != hipSuccess ||
      (axis == -1 && hipMalloc((void **)&q->d_tmp, n * 8) != hipSuccess)) {
    set_error("create_fmm_2d: hipMalloc failed");
    free_fmm_2d(q);
    return nullptr;
  }
  return q;
}

void free_fmm_2d(fmm_plan_2d *q) {
  if (!q) return;
  if (q->p1 && q->p1 != q->p0) free_fmm(q->p1);
  if (q->p0) free_fmm(q->p0);
  if (q->d_in) (void)hipFree(q->d_in);
  if (q->d_out) (void)hipFree(q->d_out);
  if (q->d_tmp) (void)hipFree(q->d_tmp);
  delete q;
}

size_t execute2D(const double *input_array, double *output_array, fmm_plan_2d *q,
                 size_t direction) {
  clear_error();
  if (!q || !input_array || !output_array || direction > 1) {
    set_error("execute2D: invalid argument");
    return (size_t)-1;
  }
  const size_t N0 = q->N0, N1 = q->N1, n = N0 * N1;
  fmm_plan_s *pp = q->p0 ? q->p0 : q->p1;
  hipStream_t st = pp->stream;
  const bool dev = is_device_ptr(input_array) && is_device_ptr(output_array);
  const double *din = dev ? input_array : q->d_in;
  double *dout = dev ? output_array : q->d_out;
  if (!dev) L2C_HIP(hipMemcpy(q->d_in, input_array, n * 8, hipMemcpyHostToDevice), (size_t)-1);
  size_t flops = 0;
  if (q->axis == 0) {
    // N1 lines of length N0, element stride N1, consecutive lines adjacent
    if (run_lines(q->p0, din, dout, direction, N1, N1, 1, 0, N1, st)) return (size_t)-1;
    flops = N1 * flops_fmm(q->p0, direction);
  } else if (q->axis == 1) {
    if (run_lines(q->p1, din, dout, direction, N0, N0, N1, 0, 1, st)) return (size_t)-1;
    flops = N0 * flops_fmm(q->p1, direction);
  } else {
    if (q->p1->stream != st) L2C_HIP(hipStreamSynchronize(st), (size_t)-1);
    if (run_lines(q->p0, din, q->d_tmp, direction, N1, N1, 1, 0, N1, st)) return (size_t)-1;
    if (run_lines(q->p1, q->d_tmp, dout, direction, N0, N0, N1, 0, 1, st)) return (size_t)-1;
    flops = N1 * flops_fmm(q->p0, direction) + N0 * flops_fmm(q->p1, direction);
  }
  L2C_HIP(hipStreamSynchronize(st), (size_t)-1);
  if (!dev) L2C_HIP(hipMemcpy(output_array, q->d_out, n * 8, hipMemcpyDeviceToHost), (size_t)-1);
  return flops;
}

}  // extern "C"
