// l2c_fmm.hip -- FMM execute kernels for gfx950 (src/C/leg2cheb.c:1299-1500
// and the near field :236-535), plus the direct O(N^2) kernel (:152-234).
//
// One transform (or one line of a batch) is three launches, both parities in
// every work-group so that each M2L matrix is read once per transform:
//
//   k_up      leaf moments (step 1) and the upward pass (step 2) inside a
//             subtree of 2^gu leaf chunks; writes the multipole store for
//             levels >= lu and, for every level l < lu, the subtree root's
//             contribution P(lu-l, .) w_root to its level-l ancestor.
//   k_top     M2L (step 3) of every row chunk above level ld, one work-group
//             each; multipoles above lu are sums of the k_up contributions.
//   k_down    per subtree of 2^gd leaf chunks: M2L below ld, L2L (step 4)
//             down the ancestor chain and the subtree, the near-field band,
//             leaf evaluation (step 5), the final scaling and one coalesced
//             store of the output (which is overwritten, not accumulated).
//
// Workspace per line: [multipoles: slot_base(L) x 2M][contributions:
// lu x 2^(lu+2) x 2M][top M2L: slot_base(ld) x 2M].
//
// Kernels are templated on the rank M (MT = M compiled in so that the
// M-loops unroll; MT = 0 reads M at run time).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "l2c_internal.h"

namespace l2c {

constexpr double kSqrtPi = 1.77245385090551602729816748334;
constexpr double k2Pi = 0.63661977236758134307553505349;  // 2/pi

struct WsLayout {
  size_t w, part, mt, total;
};
__host__ __device__ inline WsLayout ws_layout(const Geo &g) {
  const size_t M2 = 2 * (size_t)g.M;
  WsLayout l;
  l.w = 0;
  l.part = slot_base(g.L) * M2;
  l.mt = l.part + (size_t)g.lu * ((size_t)4 << g.lu) * M2;
  l.total = l.mt + slot_base(g.ld) * M2;
  return l;
}
size_t ws_doubles(const Geo &g) { return ws_layout(g).total; }

// skewed LDS indices: one pad double every 4 (input) or 8 (Lambda window)
// so that lanes 4 (input) or 8 (Lambda) doubles apart hit distinct banks
__host__ __device__ __forceinline__ int sk4(int c) { return c + (c >> 2); }
__host__ __device__ __forceinline__ int sk8(int k) { return k + (k >> 3); }

// ---------------------------------------------------------------------------
// k_up: grid (2^(lu+2), lines), 256 threads.
size_t up_shared_bytes(const Geo &g) {
  const size_t G = (size_t)1 << g.gu, M2 = 2 * g.M;
  const bool tl = 2 * (size_t)g.s * g.M * 8 <= 24 * 1024;
  return 8 * (2 * G * g.s + (tl ? 2 * (size_t)g.s * g.M : 0) + 2 * (size_t)g.M * g.M + G * M2 +
              (G / 2 + 1) * M2);
}

template <int MT>
__global__ void __launch_bounds__(256) k_up(const double *__restrict__ in, LineMap lm, Geo g,
                                            int dir, const double *__restrict__ T,
                                            const double *__restrict__ U,
                                            const double *__restrict__ P,
                                            double *__restrict__ wsall, size_t wsstride) {
  extern __shared__ double sm[];
  const int M = MT > 0 ? MT : g.M;
  const int s = g.s, L = g.L, M2 = 2 * M, MM = M * M;
  const int G = 1 << g.gu;
  const int j = blockIdx.x;
  const int tid = threadIdx.x, NT = blockDim.x;
  const double *src = in + line_off(lm, blockIdx.y);
  double *ws = wsall + blockIdx.y * wsstride;
  const WsLayout lay = ws_layout(g);
  double *w = ws + lay.w;
  double *part = ws + lay.part;
  const bool tl = 2 * s * M * 8 <= 24 * 1024;
  double *xs = sm;                         // [2][G*s]
  double *Ts = xs + 2 * G * s;             // [2][s][M] (when it fits)
  double *Us = Ts + (tl ? 2 * s * M : 0);  // [2][M][M]
  double *w0 = Us + 2 * MM;                // [G][2][M]
  double *w1 = w0 + G * M2;                // [G/2+1][2][M]
  const double *Tp = tl ? Ts : T;

  const size_t base = (size_t)j * G * 2 * s;
  for (int e = tid; e < G * 2 * s; e += NT) {
    const size_t i = base + e;
    double v = 0;
    if (i < (size_t)g.N) {
      v = src[i * lm.es];
      if (dir == C2L && i > 0) v *= (double)i;  // v_i = i u_i (src/C/leg2cheb.c:1325-1335)
    }
    xs[(e & 1) * G * s + (e >> 1)] = v;
  }
  if (tl)
    for (int e = tid; e < 2 * s * M; e += NT) Ts[e] = T[e];
  for (int e = tid; e < 2 * MM; e += NT) Us[e] = U[e];
  __syncthreads();

  // step 1: w_c[k] = sum_m T_par[m][k] x_par[c s + m], four chunks per task
  const int CPT = G >= 4 ? 4 : G;
  const int nq = G / CPT;
  for (int task = tid; task < 2 * nq * M; task += NT) {
    const int k = task % M, q = (task / M) % nq, sg = task / (M * nq);
    const double *t = Tp + sg * s * M + k;
    const double *x = xs + sg * G * s + q * CPT * s;
    const size_t wl = (slot_base(L - 1) + (size_t)j * G + q * CPT) * M2 + sg * M + k;
    double *wq = w0 + (q * CPT * 2 + sg) * M + k;
    if (CPT == 4) {
      double acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
      for (int m = 0; m < s; m++) {
        const double tv = t[m * M];
        acc0 += tv * x[m];
        acc1 += tv * x[s + m];
        acc2 += tv * x[2 * s + m];
        acc3 += tv * x[3 * s + m];
      }
      wq[0] = acc0;
      wq[M2] = acc1;
      wq[2 * M2] = acc2;
      wq[3 * M2] = acc3;
      w[wl] = acc0;
      w[wl + M2] = acc1;
      w[wl + 2 * M2] = acc2;
      w[wl + 3 * M2] = acc3;
    } else {
      double acc0 = 0, acc1 = 0;
      for (int m = 0; m < s; m++) {
        const double tv = t[m * M];
        acc0 += tv * x[m];
        if (CPT > 1) acc1 += tv * x[s + m];
      }
      wq[0] = acc0;
      w[wl] = acc0;
      if (CPT > 1) {
        wq[M2] = acc1;
        w[wl + M2] = acc1;
      }
    }
  }
  __syncthreads();

  // step 2: parent = U0 child0 + U1 child1 inside the subtree
  int n = G;
  for (int lev = L - 2; lev >= g.lu; lev--) {
    n >>= 1;
    for (int o = tid; o < n * M2; o += NT) {
      const int c = o / M2, sg = (o / M) & 1, k = o % M;
      const double *x0 = w0 + (2 * c * 2 + sg) * M;
      const double *x1 = w0 + ((2 * c + 1) * 2 + sg) * M;
      const double *u0 = Us + k * M, *u1 = Us + MM + k * M;
      double acc = 0;
      for (int l = 0; l < M; l++) acc += u0[l] * x0[l] + u1[l] * x1[l];
      w1[o] = acc;
      w[(slot_base(lev) + (size_t)j * n + c) * M2 + (o % M2)] = acc;
    }
    __syncthreads();
    double *tmp = w0;
    w0 = w1;
    w1 = tmp;
  }

  // contributions of the root (level lu, chunk j) to its ancestors:
  // part[l][j] = P(lu-l, j mod 2^(lu-l)) w_root
  for (int task = tid; task < g.lu * M2; task += NT) {
    const int lev = task / M2, o = task % M2, sg = o / M, i = o % M;
    const int k = g.lu - lev;
    const size_t d = (size_t)j & (((size_t)1 << k) - 1);
    const double *pr = P + ((((size_t)1 << k) - 2) + d) * MM + (size_t)i * M;
    const double *x = w0 + sg * M;
    double acc = 0;
    for (int l = 0; l < M; l++) acc += pr[l] * x[l];
    part[((size_t)lev * ((size_t)4 << g.lu) + j) * M2 + o] = acc;
  }
}

// ---------------------------------------------------------------------------
// M2L of row chunk r at `level` for both parities (src/C/leg2cheb.c:1419-1435):
// block b = r/2; r even: A(b,0) w[r+2] + A(b,1) w[r+3]; r odd: A(b,2) w[r+2].
// w0, w1 point at the [2][M] multipoles of chunks r+2 and r+3.
template <int MT>
__device__ __forceinline__ void m2l_rowk(const double *__restrict__ A, int level, size_t r,
                                         int k, int Mr, const double *w0, const double *w1,
                                         double &out0, double &out1) {
  const int M = MT > 0 ? MT : Mr;
  const size_t MM = (size_t)M * M;
  const double *Ab = A + (3 * total_blocks(level) + 3 * (r >> 1)) * MM;
  if ((r & 1) == 0) {
    const double *a0 = Ab + (size_t)k * M, *a1 = Ab + MM + (size_t)k * M;
    double s00 = 0, s01 = 0, s10 = 0, s11 = 0;
#pragma unroll
    for (int l = 0; l < M; l++) {
      const double x0 = a0[l], x1 = a1[l];
      s00 += x0 * w0[l];
      s01 += x0 * w0[M + l];
      s10 += x1 * w1[l];
      s11 += x1 * w1[M + l];
    }
    out0 = s00 + s10;
    out1 = s01 + s11;
  } else {
    const double *a2 = Ab + 2 * MM + (size_t)k * M;
    double s0 = 0, s1 = 0;
#pragma unroll
    for (int l = 0; l < M; l++) {
      const double x = a2[l];
      s0 += x * w0[l];
      s1 += x * w0[M + l];
    }
    out0 = s0;
    out1 = s1;
  }
}

// ---------------------------------------------------------------------------
// k_top: grid (slot_base(ld), lines), 256 threads.  M2L of every row chunk
// above level ld; multipoles above lu are summed from the k_up contributions
// in a fixed order.
template <int MT>
__global__ void __launch_bounds__(256) k_top(Geo g, const double *__restrict__ A,
                                             double *__restrict__ wsall, size_t wsstride) {
  __shared__ double wv[2 * 96];
  __shared__ double red[256];
  const int M = MT > 0 ? MT : g.M, M2 = 2 * M;
  const int tid = threadIdx.x, NT = blockDim.x;
  double *ws = wsall + blockIdx.y * wsstride;
  const WsLayout lay = ws_layout(g);
  const int q = blockIdx.x;
  int lev = 0;
  while ((int)slot_base(lev + 1) <= q) lev++;
  const size_t r = (size_t)(q - (int)slot_base(lev));
  double *mt = ws + lay.mt + (size_t)q * M2;
  if (r >= ((size_t)4 << lev) - 2) {
    for (int o = tid; o < M2; o += NT) mt[o] = 0.0;
    return;
  }
  const int nc = (r & 1) ? 1 : 2;
  if (lev >= g.lu) {
    const double *w = ws + lay.w + (slot_base(lev) + r + 2) * M2;
    for (int o = tid; o < nc * M2; o += NT) wv[o] = w[o];
  } else {
    const int k = g.lu - lev;
    const size_t nd = (size_t)1 << k;
    const int ngrp = NT / M2;
    const int o = tid % M2, grp = tid / M2;
    for (int ci = 0; ci < nc; ci++) {
      double acc = 0;
      if (grp < ngrp) {
        const double *pp = ws + lay.part +
                           ((size_t)lev * ((size_t)4 << g.lu) + (r + 2 + ci) * nd) * M2 + o;
        for (size_t d = grp; d < nd; d += ngrp) acc += pp[d * M2];
      }
      red[tid] = acc;
      __syncthreads();
      if (tid < M2) {
        double sum = 0;
        for (int gi = 0; gi < ngrp; gi++) sum += red[gi * M2 + tid];
        wv[ci * M2 + tid] = sum;
      }
      __syncthreads();
    }
  }
  __syncthreads();
  for (int k = tid; k < M; k += NT) {
    double o0, o1;
    m2l_rowk<MT>(A, lev, r, k, M, wv, wv + M2, o0, o1);
    mt[k] = o0;
    mt[M + k] = o1;
  }
}

// ---------------------------------------------------------------------------
// Near-field band for four consecutive half-rows r0..r0+3 of one parity:
// columns c = r0 + t, t < len (len equal for the four rows), term
//   L2C: (a[t-rho] a[r+c+par]) x[c]            (src/C/leg2cheb.c:259,263)
//   C2L: -(dn[t-rho] x[c]) an[r+c+par]  or  -(dn an) x with precompute
//        (src/C/leg2cheb.c:416,426,471-486)
// a[t-rho] sits in a 4-register ring m, a[r+c+par] in a 4-register ring p;
// both rings are indexed by compile-time step numbers, eight steps per trip.
// xsg: parity's input window (skew sk4, origin at the work-group's first
// half-row), aw: Lambda window (skew sk8, origin at row0), hd: band head.
// ALIGNED: rl0 % 4 == 0, so every LDS offset inside a trip is a constant.
template <int DIR, int PRE, int SG, int ALIGNED>
__device__ __forceinline__ void band_rows(const double *__restrict__ xsg,
                                          const double *__restrict__ aw,
                                          const double *__restrict__ hd, int rl0, int len,
                                          double acc[4]) {
  const int pb = 2 * rl0 + SG;
  int OX[8], OP[12];
  if (ALIGNED) {
#pragma unroll
    for (int k = 0; k < 8; k++) OX[k] = k + (k >> 2);
#pragma unroll
    for (int k = 0; k < 12; k++) OP[k] = k + ((SG + k) >> 3);
  } else {
#pragma unroll
    for (int k = 0; k < 8; k++) OX[k] = sk4(rl0 + k) - sk4(rl0);
#pragma unroll
    for (int k = 0; k < 12; k++) OP[k] = sk8(pb + k) - sk8(pb);
  }
  const double *xb = xsg + sk4(rl0);
  const double *pbp = aw + sk8(pb);
  const double *hb = hd;
  double m[4] = {0, 0, 0, 0};
  double p[4] = {pbp[OP[0]], pbp[OP[1]], pbp[OP[2]], pbp[OP[3]]};
  double a[4] = {acc[0], acc[1], acc[2], acc[3]};
#define L2C_TERM(am, ap, x, acc_)   \
  if (DIR == 0)                     \
    acc_ += ((am) * (ap)) * (x);    \
  else if (PRE)                     \
    acc_ -= ((am) * (ap)) * (x);    \
  else                              \
    acc_ -= ((am) * (x)) * (ap);
#define BAND_STEP(K, XV)                                  \
  {                                                       \
    m[(K) & 3] = hb[K];                                   \
    const double x_ = (XV);                               \
    L2C_TERM(m[(K) & 3], p[(K) & 3], x_, a[0])            \
    L2C_TERM(m[((K) + 3) & 3], p[((K) + 1) & 3], x_, a[1]) \
    L2C_TERM(m[((K) + 2) & 3], p[((K) + 2) & 3], x_, a[2]) \
    L2C_TERM(m[((K) + 1) & 3], p[((K) + 3) & 3], x_, a[3]) \
    p[(K) & 3] = pbp[OP[(K) + 4]];                        \
  }
  const int nb = len >> 3, rem = len & 7;
  for (int b = 0; b < nb; b++) {
    BAND_STEP(0, xb[OX[0]]) BAND_STEP(1, xb[OX[1]]) BAND_STEP(2, xb[OX[2]])
    BAND_STEP(3, xb[OX[3]]) BAND_STEP(4, xb[OX[4]]) BAND_STEP(5, xb[OX[5]])
    BAND_STEP(6, xb[OX[6]]) BAND_STEP(7, xb[OX[7]])
    xb += 10;
    pbp += 9;
    hb += 8;
  }
  if (rem) {
    BAND_STEP(0, xb[OX[0]]) BAND_STEP(1, rem > 1 ? xb[OX[1]] : 0.0)
    BAND_STEP(2, rem > 2 ? xb[OX[2]] : 0.0) BAND_STEP(3, rem > 3 ? xb[OX[3]] : 0.0)
    BAND_STEP(4, rem > 4 ? xb[OX[4]] : 0.0) BAND_STEP(5, rem > 5 ? xb[OX[5]] : 0.0)
    BAND_STEP(6, rem > 6 ? xb[OX[6]] : 0.0)
  }
#undef BAND_STEP
#undef L2C_TERM
  acc[0] = a[0];
  acc[1] = a[1];
  acc[2] = a[2];
  acc[3] = a[3];
}

template <int DIR, int PRE>
__device__ __forceinline__ void band_dispatch(int sg, int aligned, const double *xsg,
                                              const double *aw, const double *hd, int rl0,
                                              int len, double acc[4]) {
  if (aligned) {
    if (sg == 0)
      band_rows<DIR, PRE, 0, 1>(xsg, aw, hd, rl0, len, acc);
    else
      band_rows<DIR, PRE, 1, 1>(xsg, aw, hd, rl0, len, acc);
  } else {
    if (sg == 0)
      band_rows<DIR, PRE, 0, 0>(xsg, aw, hd, rl0, len, acc);
    else
      band_rows<DIR, PRE, 1, 0>(xsg, aw, hd, rl0, len, acc);
  }
}

// ---------------------------------------------------------------------------
// k_down: grid (2^(ld+2), lines), 128 threads (2 waves).
constexpr int kR = 4;  // band rows per register block
constexpr int kDownNT = 128;

struct DownLds {
  int xs, aw, hd, D, cres, cb, wsub, mch, out, total;  // offsets in doubles
};
__host__ __device__ inline DownLds down_lds(const Geo &g) {
  const int G = 1 << g.gd, s = g.s, M2 = 2 * g.M;
  DownLds d;
  const int nx = G * s + s + 16;  // half-index window per parity (+ tail)
  const int nxs = nx + (nx >> 2) + 2;
  const int na = 2 * G * s + 2 * s + 32;  // Lambda window (full index)
  const int nas = na + (na >> 3) + 2;
  d.xs = 0;
  d.aw = d.xs + 2 * nxs;
  d.hd = d.aw + nas;
  d.D = d.hd + 2 * s + 16;
  d.cres = d.D + 2 * g.M * g.M;
  d.cb = d.cres + (2 * G - 1) * M2;
  d.wsub = d.cb + 2 * M2;
  d.mch = d.wsub + (2 * G - 1 + 2 * (g.gd + 1)) * M2;
  d.out = d.mch + std::max(g.ld, 1) * M2;
  d.total = d.out + 2 * G * s;
  return d;
}
size_t down_shared_bytes(const Geo &g) { return 8 * (size_t)down_lds(g).total; }

template <int DIR, int PRE, int MT>
__global__ void __launch_bounds__(kDownNT) k_down(
    const double *__restrict__ in, double *__restrict__ out, LineMap lm, Geo g, int band_only,
    const double *__restrict__ A, double *__restrict__ wsall, size_t wsstride,
    const double *__restrict__ Dg, const double *__restrict__ Tg, const double *__restrict__ tab,
    const double *__restrict__ htab) {
  extern __shared__ double sm[];
  const int M = MT > 0 ? MT : g.M;
  const int s = g.s, L = g.L, N = g.N, M2 = 2 * M, MM = M * M;
  const int G = 1 << g.gd;
  const int Rr = blockIdx.x;
  const int tid = threadIdx.x, NT = blockDim.x;
  const size_t row0 = (size_t)Rr * G * 2 * s;
  if (row0 >= (size_t)N) return;
  const int nr = (int)std::min((size_t)G * 2 * s, (size_t)N - row0);
  const size_t off = line_off(lm, blockIdx.y);
  const double *src = in + off;
  double *dst = out + off;
  const double *ws = wsall + blockIdx.y * wsstride;
  const WsLayout lay = ws_layout(g);
  const DownLds ld = down_lds(g);
  double *xs = sm + ld.xs;
  const int nxs = (ld.aw - ld.xs) / 2;
  double *aw = sm + ld.aw;
  double *hd = sm + ld.hd;
  double *Ds = sm + ld.D;
  double *cres = sm + ld.cres;
  double *cb = sm + ld.cb;
  double *wsub = sm + ld.wsub;
  double *mch = sm + ld.mch;
  double *ob = sm + ld.out;

  // ---- phase 0: windows (input split by parity, Lambda, band head), the
  // subtree's multipoles and the ancestor chain's M2L vectors
  const int nwin = G * 2 * s + 2 * s + 32;
  for (int e = tid; e < nwin; e += NT) {
    const size_t i = row0 + e;
    double v = 0;
    if (i < (size_t)N) {
      v = src[i * lm.es];
      if (DIR == 1 && i > 0) v *= (double)i;  // v_i = i u_i
    }
    xs[(e & 1) * nxs + sk4(e >> 1)] = v;
    aw[sk8(e)] = i < (size_t)N ? tab[i] : 0.0;
  }
  const int nh = std::min(2 * s, DIR == 0 ? N : (N + 1) / 2);
  for (int e = tid; e < 2 * s + 16; e += NT) hd[e] = e < nh ? htab[e] : 0.0;
  if (!band_only) {
    for (int e = tid; e < 2 * MM; e += NT) Ds[e] = Dg[e];
    const double *w = ws + lay.w;
    int sbase = 0;
    for (int lam = 0; lam <= g.gd; lam++) {
      const int lev = g.ld + lam, n = (1 << lam) + 2;
      const size_t c0 = ((size_t)Rr << lam) + 2, cmax = (size_t)4 << lev;
      for (int e = tid; e < n * M2; e += NT) {
        const size_t c = c0 + e / M2;
        wsub[sbase * M2 + e] = c < cmax ? w[(slot_base(lev) + c) * M2 + e % M2] : 0.0;
      }
      sbase += n;
    }
    const double *mt = ws + lay.mt;
    for (int e = tid; e < g.ld * M2; e += NT) {
      const int lev = e / M2;
      const size_t r = (size_t)Rr >> (g.ld - lev);
      mch[e] = mt[(slot_base(lev) + r) * M2 + e % M2];
    }
  }
  __syncthreads();

  // ---- phase 1: M2L of the subtree levels ld .. L-1 into cres
  if (!band_only) {
    const int ntask = (2 * G - 1) * M;
    for (int task = tid; task < ntask; task += NT) {
      const int k = task % M, cl = task / M;  // cl: chunk slot in the subtree
      int lam = 0;
      while ((2 << lam) - 1 <= cl) lam++;
      const int ci = cl - ((1 << lam) - 1);
      const int lev = g.ld + lam;
      const size_t r = ((size_t)Rr << lam) + ci;
      double o0 = 0, o1 = 0;
      if (r < ((size_t)4 << lev) - 2) {
        const double *wb = wsub + ((1 << lam) - 1 + 2 * lam + ci) * M2;
        m2l_rowk<MT>(A, lev, r, k, M, wb, wb + M2, o0, o1);
      }
      cres[cl * M2 + k] = o0;
      cres[cl * M2 + M + k] = o1;
    }
  }
  __syncthreads();

  // ---- phase 2: L2L down the ancestor chain, then inside the subtree
  if (!band_only) {
    double *cur = cb, *nxt = cb + M2;
    for (int o = tid; o < M2; o += NT) cur[o] = 0.0;
    __syncthreads();
    for (int lev = 0; lev < g.ld; lev++) {
      const size_t r = (size_t)Rr >> (g.ld - lev);
      const bool valid = r < ((size_t)4 << lev) - 2;
      const int p = (int)(r & 1);
      for (int o = tid; o < M2; o += NT) {
        const int sg = o / M, k = o % M;
        double v = 0;
        if (valid) {
          const double *dr = Ds + p * MM + k * M;
          const double *par = cur + sg * M;
#pragma unroll
          for (int jj = 0; jj < M; jj++) v += dr[jj] * par[jj];
          v += mch[lev * M2 + o];
        }
        nxt[o] = v;
      }
      __syncthreads();
      double *t = cur;
      cur = nxt;
      nxt = t;
    }
    // root chunk (level ld)
    if (g.ld > 0 && (size_t)Rr < ((size_t)4 << g.ld) - 2) {
      const int p = Rr & 1;
      for (int o = tid; o < M2; o += NT) {
        const int sg = o / M, k = o % M;
        const double *dr = Ds + p * MM + k * M;
        const double *par = cur + sg * M;
        double v = 0;
#pragma unroll
        for (int jj = 0; jj < M; jj++) v += dr[jj] * par[jj];
        cres[o] += v;
      }
    }
    __syncthreads();
    for (int lam = 1; lam <= g.gd; lam++) {
      const int n = 1 << lam, lev = g.ld + lam;
      const int b0 = n - 1, bp = (n >> 1) - 1;
      for (int o = tid; o < n * M2; o += NT) {
        const int ci = o / M2, oo = o % M2, sg = oo / M, k = oo % M;
        const size_t r = ((size_t)Rr << lam) + ci;
        if (r < ((size_t)4 << lev) - 2) {
          const double *dr = Ds + (int)(r & 1) * MM + k * M;
          const double *par = cres + (bp + (ci >> 1)) * M2 + sg * M;
          double v = 0;
#pragma unroll
          for (int jj = 0; jj < M; jj++) v += dr[jj] * par[jj];
          cres[(b0 + ci) * M2 + oo] += v;
        }
      }
      __syncthreads();
    }
  }

  // ---- phase 3: band + leaf evaluation, paired row groups per thread
  const int ngrp = (s + kR - 1) / kR, npair = (ngrp + 1) / 2;
  const int ntask = 2 * G * npair;
  const int nLm1 = g.nL - 1;
  const size_t nvalid = ((size_t)2 << L) - 2;
  const size_t c0 = (size_t)Rr * G * s;
  const int aligned = (s & 3) == 0;
  for (int task = tid; task < ntask; task += NT) {
    const int pp = task % npair, gsg = task / npair;
    const int gch = gsg % G, sg = gsg / G;
    const size_t beta = (size_t)Rr * G + gch;
    const int cnt = (N - sg + 1) / 2;  // elements of parity sg
    const size_t cend = beta < (size_t)nLm1 ? (beta + 2) * s : (size_t)cnt;
    for (int half = 0; half < 2; half++) {
      const int gam = half == 0 ? pp : ngrp - 1 - pp;
      if (half == 1 && gam == pp) break;
      const size_t r0 = beta * s + (size_t)gam * kR;
      if (r0 >= (size_t)cnt) continue;
      const int rl0 = (int)(r0 - c0);
      const int len = (int)(cend - r0);
      double acc[kR];
      const double *xsg = xs + sg * nxs;
      if (DIR == 0) {
        acc[0] = acc[1] = acc[2] = acc[3] = 0.0;
      } else {
#pragma unroll
        for (int rho = 0; rho < kR; rho++)
          acc[rho] = kSqrtPi * xsg[sk4(rl0 + rho)] * aw[sk8(2 * (rl0 + rho) + sg)];
      }
      band_dispatch<DIR, PRE>(sg, aligned, xsg, aw, hd, rl0, len, acc);
      // far field: local expansion of leaf chunk beta at the rows' nodes
      double far[kR] = {0, 0, 0, 0};
      if (!band_only && beta < nvalid) {
        const double *c = cres + (G - 1 + gch) * M2 + sg * M;
        if (g.nodal) {
          // Lagrange basis: far = sum_k T_par[m][k] c[k] with the plan's T
#pragma unroll
          for (int rho = 0; rho < kR; rho++) {
            const int ml = min(gam * kR + rho, s - 1);
            const double *tr = Tg + ((size_t)sg * s + ml) * M;
            double f = 0;
            for (int k = 0; k < M; k++) f += tr[k] * c[k];
            far[rho] = f;
          }
        } else {
          // Chebyshev basis: T_k(X) by the three-term recurrence
#pragma unroll
          for (int rho = 0; rho < kR; rho++) {
            const double X = -1 + (double)(2 * (gam * kR + rho) + sg) / (double)s;
            const double X2 = 2 * X;
            double tm = 1, tc = X;
            double f = c[0] + X * c[1];
            for (int k = 2; k < M; k++) {
              const double tn = tc * X2 - tm;
              f += tn * c[k];
              tm = tc;
              tc = tn;
            }
            far[rho] = f;
          }
        }
      }
#pragma unroll
      for (int rho = 0; rho < kR; rho++) {
        const int ml = gam * kR + rho;
        const size_t r = r0 + rho;
        if (ml < s && r < (size_t)cnt) {
          const size_t i = 2 * r + sg;
          double v;
          if (DIR == 0) {
            v = (far[rho] + acc[rho]) * k2Pi;
            if (i == 0) v *= 0.5;
          } else {
            v = (far[rho] + acc[rho]) * ((double)i + 0.5);
          }
          ob[i - row0] = v;
        }
      }
    }
  }
  __syncthreads();
  // ---- phase 4: coalesced store
  for (int e = tid; e < nr; e += NT) dst[(row0 + e) * lm.es] = ob[e];
}

// ---------------------------------------------------------------------------
// direct O(N^2) sums, one row per thread (src/C/leg2cheb.c:152-234)
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

// Workspace for a batch of lines: grown on demand, owned by the plan.
static int ensure_ws(fmm_plan_s *p, size_t nlines, double **w) {
  const size_t need = std::max<size_t>(nlines, 1) * ws_doubles(p->geo);
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

template <int MT>
static void launch_fmm_m(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir,
                         const LineMap &lm, size_t nlines, hipStream_t st, int band_only,
                         double *ws) {
  const Geo &g = p->geo;
  const size_t wsstride = ws_doubles(g);
  if (!band_only) {
    dim3 gu((unsigned)((size_t)4 << g.lu), (unsigned)nlines);
    prof_begin(p->prof, K_UP_SUBTREE, st);
    k_up<MT><<<gu, 256, up_shared_bytes(g), st>>>(d_in, lm, g, (int)dir, p->t.T, p->t.U, p->t.P,
                                                  ws, wsstride);
    prof_end(p->prof, st);
    if (g.ld > 0) {
      dim3 gt((unsigned)slot_base(g.ld), (unsigned)nlines);
      prof_begin(p->prof, K_UP_TOP, st);
      k_top<MT><<<gt, 256, 0, st>>>(g, p->t.A[dir], ws, wsstride);
      prof_end(p->prof, st);
    }
  }
  dim3 gd((unsigned)((size_t)4 << g.ld), (unsigned)nlines);
  const size_t shm = down_shared_bytes(g);
  const double *tab = dir == L2C ? p->t.a : p->t.an;
  const double *htab = dir == L2C ? p->t.a : p->t.dn;
  const double *A = p->t.A[dir];
  prof_begin(p->prof, K_DOWN_BAND, st);
  if (dir == L2C)
    k_down<0, 0, MT><<<gd, kDownNT, shm, st>>>(d_in, d_out, lm, g, band_only, A, ws, wsstride,
                                               p->t.D, p->t.T, tab, htab);
  else if (p->precompute)
    k_down<1, 1, MT><<<gd, kDownNT, shm, st>>>(d_in, d_out, lm, g, band_only, A, ws, wsstride,
                                               p->t.D, p->t.T, tab, htab);
  else
    k_down<1, 0, MT><<<gd, kDownNT, shm, st>>>(d_in, d_out, lm, g, band_only, A, ws, wsstride,
                                               p->t.D, p->t.T, tab, htab);
  prof_end(p->prof, st);
}

static int launch_fmm(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir,
                      const LineMap &lm, size_t nlines, hipStream_t st, int band_only) {
  double *ws = nullptr;
  if (ensure_ws(p, nlines, &ws)) return -1;
  if (p->geo.M == 18)
    launch_fmm_m<18>(p, d_in, d_out, dir, lm, nlines, st, band_only, ws);
  else
    launch_fmm_m<0>(p, d_in, d_out, dir, lm, nlines, st, band_only, ws);
  L2C_HIP(hipGetLastError(), -1);
  return 0;
}

int run_fmm(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir, size_t stride,
            hipStream_t st, int band_only) {
  LineMap lm{1, 0, 0, stride};
  return launch_fmm(p, d_in, d_out, dir, lm, 1, st, band_only);
}

int run_direct(const direct_plan_s *dp, const double *d_in, double *d_out, size_t dir,
               size_t stride, hipStream_t st, size_t precompute) {
  (void)precompute;
  LineMap lm{1, 0, 0, stride};
  const int N = (int)dp->N;
  prof_begin(dp->prof, K_DIRECT, st);
  k_direct<<<dim3((N + 255) / 256, 1), 256, 0, st>>>(d_in, d_out, lm, N, (int)dir, dp->d_a,
                                                     dp->d_an, dp->d_dn);
  prof_end(dp->prof, st);
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
      if (l0 % lm.n0 != 0 && nlines > 65535) {
        set_error("run_lines: batches above 65535 lines need n0 to divide 65535-line chunks");
        return -1;
      }
      const size_t base = (l0 / lm.n0) * lm.ls1 + (l0 % lm.n0) * lm.ls0;
      prof_begin(dp->prof, K_DIRECT, st);
      k_direct<<<dim3((N + 255) / 256, (unsigned)nl), 256, 0, st>>>(
          d_in + base, d_out + base, lm, N, (int)dir, dp->d_a, dp->d_an, dp->d_dn);
      prof_end(dp->prof, st);
      L2C_HIP(hipGetLastError(), -1);
    }
    return 0;
  }
  if (nlines > 65535) {
    set_error("run_lines: at most 65535 lines per call for FMM plans");
    return -1;
  }
  return launch_fmm(p, d_in, d_out, dir, lm, nlines, st, 0);
}

}  // namespace l2c
