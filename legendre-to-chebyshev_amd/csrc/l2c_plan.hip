// l2c_plan.hip -- plan construction for the gfx950 FMM Legendre<->Chebyshev
// transform: every table is built once and left in HBM.
//
//   a, an, dn   Lambda tables (k_fill_tables)          src/C/leg2cheb.c:758-795
//   A[dir]      M2L matrices: kernel samples at Chebyshev nodes followed by a
//               2-D DCT-II (k_lambda_minus, k_build_A)  src/C/leg2cheb.c:924-995
//   T           leaf Vandermonde / Lagrange matrix       src/C/leg2cheb.c:621-648, 1000-1031
//   U, D        child->parent and parent->child operators (B tables)
//                                                        src/C/leg2cheb.c:6-123, 1036-1105
//   P           composite child->ancestor operators for the top of the tree
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "l2c_internal.h"

namespace l2c {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
void clear_error() { g_err.clear(); }

// ---------------------------------------------------------------------------
// Host tables

// Exact Lambda(k), k < 256, from a double-double recurrence
// Lambda(k) = Lambda(k-1) (2k-1)/(2k), Lambda(0) = sqrt(pi); correctly rounded
// (the reference tabulates the same values, src/C/lambda.c:102-188).
static void lamint_table(double *t) {
  // pi as double-double, sqrt by one Newton step
  const double pi_hi = 3.141592653589793116, pi_lo = 1.2246467991473532e-16;
  double s0 = std::sqrt(pi_hi);
  double e = std::fma(-s0, s0, pi_hi) + pi_lo;
  double s1 = e / (2 * s0);
  double hi = s0 + s1, lo = s1 - (hi - s0);
  t[0] = hi + lo;
  for (int k = 1; k < 256; k++) {
    // (hi,lo) *= (2k-1)
    double m = 2.0 * k - 1;
    double p = hi * m;
    double pe = std::fma(hi, m, -p) + lo * m;
    hi = p + pe;
    lo = pe - (hi - p);
    // (hi,lo) /= 2k
    double d = 2.0 * k;
    double q1 = hi / d;
    double r = std::fma(-q1, d, hi) + lo;
    double q2 = r / d;
    hi = q1 + q2;
    lo = q2 - (hi - q1);
    t[k] = hi + lo;
  }
}

static const double *lamint() {
  static double tab[256];
  static bool done = false;
  if (!done) {
    lamint_table(tab);
    done = true;
  }
  return tab;
}

// Exact B(0): row k = Chebyshev coefficients of T_k((x-1)/2) (lower
// triangular; dyadic rationals).  Same matrix as BMe (src/C/leg2cheb.c:6-63)
// and W^-1 A(1/2,-1/2) W (src/python/binommat.py:33-38).
static void binom_B0(int M, double *B0) {
  std::fill(B0, B0 + (size_t)M * M, 0.0);
  if (M < 1) return;
  B0[0] = 1;
  if (M < 2) return;
  B0[M] = -0.5;
  B0[M + 1] = 0.5;
  std::vector<double> xp(M + 1);
  for (int k = 1; k + 1 < M; k++) {
    const double *Pk = &B0[(size_t)k * M], *Pm = &B0[(size_t)(k - 1) * M];
    double *Pn = &B0[(size_t)(k + 1) * M];
    std::fill(xp.begin(), xp.end(), 0.0);
    for (int l = 0; l <= k; l++) {
      if (Pk[l] == 0.0) continue;
      if (l == 0)
        xp[1] += Pk[0];
      else {
        xp[l + 1] += 0.5 * Pk[l];
        xp[l - 1] += 0.5 * Pk[l];
      }
    }
    for (int l = 0; l <= k + 1; l++) Pn[l] = xp[l] - (l <= k ? Pk[l] : 0.0) - (l < k ? Pm[l] : 0.0);
  }
}

static void matmul(int M, const double *X, const double *Y, double *Z) {
  for (int i = 0; i < M; i++)
    for (int j = 0; j < M; j++) {
      double acc = 0;
      for (int k = 0; k < M; k++) acc += X[i * M + k] * Y[k * M + j];
      Z[i * M + j] = acc;
    }
}

// ---------------------------------------------------------------------------
// Device plan kernels

// a[k] = Lambda(k); an, dn for C2L.  One thread per group of 8 indices:
// below 256 the exact table, beyond it the series value at the group head and
// a[j] = a[j-1] - a[j-1]/2/j inside the group (src/C/leg2cheb.c:767-779).
__global__ void k_fill_tables(double *a, double *an, int N,
                              const double *__restrict__ lamint) {
#pragma clang fp contract(off)
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0 = g * 8;
  if (i0 >= N) return;
  const int i1 = min(i0 + 8, N);
  if (i0 < 256) {
    for (int i = i0; i < i1; i++) a[i] = lamint[i];
  } else {
    double v = lambda((double)i0);
    a[i0] = v;
    for (int j = i0 + 1; j < i1; j++) {
      v = v - 0.5 * v / (double)j;
      a[j] = v;
    }
  }
  if (an) {
    for (int i = i0; i < i1; i++)
      an[i] = i == 0 ? 1.12837916709551257390 /* 2/sqrt(pi) */
                     : 1 / (a[i] * (double)(2 * (size_t)i * i + i));
  }
}

// dn needs a[j-1] for j < (N+1)/2: run after a[] is complete.
__global__ void k_fill_dn(const double *a, double *dn, int n) {
#pragma clang fp contract(off)
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  dn[j] = j == 0 ? 0.0 : a[j - 1] / (double)(2 * j);
}

// Dense parity blocks of the direct sums (src/C/leg2cheb.c:152-234) for
// batched lines: out[2r+p] = sum_c K[p][r][c] u[2c+p], r, c < NPP, with the
// final scaling (2/pi and b_0/2 for L2C; (i+1/2) for C2L) and the C2L input
// weights v_j = j u_j (v_0 = u_0) folded in; zero below the diagonal and past
// the length.  Stored in MFMA A-operand order: the 16 x 4 tile (row tile
// R = r/16, k-step c/4) of parity p is 64 consecutive doubles, lane
// (r % 16) + 16 (c % 4), at ((p NR + R) (NPP/4) + c/4) 64, NR = NPP/16.
// Grid (NPP, 2), one row per work-group.
__global__ void k_build_dense(double *K, const double *__restrict__ a,
                              const double *__restrict__ an, const double *__restrict__ dn,
                              int N, int NPP, int dir) {
#pragma clang fp contract(off)
  const int r = blockIdx.x, p = blockIdx.y, i = 2 * r + p;
  for (int c = threadIdx.x; c < NPP; c += blockDim.x) {
    const int j = 2 * c + p;
    double v = 0.0;
    if (c >= r && j < N) {
      if (dir == L2C) {
        v = 0.63661977236758134307553505349 * a[c - r] * a[r + c + p];
        if (i == 0) v *= 0.5;
      } else {
        const double wj = j == 0 ? 1.0 : (double)j;
        v = c == r ? 1.77245385090551602729816748334 * an[i] * wj
                   : -(dn[c - r] * an[r + c + p]) * wj;
        v *= (double)i + 0.5;
      }
    }
    const size_t tile = ((size_t)p * (NPP >> 4) + (r >> 4)) * (NPP >> 2) + (c >> 2);
    K[tile * 64 + (r & 15) + 16 * (c & 3)] = v;
  }
}

// Lambda((y-x)/2) for block 0 of each level, q-p in {0,1}: it depends on y-x
// only, so the reference computes it once per level and reuses it for every
// block (src/C/leg2cheb.c:948-955).  Entries with j >= M-i take the value at
// the persymmetric partner, as the reference stores it.
__global__ void k_lambda_minus(double *lm, const double *__restrict__ xj, int L, int s,
                               int M) {
#pragma clang fp contract(off)
  const int level = blockIdx.x >> 1, qd = blockIdx.x & 1;
  const size_t h = (size_t)s << (L - level - 1);
  const size_t x0 = h;                              // block 0, p = 0
  const size_t y0 = 2 * (2 * h + qd * h) + h;       // q = q-p
  const int MM = M * M;
  for (int e = threadIdx.x; e < MM; e += blockDim.x) {
    int i = e / M, j = e % M;
    if (j >= M - i) {
      const int ti = M - 1 - j;
      j = M - 1 - i;
      i = ti;
    }
    const double x = (double)x0 + xj[i] * (double)h;
    const double y = (double)y0 + xj[j] * (double)h;
    lm[((size_t)level * 2 + qd) * MM + e] = lambda((y - x) / 2);
  }
}

// One work-group per submatrix (level, block, r): kernel samples, then the
// scaled 2-D DCT-II (modal) or the samples themselves (nodal).
// L2C kernel: Lambda((y-x)/2) Lambda((x+y)/2); C2L kernel:
// 2 Lambda((y-x)/2) / (Lambda((x+y)/2) (x+y)(x+y+1)(x-y+1))
// (src/C/leg2cheb.c:964-969).  DCT-II: (4/M^2) C F C^T, row 0 and column 0
// halved (src/C/dct.c:582-598).
// Block blockIdx.x of A (3 M x M, rows sub M + row) into At with its rows
// across lanes: element (r, col) at col 3M + r.
__global__ void k_rows_across_lanes(const double *__restrict__ A, double *__restrict__ At, int M) {
  const size_t n = 3 * (size_t)M * M, base = blockIdx.x * n;
  for (int e = threadIdx.x; e < (int)n; e += blockDim.x) {
    const int r = e / M, col = e % M;
    At[base + (size_t)col * 3 * M + r] = A[base + e];
  }
}

__global__ void k_build_A(double *A0, double *A1, const double *__restrict__ lm,
                          const double *__restrict__ xj, const double *__restrict__ ctab,
                          int L, int s, int M, int lagrange) {
#pragma clang fp contract(off)
  extern __shared__ double sm[];
  const int MM = M * M;
  double *F0 = sm, *F1 = sm + MM, *tmp = sm + 2 * MM;
  const size_t kk = blockIdx.x;
  int level = 0;
  while (level + 1 < L && 3 * total_blocks(level + 1) <= kk) level++;
  const size_t rel = kk - 3 * total_blocks(level);
  const size_t b = rel / 3;
  const int r = (int)(rel % 3);
  const int p = r == 2 ? 1 : 0, q = r == 0 ? 0 : 1;
  const size_t h = (size_t)s << (L - level - 1);
  const size_t ij0 = 2 * b * h, ij1 = ij0 + 2 * h;
  const size_t x0 = 2 * (ij0 + p * h) + h, y0 = 2 * (ij1 + q * h) + h;
  const double *lmq = &lm[((size_t)level * 2 + (q - p)) * MM];
  for (int e = threadIdx.x; e < MM; e += blockDim.x) {
    const int i = e / M, j = e % M;
    const int ci = min(i, j), cj = max(i, j);  // Lambda((x+y)/2) is symmetric
    const double xs = (double)x0 + xj[ci] * (double)h;
    const double ys = (double)y0 + xj[cj] * (double)h;
    const double m1 = lambda((xs + ys) / 2);
    const double x = (double)x0 + xj[i] * (double)h;
    const double y = (double)y0 + xj[j] * (double)h;
    const double fpq = lmq[e];
    if (A0) F0[e] = fpq * m1;
    if (A1) F1[e] = 2 * (fpq / (m1 * (x + y) * (x + y + 1) * (x - y + 1)));
  }
  __syncthreads();
  for (int d = 0; d < 2; d++) {
    double *Ad = d == 0 ? A0 : A1;
    if (!Ad) continue;
    const double *F = d == 0 ? F0 : F1;
    double *out = &Ad[kk * MM];
    if (lagrange) {
      for (int e = threadIdx.x; e < MM; e += blockDim.x) out[e] = F[e];
      continue;
    }
    {
#pragma clang fp contract(fast)
      for (int e = threadIdx.x; e < MM; e += blockDim.x) {
        const int i = e / M, l = e % M;
        double acc = 0;
        for (int j = 0; j < M; j++) acc += F[i * M + j] * ctab[l * M + j];
        tmp[e] = acc;
      }
      __syncthreads();
      const double sc = 4.0 / ((double)M * (double)M);
      for (int e = threadIdx.x; e < MM; e += blockDim.x) {
        const int k = e / M, l = e % M;
        double acc = 0;
        for (int i = 0; i < M; i++) acc += ctab[k * M + i] * tmp[i * M + l];
        double v = acc * sc;
        if (k == 0) v *= 0.5;
        if (l == 0) v *= 0.5;
        out[e] = v;
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Host plan construction

template <typename T>
static int dmalloc(T **p, size_t n, size_t *acct) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return -1;
  }
  *p = (T *)q;
  if (acct) *acct += std::max<size_t>(n, 1) * sizeof(T);
  return 0;
}

static void dfree(void *p) {
  if (p) (void)hipFree(p);
}

// ---------------------------------------------------------------------------
// Split form of the M2L matrices (precision form 1)

// Largest |coefficient| of each block (3 M^2 values); one wave per block.
__global__ void k_block_absmax(const double *__restrict__ A, double *bmax, int n3) {
  const size_t b = blockIdx.x;
  double m = 0.0;
  for (int e = threadIdx.x; e < n3; e += 64) m = fmax(m, fabs(A[b * n3 + e]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if (threadIdx.x == 0) bmax[b] = m;
}

// Per coefficient position e: the largest |A_e| / (block max) over the blocks
// b = blockIdx.y, blockIdx.y + gridDim.y, ...; partial maxima per grid row.
__global__ void k_pos_ratio(const double *__restrict__ A, const double *__restrict__ bmax,
                            double *part, int nb, int n3) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n3) return;
  double m = 0.0;
  for (int b = blockIdx.y; b < nb; b += gridDim.y) {
    const double d = bmax[b];
    if (d > 0) m = fmax(m, fabs(A[(size_t)b * n3 + e]) / d);
  }
  part[(size_t)blockIdx.y * n3 + e] = m;
}

__device__ inline unsigned short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// One record per block: slot i holds coefficient src[i] of the block in its
// tier's type.
__global__ void k_pack_split(const double *__restrict__ A, unsigned char *rec,
                             const unsigned short *__restrict__ src, int n64, int n32, int n16,
                             int n3, int rec_bytes) {
  const size_t b = blockIdx.x;
  const double *Ab = A + b * n3;
  unsigned char *r = rec + b * rec_bytes;
  double *r64 = (double *)r;
  float *r32 = (float *)(r + 8 * (size_t)n64);
  unsigned short *r16 = (unsigned short *)(r + 8 * (size_t)n64 + 4 * (size_t)n32);
  for (int i = threadIdx.x; i < n64 + n32 + n16; i += blockDim.x) {
    const double v = Ab[src[i]];
    if (i < n64)
      r64[i] = v;
    else if (i < n64 + n32)
      r32[i - n64] = (float)v;
    else
      r16[i - n64 - n32] = bf16_rne((float)v);
  }
}

// Storage tier of a coefficient position from its largest magnitude relative
// to its block's largest, so that the stored value's error stays below
// 2^-53 of the block's largest coefficient: FP32 (24-bit significand) below
// 2^-29, BF16 (8-bit) below 2^-44, dropped below 2^-53.
static int split_tier(double ratio) {
  if (ratio > std::ldexp(1.0, -29)) return 0;
  if (ratio > std::ldexp(1.0, -44)) return 1;
  if (ratio > std::ldexp(1.0, -53)) return 2;
  return 3;
}

static int build_split(fmm_plan_s *p, int dir) {
  hipStream_t st = p->stream;
  const int M = (int)p->M, Mp = M + 1, n3 = 3 * M * M;
  const int nb = (int)p->Nb;
  const double *A = p->t.A[dir];
  double *bmax = nullptr, *part = nullptr;
  const int rows = std::min(nb, 256);
  if (dmalloc(&bmax, (size_t)nb, nullptr) || dmalloc(&part, (size_t)rows * n3, nullptr)) {
    dfree(bmax);
    return -1;
  }
  k_block_absmax<<<nb, 64, 0, st>>>(A, bmax, n3);
  k_pos_ratio<<<dim3((n3 + 255) / 256, rows), 256, 0, st>>>(A, bmax, part, nb, n3);
  std::vector<double> hp((size_t)rows * n3);
  const hipError_t e1 = hipMemcpyAsync(hp.data(), part, hp.size() * 8, hipMemcpyDeviceToHost, st);
  const hipError_t e2 = hipStreamSynchronize(st);
  dfree(bmax);
  dfree(part);
  if (e1 != hipSuccess || e2 != hipSuccess) {
    set_error("build_split: ratio reduction failed");
    return -1;
  }
  std::vector<int> tier(n3);
  int cnt[4] = {0, 0, 0, 0};
  for (int e = 0; e < n3; e++) {
    double r = 0;
    for (int y = 0; y < rows; y++) r = std::max(r, hp[(size_t)y * n3 + e]);
    tier[e] = split_tier(r);
    cnt[tier[e]]++;
  }
  std::vector<unsigned short> src, pos;
  for (int t = 0; t < 4; t++)
    for (int e = 0; e < n3; e++)
      if (tier[e] == t) {
        src.push_back((unsigned short)e);
        pos.push_back((unsigned short)((e / M) * Mp + e % M));
      }
  DevTables::Split &sp = p->t.sp[dir];
  sp.n64 = cnt[0];
  sp.n32 = cnt[1];
  sp.n16 = cnt[2];
  sp.rec_bytes = (8 * cnt[0] + 4 * cnt[1] + 2 * cnt[2] + 15) / 16 * 16;
  unsigned short *d_src = nullptr;
  auto fail = [&]() {
    dfree(d_src);
    dfree(sp.rec);
    dfree(sp.pos);
    sp.rec = nullptr;
    sp.pos = nullptr;
    return -1;
  };
  if (dmalloc(&sp.rec, (size_t)nb * sp.rec_bytes, &p->hbm_bytes) ||
      dmalloc(&sp.pos, (size_t)n3, &p->hbm_bytes) || dmalloc(&d_src, (size_t)n3, nullptr))
    return fail();
  L2C_HIP(hipMemcpyAsync(sp.pos, pos.data(), n3 * sizeof(unsigned short), hipMemcpyHostToDevice, st), fail());
  L2C_HIP(hipMemcpyAsync(d_src, src.data(), n3 * sizeof(unsigned short), hipMemcpyHostToDevice, st), fail());
  L2C_HIP(hipMemsetAsync(sp.rec, 0, (size_t)nb * sp.rec_bytes, st), fail());
  k_pack_split<<<nb, 256, 0, st>>>(A, sp.rec, d_src, sp.n64, sp.n32, sp.n16, n3, sp.rec_bytes);
  L2C_HIP(hipGetLastError(), fail());
  L2C_HIP(hipStreamSynchronize(st), fail());
  dfree(d_src);
  return 0;
}


// Lambda tables on the device: a (N), and an (N), dn ((N+1)/2) when C2L.
static int build_lambda_tables(size_t N, bool c2l, double **a, double **an, double **dn,
                               hipStream_t st, size_t *acct) {
  if (dmalloc(a, N + 8, acct)) return -1;
  if (c2l) {
    if (dmalloc(an, N + 8, acct)) return -1;
    if (dmalloc(dn, (N + 1) / 2 + 8, acct)) return -1;
  }
  double *d_lamint = nullptr;
  if (dmalloc(&d_lamint, 256, nullptr)) return -1;
  L2C_HIP(hipMemcpyAsync(d_lamint, lamint(), 256 * sizeof(double), hipMemcpyHostToDevice, st), -1);
  const int groups = (int)((N + 7) / 8);
  k_fill_tables<<<(groups + 255) / 256, 256, 0, st>>>(*a, c2l ? *an : nullptr, (int)N,
                                                      d_lamint);
  L2C_HIP(hipGetLastError(), -1);
  if (c2l) {
    const int nd = (int)((N + 1) / 2);
    k_fill_dn<<<(nd + 255) / 256, 256, 0, st>>>(*a, *dn, nd);
    L2C_HIP(hipGetLastError(), -1);
  }
  L2C_HIP(hipStreamSynchronize(st), -1);
  dfree(d_lamint);
  return 0;
}

direct_plan_s *make_direct(size_t N, size_t direction, size_t precompute) {
  direct_plan_s *dp = new direct_plan_s();
  dp->N = N;
  dp->direction = direction;
  dp->precompute = precompute;
  if (const char *e = getenv("L2C_DIRECT_LINES_Y")) {
    const long v = atol(e);
    if (v >= 1 && v <= 65535) dp->lines_y = (unsigned)v;
  }
  if (const char *e = getenv("L2C_DIRECT_BAND")) dp->direct_band = atoi(e) != 0;
  (void)hipGetDevice(&dp->device);
  if (hipStreamCreateWithFlags(&dp->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    delete dp;
    return nullptr;
  }
  const bool c2l = direction == C2L || direction == BOTH;
  if (build_lambda_tables(N, c2l, &dp->d_a, &dp->d_an, &dp->d_dn, dp->stream, nullptr)) {
    destroy_direct(dp);
    return nullptr;
  }
  if (N <= (size_t)kGemmMaxN) {
    dp->NPP = (int)(((N + 1) / 2 + 15) / 16 * 16);
    const size_t nk = 2 * (size_t)dp->NPP * dp->NPP;
    for (int d = 0; d < 2; d++) {
      if ((d == L2C && direction == C2L) || (d == C2L && !c2l)) continue;
      if (dmalloc(&dp->d_K[d], nk, nullptr)) {
        destroy_direct(dp);
        return nullptr;
      }
      k_build_dense<<<dim3(dp->NPP, 2), 64, 0, dp->stream>>>(dp->d_K[d], dp->d_a, dp->d_an,
                                                             dp->d_dn, (int)N, dp->NPP, d);
      L2C_HIP(hipGetLastError(), (destroy_direct(dp), nullptr));
    }
    L2C_HIP(hipStreamSynchronize(dp->stream), (destroy_direct(dp), nullptr));
  }
  return dp;
}

void destroy_direct(direct_plan_s *dp) {
  if (!dp) return;
  dfree(dp->d_a);
  dfree(dp->d_an);
  dfree(dp->d_dn);
  dfree(dp->d_K[0]);
  dfree(dp->d_K[1]);
  dfree(dp->d_in);
  dfree(dp->d_out);
  if (dp->h_in) (void)hipHostFree(dp->h_in);
  if (dp->h_out) (void)hipHostFree(dp->h_out);
  if (dp->stream) (void)hipStreamDestroy(dp->stream);
  delete dp;
}

static void destroy_fmm(fmm_plan_s *p) {
  if (!p) return;
  // nothing may still read the tables or the workspace
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  drop_graphs(p);
  dfree(p->t.a);
  dfree(p->t.an);
  dfree(p->t.dn);
  dfree(p->t.A[0]);
  dfree(p->t.A[1]);
  dfree(p->t.At[0]);
  dfree(p->t.At[1]);
  for (int d = 0; d < 2; d++) {
    dfree(p->t.sp[d].rec);
    dfree(p->t.sp[d].pos);
    dfree(p->t.sl[d].q);
    dfree(p->t.sl[d].sc);
  }
  dfree(p->t.T);
  dfree(p->t.U);
  dfree(p->t.D);
  dfree(p->t.P);
  dfree(p->t.Q);
  dfree(p->t.head[0]);
  dfree(p->t.head[1]);
  dfree(p->d_w);
  dfree(p->d_wl);
  for (double *r : p->wl_retired) dfree(r);
  dfree(p->d_in);
  dfree(p->d_out);
  if (p->h_in) (void)hipHostFree(p->h_in);
  if (p->h_out) (void)hipHostFree(p->h_out);
  if (p->dplan) p->dplan->prof = nullptr;
  destroy_direct(p->dplan);
  if (p->prof) {
    for (auto e : p->prof->pool) (void)hipEventDestroy(e);
    delete p->prof;
  }
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

// Work-group shapes: leaf chunks per upward / downward work-group, shrunk
// until the work-group's LDS fits the budget.
// single: the cuts of single-line calls, else those of line calls.  Line
// calls (their grids hold every line's work-groups) take the largest
// subtrees; single transforms take cuts that grow with the depth, gu = gd =
// (L - 4) / 2 within [1, 4] (gd >= 2 for k_down8r), measured best at
// L = 4 ... 12 (s = 64: 2^16 23.1 us at (2, 2) against 26.8 us at (4, 4);
// 2^18 26.6 against 28.6 us at (3, 3); 2^20 (4, 4) best), while line calls
// lose with them (64 lines of 2^16: 174 K against 432 K lines/s).
static bool choose_cuts(Geo &g, bool single) {
  const int Lm1 = g.L - 1;
  // k_down8 (s = 16, 32, 64; L2C_DOWN8=0 selects k_down): 16 leaf chunks and
  // up to 80 KiB of LDS (two work-groups per CU); k_down: 8 and 64 KiB
  const char *e8 = getenv("L2C_DOWN8");
  g.fast = down8_eligible(g) && !(e8 && atoi(e8) == 0) ? 1 : 0;
  const int dcut = (g.L - 4) / 2;
  int gu = 4, gd = g.fast ? 4 : 3;
  if (single) {
    gu = std::max(1, std::min(4, dcut));
    if (g.fast) gd = std::max(2, std::min(4, dcut));
  }
  if (const char *e = getenv("L2C_GU")) gu = atoi(e);
  if (const char *e = getenv("L2C_GD")) gd = atoi(e);
  // k_up holds its composite operators in LDS (lu of them): up to 80 KiB
  const size_t budget = 80 * 1024, dbudget = (g.fast ? 80 : 64) * 1024;
  g.gu = std::max(0, std::min(std::min(gu, 4), Lm1));  // k_up: at most 16 leaf chunks
  g.gd = std::max(0, std::min(std::min(gd, g.fast ? 4 : Lm1), Lm1));
  for (;;) {
    g.lu = Lm1 - g.gu;
    g.ld = Lm1 - g.gd;
    bool shrink = false;
    if (up_shared_bytes(g) > budget && g.gu > 0) {
      g.gu--;
      shrink = true;
    } else if (down_shared_bytes(g) > dbudget && g.gd > 0) {
      g.gd--;
      shrink = true;
    }
    if (!shrink) break;
  }
  return up_shared_bytes(g) <= budget && down_shared_bytes(g) <= dbudget;
}

fmm_plan_s *create_fmm_impl(size_t N, size_t maxs, size_t M, size_t direction, size_t lagrange,
                            size_t precompute, size_t v) {
  clear_error();
  if (N == 0 || maxs == 0 || direction > 2) {
    set_error("create_fmm: need N > 0, maxs > 0, direction in {0,1,2}");
    return nullptr;
  }
  if (N > ((size_t)1 << 30)) {
    set_error("create_fmm: N > 2^30 not supported");
    return nullptr;
  }
  fmm_plan_s *p = new fmm_plan_s();
  p->N = N;
  p->maxs = maxs;
  p->M = M;
  p->direction = direction;
  p->lagrange = lagrange;
  p->precompute = precompute;
  if (const char *e = getenv("L2C_DOWN_SPLIT")) p->down_split = std::max(0, std::min(2, atoi(e)));
  if (const char *e = getenv("L2C_M2L_MMA")) p->m2l_mma_lines = std::max(0, atoi(e));
  if (const char *e = getenv("L2C_DOWN_ROLES")) p->down_roles = atoi(e) != 0;
  if (const char *e = getenv("L2C_ROLES_SHIFT")) p->roles_shift = atoi(e);
  if (const char *e = getenv("L2C_DIRECT_MAXN")) p->direct_max_n = std::max(0, atoi(e));
  if (const char *e = getenv("L2C_DIRECT_BAND")) p->direct_band = atoi(e) != 0;
  if (const char *e = getenv("L2C_DIRECT_LINES")) p->direct_line_elems = std::max(0, atoi(e));
  if (const char *e = getenv("L2C_M2L_T")) p->m2l_t = atoi(e) != 0;
  (void)hipGetDevice(&p->device);
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    destroy_fmm(p);
    return nullptr;
  }
  // geometry: src/C/leg2cheb.c:860-871
  const int L = (int)std::ceil(std::log2((double)N / (double)maxs)) - 2;
  if (L < 1) {
    if (v > 1) printf("Levels < 1. Using only direct method\n");
    p->L = 0;
    p->Nn = N;
    p->dplan = make_direct(N, direction, precompute);
    if (!p->dplan) {
      destroy_fmm(p);
      return nullptr;
    }
    return p;
  }
  if (M < 2 || M > 48) {
    set_error("create_fmm: M must be in [2, 48]");
    destroy_fmm(p);
    return nullptr;
  }
  const size_t s = (size_t)std::ceil((double)N / (double)((size_t)1 << (L + 2)));
  const size_t Nn = s << (L + 2);
  if (s > 512) {
    set_error("create_fmm: leaf size s > 512 (maxs > 512) not supported");
    destroy_fmm(p);
    return nullptr;
  }
  p->L = L;
  p->s = s;
  p->Nn = Nn;
  p->Nb = total_blocks(L);
  p->Nc = total_sub(L);
  Geo &g = p->geo;
  g.N = (int)N;
  g.Nn = (int)Nn;
  g.L = L;
  g.s = (int)s;
  g.M = (int)M;
  g.nL = (int)(N / (2 * s));
  g.nodal = lagrange ? 1 : 0;
  p->geo1 = g;
  if (!choose_cuts(g, false) || !choose_cuts(p->geo1, true)) {
    set_error("create_fmm: leaf size s and rank M need more than 64 KiB of LDS per work-group");
    destroy_fmm(p);
    return nullptr;
  }
  // the composite operators serve both geometries
  p->lu_tab = std::max(g.lu, p->geo1.lu);
  p->ld_tab = std::max(g.ld, p->geo1.ld);
  if (v > 1) {
    printf("N %zu\nNum levels %d\nNum submatrices %zu\nNum blocks %zu\n", N, L, p->Nc, p->Nb);
    printf("Given max s %zu \nComputed s %zu \nComputed N %zu\nLagrange %zu\n", maxs, s, Nn,
           lagrange);
    printf("Precomputed matrix entries %zu\n", precompute);
  }
  hipStream_t st = p->stream;
  size_t *acct = &p->hbm_bytes;
  const bool l2c = direction == L2C || direction == BOTH;
  const bool c2l = direction == C2L || direction == BOTH;
  if (build_lambda_tables(N, c2l, &p->t.a, &p->t.an, &p->t.dn, st, acct)) {
    destroy_fmm(p);
    return nullptr;
  }
  // band heads, read as uniform (scalar) loads by the band loop
  for (int d = 0; d < 2; d++) {
    if ((d == 0 && !l2c) || (d == 1 && !c2l)) continue;
    const size_t nhd = 2 * s + 16, nh = std::min(2 * s, d == 0 ? N : (N + 1) / 2);
    if (dmalloc(&p->t.head[d], nhd, acct)) {
      destroy_fmm(p);
      return nullptr;
    }
    L2C_HIP(hipMemsetAsync(p->t.head[d], 0, nhd * sizeof(double), st), (destroy_fmm(p), nullptr));
    L2C_HIP(hipMemcpyAsync(p->t.head[d], d == 0 ? p->t.a : p->t.dn, nh * sizeof(double),
                           hipMemcpyDeviceToDevice, st),
            (destroy_fmm(p), nullptr));
  }

  // host tables: nodes, cosine table, T, U, D, P
  const int Mi = (int)M;
  const size_t MM = M * M;
  std::vector<double> xj(M), ctab(MM), T(2 * s * M), B0(MM), U(2 * MM), D(2 * MM);
  for (int i = 0; i < Mi; i++) xj[i] = std::cos((i + 0.5) * M_PI / Mi);
  for (int k = 0; k < Mi; k++)
    for (int i = 0; i < Mi; i++) ctab[k * M + i] = std::cos(M_PI * (double)k * ((double)i + 0.5) / (double)Mi);
  if (!lagrange) {
    // src/C/leg2cheb.c:621-648
    std::vector<double> Tm(2 * s * M);
    for (size_t i = 0; i < 2 * s; i++) {
      const double x = -1 + (double)i / (double)s, x2 = 2 * x;
      Tm[i * M] = 1;
      Tm[i * M + 1] = x;
      for (size_t j = 2; j < M; j++) Tm[i * M + j] = Tm[i * M + j - 1] * x2 - Tm[i * M + j - 2];
    }
    for (size_t i = 0; i < s; i++)
      for (size_t j = 0; j < M; j++) {
        T[i * M + j] = Tm[2 * i * M + j];
        T[(s + i) * M + j] = Tm[(2 * i + 1) * M + j];
      }
    binom_B0(Mi, B0.data());
    // U(0) = B(0); U(1) = B(0) with the odd diagonals negated; D(p) = U(p)^T
    for (int k = 0; k < Mi; k++)
      for (int l = 0; l < Mi; l++) {
        U[k * M + l] = B0[k * M + l];
        U[MM + k * M + l] = ((k + l) & 1) ? -B0[k * M + l] : B0[k * M + l];
      }
  } else {
    // src/C/leg2cheb.c:997-1031, :1081-1105: barycentric Lagrange at
    // Chebyshev-Gauss nodes
    std::vector<double> wj(M), Bn(2 * MM);
    for (int j = 0; j < Mi; j++) wj[j] = ((j % 2) ? -1.0 : 1.0) * std::sin((j + 0.5) * M_PI / Mi);
    for (size_t i = 0; i < s; i++) {
      const double xe = -1 + (double)(2 * i) / (double)s, xo = -1 + (double)(2 * i + 1) / (double)s;
      double se = 0, so = 0;
      for (size_t j = 0; j < M; j++) {
        const double e = wj[j] / (xe - xj[j]), o = wj[j] / (xo - xj[j]);
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
    for (int i = 0; i < Mi; i++) {
      const double a0 = (xj[i] - 1) / 2, a1 = (xj[i] + 1) / 2;
      double s0 = 0, s1 = 0;
      for (int j = 0; j < Mi; j++) {
        const double e0 = wj[j] / (a0 - xj[j]), e1 = wj[j] / (a1 - xj[j]);
        Bn[i * M + j] = e0;
        Bn[MM + i * M + j] = e1;
        s0 += e0;
        s1 += e1;
      }
      for (int j = 0; j < Mi; j++) {
        Bn[i * M + j] /= s0;
        Bn[MM + i * M + j] /= s1;
      }
    }
    // upward: parent = Bn0^T child0 + Bn1^T child1 ; downward: child_p = Bn_p parent
    for (int q = 0; q < 2; q++)
      for (int i = 0; i < Mi; i++)
        for (int j = 0; j < Mi; j++) U[q * MM + i * M + j] = Bn[q * MM + j * M + i];
  }
  for (int q = 0; q < 2; q++)
    for (int i = 0; i < Mi; i++)
      for (int j = 0; j < Mi; j++)
        D[q * MM + i * M + j] = U[q * MM + j * M + i];  // D(p) = U(p)^T
  // composite operators P(k, d), k = 1..lu, d < 2^k:
  // P(1,q) = U(q), P(k,d) = U(d >> (k-1)) P(k-1, d mod 2^(k-1))
  const int lu = p->lu_tab;
  const size_t nP = lu > 0 ? (((size_t)2 << lu) - 2) : 0;
  std::vector<double> P(std::max<size_t>(nP, 1) * MM);
  for (int k = 1; k <= lu; k++) {
    const size_t base = ((size_t)1 << k) - 2;
    for (size_t d = 0; d < ((size_t)1 << k); d++) {
      double *dst = &P[(base + d) * MM];
      const int top = (int)(d >> (k - 1));
      if (k == 1) {
        std::copy(&U[top * MM], &U[top * MM] + MM, dst);
      } else {
        const size_t pb = ((size_t)1 << (k - 1)) - 2;
        matmul(Mi, &U[top * MM], &P[(pb + (d & (((size_t)1 << (k - 1)) - 1))) * MM], dst);
      }
    }
  }
  // composite downward operators Q(k, d), k = 1..ld, d < 2^k: the L2L chain
  // from a level-l chunk down k levels to its descendant whose low k index
  // bits are d.  Q(1,q) = D(q), Q(k,d) = D(d & 1) Q(k-1, d >> 1).
  const int ldq = p->ld_tab;
  const size_t nQ = ldq > 0 ? (((size_t)2 << ldq) - 2) : 0;
  std::vector<double> Q(std::max<size_t>(nQ, 1) * MM);
  for (int k = 1; k <= ldq; k++) {
    const size_t base = ((size_t)1 << k) - 2;
    for (size_t d = 0; d < ((size_t)1 << k); d++) {
      double *dst = &Q[(base + d) * MM];
      if (k == 1) {
        std::copy(&D[d * MM], &D[d * MM] + MM, dst);
      } else {
        const size_t pb = ((size_t)1 << (k - 1)) - 2;
        matmul(Mi, &D[(d & 1) * MM], &Q[(pb + (d >> 1)) * MM], dst);
      }
    }
  }

  if (dmalloc(&p->t.T, 2 * s * M, acct) || dmalloc(&p->t.U, 2 * MM, acct) ||
      dmalloc(&p->t.D, 2 * MM, acct) || dmalloc(&p->t.P, std::max<size_t>(nP, 1) * MM, acct) ||
      dmalloc(&p->t.Q, std::max<size_t>(nQ, 1) * MM, acct)) {
    destroy_fmm(p);
    return nullptr;
  }
  // the device copy of every Q matrix is column-major: the chain threads of
  // one level read rows k = 0..M-1 of it, which then lie in consecutive words
  std::vector<double> Qc(Q.size());
  for (size_t m = 0; m < nQ; m++)
    for (int k = 0; k < Mi; k++)
      for (int l = 0; l < Mi; l++) Qc[m * MM + (size_t)l * M + k] = Q[m * MM + (size_t)k * M + l];
  if (nQ)
    L2C_HIP(hipMemcpyAsync(p->t.Q, Qc.data(), nQ * MM * 8, hipMemcpyHostToDevice, st), (destroy_fmm(p), nullptr));
  L2C_HIP(hipMemcpyAsync(p->t.T, T.data(), T.size() * 8, hipMemcpyHostToDevice, st), (destroy_fmm(p), nullptr));
  L2C_HIP(hipMemcpyAsync(p->t.U, U.data(), U.size() * 8, hipMemcpyHostToDevice, st), (destroy_fmm(p), nullptr));
  L2C_HIP(hipMemcpyAsync(p->t.D, D.data(), D.size() * 8, hipMemcpyHostToDevice, st), (destroy_fmm(p), nullptr));
  if (nP)
    L2C_HIP(hipMemcpyAsync(p->t.P, P.data(), nP * MM * 8, hipMemcpyHostToDevice, st), (destroy_fmm(p), nullptr));

  // M2L matrices
  const size_t Nc = p->Nc;
  if ((l2c && dmalloc(&p->t.A[0], Nc * MM, acct)) || (c2l && dmalloc(&p->t.A[1], Nc * MM, acct))) {
    destroy_fmm(p);
    return nullptr;
  }
  double *d_xj = nullptr, *d_ctab = nullptr, *d_lm = nullptr;
  if (dmalloc(&d_xj, M, nullptr) || dmalloc(&d_ctab, MM, nullptr) ||
      dmalloc(&d_lm, (size_t)L * 2 * MM, nullptr)) {
    destroy_fmm(p);
    return nullptr;
  }
  L2C_HIP(hipMemcpyAsync(d_xj, xj.data(), M * 8, hipMemcpyHostToDevice, st), (destroy_fmm(p), nullptr));
  L2C_HIP(hipMemcpyAsync(d_ctab, ctab.data(), MM * 8, hipMemcpyHostToDevice, st), (destroy_fmm(p), nullptr));
  k_lambda_minus<<<L * 2, 256, 0, st>>>(d_lm, d_xj, L, (int)s, Mi);
  L2C_HIP(hipGetLastError(), (destroy_fmm(p), nullptr));
  const size_t shm = 3 * MM * sizeof(double);
  k_build_A<<<(unsigned)Nc, 256, shm, st>>>(p->t.A[0], p->t.A[1], d_lm, d_xj, d_ctab, L, (int)s,
                                            Mi, (int)lagrange);
  L2C_HIP(hipGetLastError(), (destroy_fmm(p), nullptr));
  if (p->m2l_t && M <= 21) {
    // the row-across-lanes copy for k_m2l_t (Nc / 3 blocks of 3 M^2)
    for (int d = 0; d < 2; d++) {
      if (!p->t.A[d]) continue;
      if (dmalloc(&p->t.At[d], Nc * MM, acct)) {
        // the copy only speeds up the M2L: without room for it k_m2l reads A
        (void)hipGetLastError();
        clear_error();
        p->t.At[d] = nullptr;
        continue;
      }
      k_rows_across_lanes<<<(unsigned)(Nc / 3), 256, 0, st>>>(p->t.A[d], p->t.At[d], Mi);
      L2C_HIP(hipGetLastError(), (destroy_fmm(p), nullptr));
    }
  }

  // multipole store and staging
  const size_t slots = slot_base(L);
  (void)slots;
  p->wl_cap = std::max(ws_doubles(g), ws_doubles(p->geo1));
  // (host staging buffers are allocated by the first call with host arrays)
  if (dmalloc(&p->d_wl, p->wl_cap, acct)) {
    destroy_fmm(p);
    return nullptr;
  }
  L2C_HIP(hipStreamSynchronize(st), (destroy_fmm(p), nullptr));
  dfree(d_xj);
  dfree(d_ctab);
  dfree(d_lm);
  return p;
}

}  // namespace l2c

using namespace l2c;

// ---------------------------------------------------------------------------
// C ABI: plan lifetime, helpers

extern "C" {

const char *l2c_last_error(void) { return l2c::g_err.c_str(); }

fmm_plan *create_fmm(const size_t N, const size_t maxs, const size_t M, const size_t direction,
                     const size_t lagrange, const size_t precompute, const size_t v) {
  return create_fmm_impl(N, maxs, M, direction, lagrange, precompute, v);
}

void free_fmm(fmm_plan *plan) { destroy_fmm(plan); }

direct_plan *create_direct(size_t N, size_t direction, const size_t precompute) {
  clear_error();
  if (N == 0 || direction > 2) {
    set_error("create_direct: need N > 0 and direction in {0,1,2}");
    return nullptr;
  }
  return make_direct(N, direction, precompute);
}

void free_direct(direct_plan *dplan) { destroy_direct(dplan); }

int l2c_get_plan_info(const fmm_plan *p, l2c_plan_info *info) {
  if (!p || !info) {
    set_error("l2c_get_plan_info: NULL argument");
    return -1;
  }
  info->N = p->N;
  info->Nn = p->Nn;
  info->L = p->L;
  info->s = p->s;
  info->M = p->M;
  info->maxs = p->maxs;
  info->direction = p->direction;
  info->lagrange = p->lagrange;
  info->Nc = p->Nc;
  info->Nb = p->Nb;
  info->hbm_bytes = p->hbm_bytes;
  info->a_bytes = p->Nc * p->M * p->M * sizeof(double);
  info->subtree_up = p->L ? 1 << p->geo1.gu : 0;
  info->subtree_down = p->L ? 1 << p->geo1.gd : 0;
  return 0;
}

size_t l2c_get_plan_table(const fmm_plan *p, int table, size_t direction, double *dst,
                          size_t count) {
  clear_error();
  if (!p || direction > 1) {
    set_error("l2c_get_plan_table: NULL plan or direction not 0/1");
    return (size_t)-1;
  }
  const size_t N = p->N, M = p->M, MM = M * M, s = p->s;
  const double *src = nullptr;
  size_t n = 0;
  const bool fmm = p->L >= 1;
  const direct_plan_s *dq = p->dplan;
  switch (table) {
    case L2C_TAB_LAMBDA:
      src = fmm ? p->t.a : dq->d_a;
      n = N;
      break;
    case L2C_TAB_AN:
      src = fmm ? p->t.an : dq->d_an;
      n = N;
      break;
    case L2C_TAB_DN:
      src = fmm ? p->t.dn : dq->d_dn;
      n = (N + 1) / 2;
      break;
    case L2C_TAB_HEAD:
      src = fmm ? p->t.head[direction] : nullptr;
      n = 2 * s + 16;
      break;
    case L2C_TAB_M2L:
      src = fmm ? p->t.A[direction] : nullptr;
      n = p->Nc * MM;
      break;
    case L2C_TAB_T:
      src = fmm ? p->t.T : nullptr;
      n = 2 * s * M;
      break;
    case L2C_TAB_U:
      src = fmm ? p->t.U : nullptr;
      n = 2 * MM;
      break;
    case L2C_TAB_D:
      src = fmm ? p->t.D : nullptr;
      n = 2 * MM;
      break;
    case L2C_TAB_P:
      src = fmm && p->lu_tab > 0 ? p->t.P : nullptr;
      n = fmm && p->lu_tab > 0 ? (((size_t)2 << p->lu_tab) - 2) * MM : 0;
      break;
    case L2C_TAB_Q:
      src = fmm && p->ld_tab > 0 ? p->t.Q : nullptr;
      n = fmm && p->ld_tab > 0 ? (((size_t)2 << p->ld_tab) - 2) * MM : 0;
      break;
    default:
      set_error("l2c_get_plan_table: unknown table");
      return (size_t)-1;
  }
  if (!src) {
    set_error("l2c_get_plan_table: the plan holds no such table (direction not planned, "
              "direct-path plan or no composite levels)");
    return (size_t)-1;
  }
  if (!dst) return n;
  if (count < n) {
    set_error("l2c_get_plan_table: destination holds fewer doubles than the table");
    return (size_t)-1;
  }
  hipStream_t st = fmm ? p->stream : dq->stream;
  L2C_HIP(hipStreamSynchronize(st), (size_t)-1);
  L2C_HIP(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost), (size_t)-1);
  return n;
}

// src/C/leg2cheb.c:125-150
size_t get_number_of_blocks(const size_t level) { return nblocks(level); }
size_t get_h(const size_t level, const size_t L) { return (size_t)1 << (L - level - 1); }
size_t get_number_of_submatrices(const size_t level) { return 3 * nblocks(level); }
size_t get_total_number_of_submatrices(const size_t L) { return total_sub(L); }
size_t get_total_number_of_blocks(const size_t L) { return total_blocks(L); }
void get_ij(size_t *ij, const size_t level, const size_t block, const size_t s, const size_t L) {
  const size_t h = get_h(level, L);
  ij[0] = 2 * block * s * h;
  ij[1] = ij[0] + 2 * s * h;
}

// src/C/lambda.c
double _Lambda0(const double z) { return std::exp(std::lgamma(z + 0.5) - std::lgamma(z + 1)); }
double Lambda(const double z) { return lambda(z); }
double LambdaI(const size_t z) { return z > 255 ? lambda((double)z) : lamint()[z]; }
void Lambdavec(const double *z, double *w, const size_t N) {
  for (size_t i = 0; i < N; i++) w[i] = lambda(z[i]);
}

// src/C/leg2cheb.c:537-619: compact triangular B(0) (upper == false: lower
// triangle, rows packed; x holds the two children back to back) or B(0)^T
// (upper == true: upper triangle, rows packed; b holds the two children and
// is accumulated into).
void matvectri(const double *A, const double *x, double *b, const size_t m, const bool upper) {
  if (!upper) {
    const double *xp = x + m;
    size_t off = 0;
    for (size_t i = 0; i < m; i++) {
      double acc = 0;
      for (size_t j = 0; j <= i; j++) {
        const double z = ((i + j) & 1) ? x[j] - xp[j] : x[j] + xp[j];
        acc += A[off + j] * z;
      }
      b[i] = acc;
      off += i + 1;
    }
  } else {
    double *bp = b + m;
    size_t off = 0;
    for (size_t i = 0; i < m; i++) {
      double se = 0, so = 0;
      for (size_t j = i; j < m; j++) {
        if (((j - i) & 1) == 0)
          se += A[off + (j - i)] * x[j];
        else
          so += A[off + (j - i)] * x[j];
      }
      b[i] += se + so;
      bp[i] += se - so;
      off += m - i;
    }
  }
}

// src/C/dct.c:582-598 for M = 18, as a direct sum.
void dct2(double *input, double *output) {
  const int M = 18;
  double C[18 * 18], tmp[18 * 18];
  for (int k = 0; k < M; k++)
    for (int i = 0; i < M; i++) C[k * M + i] = std::cos(M_PI * (double)k * ((double)i + 0.5) / M);
  for (int i = 0; i < M; i++)
    for (int l = 0; l < M; l++) {
      double acc = 0;
      for (int j = 0; j < M; j++) acc += input[i * M + j] * C[l * M + j];
      tmp[i * M + l] = acc;
    }
  for (int k = 0; k < M; k++)
    for (int l = 0; l < M; l++) {
      double acc = 0;
      for (int i = 0; i < M; i++) acc += C[k * M + i] * tmp[i * M + l];
      output[k * M + l] = acc / 81.0;
    }
  for (int i = 0; i < M; i++) {
    output[i] *= 0.5;
    output[i * M] *= 0.5;
  }
}

void *l2c_malloc_device(size_t bytes) {
  void *p = nullptr;
  L2C_HIP(hipMalloc(&p, bytes), nullptr);
  return p;
}
int l2c_free_device(void *p) {
  L2C_HIP(hipFree(p), -1);
  return 0;
}
int l2c_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  L2C_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), -1);
  return 0;
}
int l2c_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  L2C_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), -1);
  return 0;
}
int l2c_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  L2C_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), -1);
  L2C_HIP(hipStreamSynchronize(nullptr), -1);  // complete on return
  return 0;
}
int l2c_synchronize(fmm_plan *p) {
  if (!p) return -1;
  L2C_HIP(hipStreamSynchronize(p->stream), -1);
  return 0;
}
int l2c_set_precision(fmm_plan *p, int form) {
  clear_error();
  if (!p || form < 0 || form > 2) {
    set_error("l2c_set_precision: form must be 0 (plain FP64), 1 (split M2L matrices) or 2 "
              "(M2L as BF16 slices on MFMA)");
    return -1;
  }
  if (form == 2 && p->L >= 1 && p->M != 18) {
    set_error("l2c_set_precision: form 2 needs M = 18");
    return -1;
  }
  drop_graphs(p);  // their launches carry the previous form's arguments
  if (form == 1 && p->L >= 1) {
    L2C_HIP(hipStreamSynchronize(p->stream), -1);
    for (int d = 0; d < 2; d++)
      if (p->t.A[d] && !p->t.sp[d].rec && build_split(p, d)) return -1;
  }
  if (form == 2 && p->L >= 1) {
    L2C_HIP(hipStreamSynchronize(p->stream), -1);
    for (int d = 0; d < 2; d++)
      if (p->t.A[d] && build_slices(p, d)) return -1;
  }
  p->precision = form;
  return 0;
}

}  // extern "C"
