// l2c_fmm.hip -- FMM execute kernels for gfx950 (src/C/leg2cheb.c:1299-1500
// and the near field :236-535), plus the direct O(N^2) kernel (:152-234).
//
// One transform (or one line of a batch) is three launches, both parities in
// every work-group:
//
//   k_up    leaf moments (step 1) on FP64 MFMA and the upward pass (step 2)
//           inside a subtree of 2^gu <= 16 leaf chunks; writes the multipole
//           store for levels >= lu and, for each level l < lu, the subtree
//           root's contribution P(lu-l, .) w_root to its level-l ancestor.
//   k_m2l   M2L (step 3) of every block of every level, streaming the M2L
//           matrices through LDS with coalesced loads (each matrix read once
//           for both parities); blocks above lu first sum the k_up
//           contributions into their multipoles.  HBM-bound.
//   k_down  per subtree of 2^gd leaf chunks: L2L (step 4) down the ancestor
//           chain and the subtree, the near-field band (FP64 VALU), leaf
//           evaluation (step 5), final scaling, one coalesced store of the
//           output (which is overwritten, not accumulated).
//
// Workspace per line: [multipoles: slot_base(L) x 2M][contributions:
// lu x 2^(lu+2) x 2M][M2L results: slot_base(L) x 2M]; chunk slot
// slot_base(l) + r holds [parity][M].
//
// Kernels are templated on the rank M (MT = M compiled in so that the
// M-loops unroll; MT = 0 reads M at run time).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "l2c_internal.h"

namespace l2c {

constexpr double kSqrtPi = 1.77245385090551602729816748334;
constexpr double k2Pi = 0.63661977236758134307553505349;  // 2/pi

typedef double d4 __attribute__((ext_vector_type(4)));

struct WsLayout {
  size_t w, part, m2l, total;
};
__host__ __device__ inline WsLayout ws_layout(const Geo &g) {
  const size_t M2 = 2 * (size_t)g.M;
  WsLayout l;
  l.w = 0;
  l.part = slot_base(g.L) * M2;
  l.m2l = l.part + (size_t)g.lu * ((size_t)4 << g.lu) * M2;
  l.total = l.m2l + slot_base(g.L) * M2;
  return l;
}
size_t ws_doubles(const Geo &g) { return ws_layout(g).total; }

// skewed LDS indices: one pad double every 4 (input) or 8 (Lambda window)
// so that lanes 4 (input) or 8 (Lambda) doubles apart hit distinct banks
__host__ __device__ __forceinline__ int sk4(int c) { return c + (c >> 2); }
__host__ __device__ __forceinline__ int sk8(int k) { return k + (k >> 3); }


// ---------------------------------------------------------------------------
// k_up: grid (2^(lu+2), lines), 256 threads.  G = 2^gu <= 16 leaf chunks.
//
// Leaf moments W_par (G x M) = X_par (G x s) T_par (s x M) run on
// v_mfma_f64_16x16x4f64 (A operand: lane l holds A[l%16][l/16]; B: lane l
// holds B[l/16][l%16]; D: lane l, register v holds D[l/16 + 4v][l%16]).
// One wave per (parity, 16-column tile of M).
constexpr int kUpNT = 256;

__host__ __device__ inline int up_xstride(int s) { return s | 1; }  // odd: conflict-free rows
// LDS: input [2][16][xst], moments [16][2][M], parents [8+1][2][M], the
// upward operators U(0), U(1) with rows padded to M+1 (the composite
// operators P are read from L2 by the threads that apply them)
size_t up_shared_bytes(const Geo &g) {
  const size_t M2 = 2 * g.M, Mp = g.M + 1;
  return 8 * (2 * 16 * (size_t)up_xstride(g.s) + 16 * M2 + 9 * M2 + 2 * (size_t)g.M * Mp);
}

// LM: the subtree levels on MFMA (MT > 0; chosen for subtrees of 2^3 and
// 2^4 leaf chunks in calls of up to 8 lines) or on VALU.
template <int MT, bool LM>
__global__ void __launch_bounds__(kUpNT) __attribute__((amdgpu_waves_per_eu(4))) k_up(
    const double *__restrict__ in, LineMap lm, Geo g,
                                              int dir, const double *__restrict__ T,
                                              const double *__restrict__ U,
                                              const double *__restrict__ P,
                                              double *__restrict__ wsall, size_t wsstride) {
  extern __shared__ double sm[];
  const int M = MT > 0 ? MT : g.M;
  const int s = g.s, L = g.L, M2 = 2 * M, MM = M * M;
  const int G = 1 << g.gu;
  const int j = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double *src = in + line_off(lm, blockIdx.y);
  const size_t Nl = (size_t)line_len(lm, blockIdx.y, g.N);  // zeros past the line's end
  double *ws = wsall + blockIdx.y * wsstride;
  const WsLayout lay = ws_layout(g);
  double *w = ws + lay.w;
  double *part = ws + lay.part;
  const int xst = up_xstride(s);
  const int Mp = M + 1;
  double *xs = sm;                  // [2][16][xst]
  double *w0 = xs + 2 * 16 * xst;   // [16][2][M]
  double *w1 = w0 + 16 * M2;        // [8+1][2][M]
  double *Ul = w1 + 9 * M2;         // [2][M][M+1]: U(0), U(1)

  // ---- loads, all issued before any use: input, this thread's rows of the
  // upward operators (k = tid % M) and of its composite operator P
  const int nin = G * 2 * s;
  const size_t base = (size_t)j * G * 2 * s;
  double vin[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int e = q * kUpNT + tid;
    const size_t i = base + e;
    vin[q] = (e < nin && i < Nl) ? src[i * lm.es] : 0.0;
  }
  // B operand of this wave's first MFMA task (16 k-steps = the first 64 rows of T)
  const int ntile = (M + 15) / 16;
  constexpr int QT = 16;
  double tb[QT];
  {
    const int sg = wave / ntile, nt = wave % ntile;
    const int kb = nt * 16 + (lane & 15), ka = lane >> 4;
    const double *tcol = T + (size_t)sg * s * M + kb;
#pragma unroll
    for (int q = 0; q < QT; q++) {
      const int m = 4 * q + ka;
      tb[q] = (wave < 2 * ntile && m < s && kb < M) ? tcol[(size_t)m * M] : 0.0;
    }
  }
  const int ku = tid % M;
  const int np = g.lu * M2;
  // A operand of the upward levels on MFMA (LM): rows rt*16 + (lane %
  // 16) of [U(0) U(1)] (M x 2M), k-steps kk (columns 4 kk + lane / 16), for
  // waves rt < ntile; loaded from HBM with the input, used from registers
  constexpr int QA = LM ? (2 * MT + 3) / 4 : 1;
  double ua[QA];
  if constexpr (LM) {
    const int ra = wave * 16 + (lane & 15);
#pragma unroll
    for (int kk = 0; kk < QA; kk++) {
      const int k = 4 * kk + (lane >> 4);
      const int pu = k < M ? 0 : 1, kc = k - pu * M;
      ua[kk] = (wave < ntile && ra < M && k < M2) ? U[(size_t)pu * MM + (size_t)ra * M + kc] : 0.0;
    }
  }
  // U(0), U(1) in LDS for the VALU levels (M without a compiled kernel),
  // once per work-group; every load of the batch is in flight before the
  // first LDS store
  if constexpr (!LM) {
    const int nop = 2 * MM;
    constexpr int QO = 3;
    double vo[QO];
#pragma unroll
    for (int q = 0; q < QO; q++) {
      const int e = q * kUpNT + tid;
      vo[q] = e < nop ? U[e] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < QO; q++) {
      const int e = q * kUpNT + tid;
      if (e < nop) Ul[(e / M) * Mp + e % M] = vo[q];
    }
    for (int e = QO * kUpNT + tid; e < nop; e += kUpNT) Ul[(e / M) * Mp + e % M] = U[e];
  }
  for (int e0 = 8 * kUpNT; e0 < nin; e0 += kUpNT) {  // rare: G*2s > 2048
    const int e = e0 + tid;
    if (e < nin) {
      const size_t i = base + e;
      const double v = i < Nl ? src[i * lm.es] : 0.0;
      xs[(e & 1) * 16 * xst + (e >> 1) / s * xst + (e >> 1) % s] =
          (dir == C2L && i > 0) ? v * (double)i : v;
    }
  }
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int e = q * kUpNT + tid;
    if (e < nin) {
      const size_t i = base + e;
      const int h = e >> 1;  // half index inside the subtree
      // v_i = i u_i for C2L (src/C/leg2cheb.c:1325-1335)
      xs[(e & 1) * 16 * xst + (h / s) * xst + (h % s)] =
          (dir == C2L && i > 0) ? vin[q] * (double)i : vin[q];
    }
  }
  // rows G..15 of the MFMA A operand are zero
  for (int e = tid; e < 2 * (16 - G) * xst; e += kUpNT) {
    const int par = e / ((16 - G) * xst), rem = e % ((16 - G) * xst);
    xs[par * 16 * xst + G * xst + rem] = 0.0;
  }
  __syncthreads();

  // ---- step 1 on MFMA: W_par[i][k] = sum_m X_par[i][m] T_par[m][k]
  for (int task = wave; task < 2 * ntile; task += kUpNT / 64) {
    const int sg = task / ntile, nt = task % ntile;
    const int ia = lane & 15, ka = lane >> 4;  // A: row i, k-offset
    const int kb = nt * 16 + (lane & 15);      // B: column k
    const double *xrow = xs + sg * 16 * xst + ia * xst;
    const double *tcol = T + (size_t)sg * s * M + kb;
    d4 acc = {0, 0, 0, 0};
    int m0 = 0;
    if (task == wave) {  // B prefetched for rows < 64; two independent chains
      d4 acc1 = {0, 0, 0, 0};
#pragma unroll
      for (int q = 0; q < QT; q += 2) {
        const int m = 4 * q + ka;
        const double a0 = m < s ? xrow[m] : 0.0, a1 = m + 4 < s ? xrow[m + 4] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, tb[q], acc, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, tb[q + 1], acc1, 0, 0, 0);
      }
      acc += acc1;
      m0 = 4 * QT;
    }
    for (; m0 < s; m0 += 4) {
      const int m = m0 + ka;
      const double a = m < s ? xrow[m] : 0.0;
      const double b = (m < s && kb < M) ? tcol[(size_t)m * M] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int i = (lane >> 4) + 4 * v;
      if (i < G && kb < M) {
        w0[(i * 2 + sg) * M + kb] = acc[v];
        w[(slot_base(L - 1) + (size_t)j * G + i) * M2 + sg * M + kb] = acc[v];
      }
    }
  }
  __syncthreads();

  // ---- step 2: parent = U0 child0 + U1 child1 inside the subtree.  LM:
  // on v_mfma_f64_16x16x4f64, parents (M x 2n) = [U0 U1] (M x 2M) times the
  // children stacked (2M x 2n), column jc = 2c + parity for parent c (2n <=
  // 16); wave rt < ntile computes row tile rt in two chains of k-steps.
  // Else on VALU, thread (k = tid % M) reading its operator rows from LDS.
  const int npair = kUpNT / M;  // (chunk, parity) pairs handled per sweep
  int n = G;
  if constexpr (LM) {
    const int jc = lane & 15, cc = jc >> 1, sgc = jc & 1;
    for (int lev = L - 2; lev >= g.lu; lev--) {
      n >>= 1;
      if (wave < ntile) {
        d4 D0 = {0, 0, 0, 0}, D1 = {0, 0, 0, 0};
        double bv[QA];
#pragma unroll
        for (int kk = 0; kk < QA; kk++) {
          const int k = 4 * kk + (lane >> 4);
          const int pu = k < M ? 0 : 1, kc = k - pu * M;
          bv[kk] = (jc < 2 * n && k < M2) ? w0[((2 * cc + pu) * 2 + sgc) * M + kc] : 0.0;
        }
#pragma unroll
        for (int kk = 0; kk < QA; kk += 2) {
          D0 = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[kk], bv[kk], D0, 0, 0, 0);
          if (kk + 1 < QA) D1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[kk + 1], bv[kk + 1], D1, 0, 0, 0);
        }
        D0 += D1;
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int i = wave * 16 + (lane >> 4) + 4 * v;
          if (i < M && jc < 2 * n) {
            w1[(cc * 2 + sgc) * M + i] = D0[v];
            w[(slot_base(lev) + (size_t)j * n + cc) * M2 + sgc * M + i] = D0[v];
          }
        }
      }
      __syncthreads();
      double *tmp = w0;
      w0 = w1;
      w1 = tmp;
    }
  }
  for (int lev = L - 2; !LM && lev >= g.lu; lev--) {
    n >>= 1;
    if (tid < npair * M) {
      for (int cs = tid / M; cs < 2 * n; cs += npair) {
        const int c = cs >> 1, sg = cs & 1;
        const double *x0 = w0 + (2 * c * 2 + sg) * M;
        const double *x1 = w0 + ((2 * c + 1) * 2 + sg) * M;
        const double *u0 = Ul + ku * Mp, *u1 = Ul + (M + ku) * Mp;
        double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        if constexpr (MT > 0) {
#pragma unroll
          for (int l = 0; l < MT; l += 2) {
            a0 += u0[l] * x0[l];
            a1 += u1[l] * x1[l];
            if (l + 1 < MT) {
              a2 += u0[l + 1] * x0[l + 1];
              a3 += u1[l + 1] * x1[l + 1];
            }
          }
        } else {
          for (int l = 0; l < M; l++) {
            a0 += u0[l] * x0[l];
            a1 += u1[l] * x1[l];
          }
        }
        const double acc = (a0 + a2) + (a1 + a3);
        w1[(c * 2 + sg) * M + ku] = acc;
        w[(slot_base(lev) + (size_t)j * n + c) * M2 + sg * M + ku] = acc;
      }
    }
    __syncthreads();
    double *tmp = w0;
    w0 = w1;
    w1 = tmp;
  }

  // ---- contributions of the root (level lu, chunk j) to its ancestors:
  // part[l][j] = P(lu-l, j mod 2^(lu-l)) w_root
  for (int task = tid; task < np; task += kUpNT) {
    const int lev = task / M2, o = task % M2, sg = o / M, i = o % M;
    const double *x = w0 + sg * M;
    const int k = g.lu - lev;
    const size_t d = (size_t)j & (((size_t)1 << k) - 1);
    const double *prow = P + ((((size_t)1 << k) - 2) + d) * MM + (size_t)i * M;
    double acc = 0;
#pragma unroll
    for (int l = 0; l < (MT > 0 ? MT : 64); l++) {
      if (MT == 0 && l >= M) break;
      acc += prow[l] * x[l];
    }
    part[((size_t)lev * ((size_t)4 << g.lu) + j) * M2 + o] = acc;
  }
}

// ---------------------------------------------------------------------------
// k_m2l: M2L of every block (src/C/leg2cheb.c:1419-1435).  Block b of level l
// holds A(b,0), A(b,1), A(b,2); row chunk 2b gets A(b,0) w[2b+2] +
// A(b,1) w[2b+3], row chunk 2b+1 gets A(b,2) w[2b+3]; both parities at once.
// grid: total_blocks(lu) work-groups for the blocks above lu (one each, with
// the multipoles summed from the k_up contributions in a fixed order), then
// one work-group per four blocks below (one block per wave).  256 threads.
constexpr int kM2lNT = 256;
// per wave: three padded matrices, two multipoles, 3M row results
__host__ __device__ inline int m2l_wave_lds(int M) { return 3 * M * (M + 1) + 4 * M + 6 * M; }
// Blocks per work-group below lu: four (one per wave) when four staged blocks
// fit in 64 KiB of LDS (M <= 22), else one block per work-group.
int m2l_bpw(const Geo &g) {
  return 8 * (size_t)(4 * m2l_wave_lds(g.M) + kM2lNT) <= 64 * 1024 ? 4 : 1;
}
size_t m2l_shared_bytes(const Geo &g) {
  return 8 * (size_t)(m2l_bpw(g) * m2l_wave_lds(g.M) + kM2lNT);
}

// Order LDS traffic between lanes of one wave (the wave's DS operations run
// in order; the fences stop the compiler from moving them across).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Evaluate one staged block with the 64 lanes of a wave (WG = false) or the
// whole work-group (WG = true).  Al: [3][M][M+1] (padded rows),
// wl: [chunk 2b+2, 2b+3][parity][M], res: [3M][2] scratch.
template <int MT, bool WG>
__device__ __forceinline__ void m2l_block_compute(const double *Al, const double *wl,
                                                  double *res, int M, int lane,
                                                  double *out_even, double *out_odd) {
  constexpr int step = WG ? kM2lNT : 64;
  for (int row = lane; row < 3 * M; row += step) {
    const int sub = row / M, k = row % M;
    const double *arow = Al + row * (M + 1);
    const double *wv = wl + (sub == 0 ? 0 : 2 * M);
    double o0 = 0, o1 = 0;
#pragma unroll
    for (int l = 0; l < (MT > 0 ? MT : 64); l++) {
      if (MT == 0 && l >= M) break;
      const double a = arow[l];
      o0 += a * wv[l];
      o1 += a * wv[M + l];
    }
    (void)k;
    res[2 * row] = o0;
    res[2 * row + 1] = o1;
  }
  if (WG)
    __syncthreads();
  else
    wave_sync();
  for (int k = lane; k < M; k += step) {
    // row chunk 2b: (A(b,0) w[2b+2]) + (A(b,1) w[2b+3]); row chunk 2b+1: A(b,2) w[2b+3]
    out_even[k] = res[2 * k] + res[2 * (M + k)];
    out_even[M + k] = res[2 * k + 1] + res[2 * (M + k) + 1];
    out_odd[k] = res[2 * (2 * M + k)];
    out_odd[M + k] = res[2 * (2 * M + k) + 1];
  }
}

// Split-form M2L block (precision form 1): record slots in tier order (FP64,
// FP32, BF16), decoded to FP64 into the padded LDS block at pos[slot]; the
// dropped positions (slots past the three tiers) are written as zeros.
struct SplitView {
  const unsigned char *rec;
  const unsigned short *pos;
  int n64, n32, n16, rec_bytes;
};
template <int NTH>
__device__ __forceinline__ void split_unpack(const SplitView &sv, size_t block, int n3, double *Al,
                                             int t) {
  // Every slot reads the aligned 8-byte word holding its value (FP64: the
  // value; FP32: one half; BF16: one quarter), so the loads of a batch carry
  // no branch and all issue before the first LDS store.
  const unsigned char *rb = sv.rec + block * sv.rec_bytes;
  const int n2 = sv.n64 + sv.n32, nt = n2 + sv.n16;
  for (int i0 = 0; i0 < n3; i0 += 8 * NTH) {
    unsigned long long wv[8];
    int d[8], o[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int i = i0 + q * NTH + t;
      d[q] = i < n3 ? (int)sv.pos[i] : -1;
      o[q] = i < sv.n64 ? 8 * i : i < n2 ? 8 * sv.n64 + 4 * (i - sv.n64) : 8 * sv.n64 + 4 * sv.n32 + 2 * (i - n2);
      const int oc = i < nt ? o[q] : 0;  // dropped slots read the record's first word
      wv[q] = *(const unsigned long long *)(rb + (oc & ~7));
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int i = i0 + q * NTH + t;
      const int sh = 8 * (o[q] & 7);
      const double v64 = __longlong_as_double((long long)wv[q]);
      const double v32 = (double)__uint_as_float((unsigned)(wv[q] >> sh));
      const double v16 = (double)__uint_as_float(((unsigned)(wv[q] >> sh) & 0xFFFFu) << 16);
      const double v = i < sv.n64 ? v64 : i < n2 ? v32 : i < nt ? v16 : 0.0;
      if (d[q] >= 0) Al[d[q]] = v;
    }
  }
}

// One block per work-group (B < total_blocks(lu): multipoles summed from the
// k_up contributions; or any block when only one fits the LDS), its lines
// line0 .. line0 + nl - 1.  sm: m2l_wave_lds(M) + kM2lNT doubles.
template <int MT>
__device__ __forceinline__ void m2l_top(const Geo &g, const double *__restrict__ A,
                                        double *__restrict__ wsall, size_t wsstride,
                                        size_t line0, int nl, const SplitView &sv, double *sm,
                                        int B) {
  const int M = MT > 0 ? MT : g.M, M2 = 2 * M, MM = M * M, Mp = M + 1;
  const int tid = threadIdx.x;
  const WsLayout lay = ws_layout(g);
  const int WL = m2l_wave_lds(M);
  int lev = 0;
  while ((int)total_blocks(lev + 1) <= B) lev++;
  const int b = B - (int)total_blocks(lev);
  double *Al = sm, *wl = sm + 3 * M * Mp, *red = sm + WL;
  if (sv.rec) {
    split_unpack<kM2lNT>(sv, (size_t)B, 3 * MM, Al, tid);
  } else {
    const double *Ab = A + (size_t)(3 * total_blocks(lev) + 3 * b) * MM;
    double va[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int e = q * kM2lNT + tid;
      va[q] = e < 3 * MM ? Ab[e] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int e = q * kM2lNT + tid;
      if (e < 3 * MM) Al[(e / M) * Mp + e % M] = va[q];
    }
    for (int e = 8 * kM2lNT + tid; e < 3 * MM; e += kM2lNT) Al[(e / M) * Mp + e % M] = Ab[e];
  }
  const int nw = 2 * M2;  // multipole doubles of chunks 2b+2, 2b+3
  for (int li = 0; li < nl; li++) {
    double *ws = wsall + (line0 + li) * wsstride;
    const double *w = ws + lay.w;
    double *mo = ws + lay.m2l;
    if (lev >= g.lu) {
      const double *wb = w + (slot_base(lev) + 2 * (size_t)b + 2) * M2;
      for (int o = tid; o < nw; o += kM2lNT) wl[o] = wb[o];
    } else {
      const size_t nd = (size_t)1 << (g.lu - lev);
      const int ngrp = std::max(1, kM2lNT / nw);
      for (int o0 = 0; o0 < nw; o0 += kM2lNT) {
        const int o = o0 + tid % std::min(nw, kM2lNT), grp = tid / std::min(nw, kM2lNT);
        double acc = 0;
        if (grp < ngrp && o < nw) {
          const int ci = o / M2, oo = o % M2;
          const double *pp = ws + lay.part +
                             ((size_t)lev * ((size_t)4 << g.lu) + (2 * (size_t)b + 2 + ci) * nd) * M2 + oo;
          for (size_t d0 = grp; d0 < nd; d0 += 8 * (size_t)ngrp) {
            double v[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
              const size_t d = d0 + (size_t)q * ngrp;
              v[q] = d < nd ? pp[d * M2] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 8; q++) acc += v[q];
          }
        }
        red[tid] = acc;
        __syncthreads();
        if (tid < std::min(nw, kM2lNT) && o0 + tid < nw) {
          double sum = 0;
          for (int gi = 0; gi < ngrp; gi++) sum += red[gi * std::min(nw, kM2lNT) + tid];
          wl[o0 + tid] = sum;
        }
        __syncthreads();
      }
    }
    __syncthreads();
    m2l_block_compute<MT, true>(Al, wl, wl + nw, M, tid,
                                mo + (slot_base(lev) + 2 * (size_t)b) * M2,
                                mo + (slot_base(lev) + 2 * (size_t)b + 1) * M2);
    __syncthreads();  // wl and the result scratch are reused by the next line
  }
}

template <int MT>
__global__ void __launch_bounds__(kM2lNT) k_m2l(Geo g, const double *__restrict__ A,
                                                double *__restrict__ wsall, size_t wsstride,
                                                int bpw, size_t nlines, int lpb, SplitView sv) {
  extern __shared__ double sm[];
  const int M = MT > 0 ? MT : g.M, M2 = 2 * M, MM = M * M, Mp = M + 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const WsLayout lay = ws_layout(g);
  const int ntop = (int)total_blocks(g.lu);
  const int ntot = (int)total_blocks(g.L);
  const int WL = m2l_wave_lds(M);
  // lines line0 .. line0 + nl - 1 share every staged block (the plan's M2L
  // matrices are read once per work-group, not once per line)
  const size_t line0 = (size_t)blockIdx.y * lpb;
  const int nl = (int)std::min<size_t>((size_t)lpb, nlines - line0);

  if ((int)blockIdx.x < ntop || bpw == 1) {
    // ---- one block per work-group: every block above lu (multipoles summed
    // from the k_up contributions), and every block when bpw == 1
    if ((int)blockIdx.x < ntot) m2l_top<MT>(g, A, wsall, wsstride, line0, nl, sv, sm, blockIdx.x);
    return;
  }

  // ---- blocks at levels >= lu: one per wave
  const int B = ntop + ((int)blockIdx.x - ntop) * 4 + wave;
  if (B >= ntot) return;  // no work-group barrier below this point
  int lev = g.lu;
  while ((int)total_blocks(lev + 1) <= B) lev++;
  const int b = B - (int)total_blocks(lev);
  double *Al = sm + wave * WL, *wl = Al + 3 * M * Mp;
  const size_t wofs = lay.w + (slot_base(lev) + 2 * (size_t)b + 2) * M2;  // chunks 2b+2, 2b+3
  const size_t oofs = lay.m2l + (slot_base(lev) + 2 * (size_t)b) * M2;
  // multipoles of the first line; later lines are fetched one line ahead
  const double *wb = wsall + line0 * wsstride + wofs;
  double w0v = lane < 2 * M2 ? wb[lane] : 0.0;  // 2*M2 <= 192
  double w1v = lane + 64 < 2 * M2 ? wb[lane + 64] : 0.0;
  double w2v = lane + 128 < 2 * M2 ? wb[lane + 128] : 0.0;
  if (sv.rec) {
    split_unpack<64>(sv, (size_t)B, 3 * MM, Al, lane);
  } else {
    const double *Ab = A + (size_t)(3 * total_blocks(lev) + 3 * b) * MM;
    constexpr int QA = MT > 0 ? (3 * MT * MT + 63) / 64 : 16;
    double va[QA];
#pragma unroll
    for (int q = 0; q < QA; q++) {
      const int e = q * 64 + lane;
      va[q] = e < 3 * MM ? Ab[e] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < QA; q++) {
      const int e = q * 64 + lane;
      if (e < 3 * MM) Al[(e / M) * Mp + e % M] = va[q];
    }
    for (int e = QA * 64 + lane; e < 3 * MM; e += 64) Al[(e / M) * Mp + e % M] = Ab[e];
  }
  for (int li = 0; li < nl; li++) {
    if (lane < 2 * M2) wl[lane] = w0v;
    if (lane + 64 < 2 * M2) wl[lane + 64] = w1v;
    if (lane + 128 < 2 * M2) wl[lane + 128] = w2v;
    if (li + 1 < nl) {
      const double *wn = wsall + (line0 + li + 1) * wsstride + wofs;
      w0v = lane < 2 * M2 ? wn[lane] : 0.0;
      w1v = lane + 64 < 2 * M2 ? wn[lane + 64] : 0.0;
      w2v = lane + 128 < 2 * M2 ? wn[lane + 128] : 0.0;
    }
    wave_sync();
    double *mo = wsall + (line0 + li) * wsstride + oofs;
    m2l_block_compute<MT, false>(Al, wl, wl + 2 * M2, M, lane, mo, mo + M2);
    wave_sync();  // wl and the result scratch are reused by the next line
  }
}

// M2L of the blocks at levels >= lu, one block per wave, from At (rows across
// lanes): lane r < 3M holds row r of the block (r = sub M + row) in
// registers after M coalesced loads, so no block is staged in LDS and the
// work-group needs only the multipoles' 4M doubles per wave; the rows' sums
// are combined across lanes (chunk 2b: rows of sub 0 plus sub 1; chunk
// 2b+1: sub 2) in the order k_m2l uses, so the results are the same.  The
// first ntop work-groups take the blocks above lu as k_m2l does (one per
// work-group, from A).  Grid (ntop + (ntot - ntop + 3) / 4, line groups),
// kM2lNT threads, 8 (m2l_wave_lds(M) + kM2lNT) bytes of LDS.
template <int MT>
__global__ void __launch_bounds__(kM2lNT) k_m2l_t(Geo g, const double *__restrict__ A,
                                                  const double *__restrict__ At,
                                                  double *__restrict__ wsall, size_t wsstride,
                                                  size_t nlines, int lpb) {
  static_assert(MT > 0 && 3 * MT <= 64, "rows of a block across one wave");
  constexpr int M = MT, M2 = 2 * MT, MM = MT * MT;
  extern __shared__ double sm[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ntop = (int)total_blocks(g.lu), ntot = (int)total_blocks(g.L);
  const size_t line0 = (size_t)blockIdx.y * lpb;
  const int nl = (int)std::min<size_t>((size_t)lpb, nlines - line0);
  if ((int)blockIdx.x < ntop) {
    m2l_top<MT>(g, A, wsall, wsstride, line0, nl, SplitView{}, sm, blockIdx.x);
    return;
  }
  const int B = ntop + ((int)blockIdx.x - ntop) * 4 + wave;
  if (B >= ntot) return;  // no work-group barrier below this point
  int lev = g.lu;
  while ((int)total_blocks(lev + 1) <= B) lev++;
  const int b = B - (int)total_blocks(lev);
  const WsLayout lay = ws_layout(g);
  const size_t wofs = lay.w + (slot_base(lev) + 2 * (size_t)b + 2) * M2;  // chunks 2b+2, 2b+3
  const size_t oofs = lay.m2l + (slot_base(lev) + 2 * (size_t)b) * M2;
  const bool rowl = lane < 3 * M;
  const double *Ab = At + (size_t)B * 3 * MM;
  double a[M];
#pragma unroll
  for (int q = 0; q < M; q++) a[q] = rowl ? Ab[q * 3 * M + lane] : 0.0;
  double *wl = sm + wave * 2 * M2;
  const int sub = lane / M;
  const double *wv = wl + (sub == 0 ? 0 : M2);  // [chunk][parity][M]
  for (int li = 0; li < nl; li++) {
    const double *wb = wsall + (line0 + li) * wsstride + wofs;
    wl[lane] = wb[lane];
    if (lane + 64 < 2 * M2) wl[lane + 64] = wb[lane + 64];
    wave_sync();
    double o0 = 0, o1 = 0;
#pragma unroll
    for (int q = 0; q < M; q++) {
      o0 += a[q] * wv[q];
      o1 += a[q] * wv[M + q];
    }
    const double p0 = __shfl(o0, lane + M, 64), p1 = __shfl(o1, lane + M, 64);
    double *mo = wsall + (line0 + li) * wsstride + oofs;
    if (lane < M) {
      mo[lane] = o0 + p0;
      mo[M + lane] = o1 + p1;
    } else if (lane >= 2 * M && rowl) {
      mo[M2 + lane - 2 * M] = o0;
      mo[M2 + M + lane - 2 * M] = o1;
    }
    wave_sync();  // wl is rewritten for the next line
  }
}

// ---------------------------------------------------------------------------
// k_m2l_mma: M2L of the blocks at levels >= lu for many lines of one plan on
// MFMA, one block per wave, the lines in groups of 8 (16 MFMA columns:
// column n = line (n >> 1), parity n & 1).  A block's three submatrices are
// applied as one stacked operator E (36 x 36): rows 0..17 = [A(b,0) A(b,1)]
// give the M2L vector of chunk 2b, rows 18..35 = [0 A(b,2)] that of chunk
// 2b+1, both from the stacked multipoles S = [w(2b+2); w(2b+3)]
// (src/C/leg2cheb.c:1419-1435).
//   FORM 0: FP64, v_mfma_f64_16x16x4f64 (A lane l: E[l%16][l/16], B lane l:
//           S[l/16][l%16], D lane l reg v: [l/16 + 4v][l%16]); E's fragments
//           stay in registers across the wave's line groups.
//   FORM 2: BF16 slices, v_mfma_f32_16x16x32_bf16 (A lane l elem e:
//           E[l%16][8(l/16)+e], B lane l elem e: S[8(l/16)+e][l%16], D lane l
//           reg v: [4(l/16)+v][l%16]).  Rows of E carry a power-of-two scale
//           from plan time, columns of S one from their largest entry; both
//           are cut into kSlices slices q_i 2^(-8(i+1)) with integers
//           |q_i| <= 256, so every slice is exact in BF16, every product of
//           slices exact in FP32, and the sum over the 36 k of the pairs
//           (i, j) on one anti-diagonal i + j = d exact in one FP32
//           accumulator; anti-diagonals d < kSlices are kept and recombined
//           in FP64.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Cut x (|x| < 1) into kSlices integers q_i with x ~ sum q_i 2^(-8(i+1)).
__device__ __forceinline__ void cut_slices(double x, float q[kSlices]) {
#pragma unroll
  for (int i = 0; i < kSlices; i++) {
    const double sc = (double)(1ull << (8 * (i + 1)));
    const double qi = rint(x * sc);
    q[i] = (float)qi;
    x -= qi / sc;  // exact: qi / sc is x rounded to a coarser grid
  }
}

// Plan time (form 2): block B = ntop + blockIdx.x, one wave; lane l writes its
// fragment elements of every (row tile, k-step, slice) and, for l < 48, the
// scale of row l.
__global__ void __launch_bounds__(64) k_slice_A(const double *__restrict__ A, int M, int L, int lu,
                                                unsigned short *__restrict__ q,
                                                double *__restrict__ sc) {
  const int B = (int)total_blocks(lu) + blockIdx.x, l = threadIdx.x;
  int lev = lu;
  while ((int)total_blocks(lev + 1) <= B) lev++;
  const int b = B - (int)total_blocks(lev);
  const size_t MM = (size_t)M * M;
  const double *Ab = A + (3 * total_blocks(lev) + 3 * (size_t)b) * MM;
  auto E = [&](int r, int k) -> double {
    if (r < M) return k < M ? Ab[r * M + k] : (k < 2 * M ? Ab[MM + r * M + (k - M)] : 0.0);
    if (r < 2 * M && k >= M && k < 2 * M) return Ab[2 * MM + (r - M) * M + (k - M)];
    return 0.0;
  };
  auto row_scale = [&](int r) -> double {
    double m = 0.0;
    for (int k = 0; k < 2 * M; k++) m = fmax(m, fabs(E(r, k)));
    if (m == 0.0) return 0.0;
    int e;
    (void)frexp(m, &e);
    return ldexp(1.0, e);  // m < 2^e
  };
  const size_t blk = (size_t)blockIdx.x;
  if (l < 48) sc[blk * 48 + l] = row_scale(l);
  for (int t = 0; t < 3; t++) {
    const int r = 16 * t + (l & 15);
    const double s_r = row_scale(r);
    for (int kk = 0; kk < 2; kk++) {
      float qs[8][kSlices];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int k = 32 * kk + 8 * (l >> 4) + e;
        cut_slices(s_r > 0 ? E(r, k) / s_r : 0.0, qs[e]);
      }
      for (int i = 0; i < kSlices; i++) {
        unsigned short *dst = q + ((((blk * 3 + t) * 2 + kk) * kSlices + i) * 64 + l) * 8;
#pragma unroll
        for (int e = 0; e < 8; e++) dst[e] = (unsigned short)(__float_as_uint(qs[e][i]) >> 16);
      }
    }
  }
}

template <int FORM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_m2l_mma(Geo g, const double *__restrict__ A,
                                                 const bf16x8 *__restrict__ Aq,
                                                 const double *__restrict__ Asc,
                                                 double *__restrict__ wsall, size_t wsstride,
                                                 size_t nlines, int groups_per_wg) {
  extern __shared__ double sm[];
  constexpr int M = 18, M2 = 36, MM = 324, Mp = 19;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ntop = (int)total_blocks(g.lu), ntot = (int)total_blocks(g.L);
  const int B = ntop + (int)blockIdx.x * 4 + wave;
  if (B >= ntot) return;  // no work-group barrier in this kernel
  int lev = g.lu;
  while ((int)total_blocks(lev + 1) <= B) lev++;
  const int b = B - (int)total_blocks(lev);
  const WsLayout lay = ws_layout(g);
  const size_t sofs = lay.w + (slot_base(lev) + 2 * (size_t)b + 2) * M2;  // S: w(2b+2), w(2b+3)
  const size_t oofs = lay.m2l + (slot_base(lev) + 2 * (size_t)b) * M2;    // out: chunks 2b, 2b+1
  const size_t ngroups = (nlines + 7) / 8;
  const size_t gr0 = (size_t)blockIdx.y * groups_per_wg;
  const size_t gr1 = std::min(ngroups, gr0 + groups_per_wg);
  const int n = lane & 15;
  if (FORM == 0) {
    // the block into this wave's LDS rows, then E's 23 fragments into registers
    double *Al = sm + wave * 3 * M * Mp;
    const double *Ab = A + (size_t)(3 * total_blocks(lev) + 3 * b) * MM;
    constexpr int QA = (3 * MM + 63) / 64;
    double va[QA];
#pragma unroll
    for (int q = 0; q < QA; q++) {
      const int e = q * 64 + lane;
      va[q] = e < 3 * MM ? Ab[e] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < QA; q++) {
      const int e = q * 64 + lane;
      if (e < 3 * MM) Al[(e / M) * Mp + e % M] = va[q];
    }
    wave_sync();
    auto E = [&](int r, int k) -> double {
      if (r < M) return k < M ? Al[r * Mp + k] : Al[(M + r) * Mp + (k - M)];
      if (r < 2 * M && k >= M) return Al[(M + r) * Mp + (k - M)];  // A(b,2) rows
      return 0.0;
    };
    double ea[23];
#pragma unroll
    for (int j = 0; j < 9; j++) {
      const int k = 4 * j + (lane >> 4);
      ea[j] = E(n, k);
      ea[9 + j] = E(16 + n, k);
      if (j >= 4) ea[18 + j - 4] = E(32 + n, k);
    }
    for (size_t gr = gr0; gr < gr1; gr++) {
      const size_t line = gr * 8 + (n >> 1);
      const bool live = line < nlines;
      const double *S = wsall + (live ? line : 0) * wsstride + sofs + (n & 1) * M;
      double bv[9];
#pragma unroll
      for (int j = 0; j < 9; j++) {
        const int k = 4 * j + (lane >> 4);
        bv[j] = live ? S[(k >= M ? M2 : 0) + (k >= M ? k - M : k)] : 0.0;
      }
      d4 D0 = {0, 0, 0, 0}, D1 = D0, D2 = D0;
#pragma unroll
      for (int j = 0; j < 9; j++) {
        D0 = __builtin_amdgcn_mfma_f64_16x16x4f64(ea[j], bv[j], D0, 0, 0, 0);
        D1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ea[9 + j], bv[j], D1, 0, 0, 0);
        if (j >= 4) D2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ea[18 + j - 4], bv[j], D2, 0, 0, 0);
      }
      if (live) {
        double *o = wsall + line * wsstride + oofs + (n & 1) * M;
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r0 = (lane >> 4) + 4 * v;  // rows of tiles 0, 1, 2: r0, 16 + r0, 32 + r0
          o[r0] = D0[v];                                          // r0 < 16: chunk 2b
          o[16 + r0 < M ? 16 + r0 : M2 + 16 + r0 - M] = D1[v];     // 16..17: 2b; 18..31: 2b+1
          if (32 + r0 < 2 * M) o[M2 + 32 + r0 - M] = D2[v];         // 32..35: 2b+1
        }
      }
    }
  } else {
    const size_t blk = (size_t)(B - ntop);
    const bf16x8 *Ablk = Aq + blk * 3 * 2 * kSlices * 64;
    for (size_t gr = gr0; gr < gr1; gr++) {
      const size_t line = gr * 8 + (n >> 1);
      const bool live = line < nlines;
      const double *S = wsall + (live ? line : 0) * wsstride + sofs + (n & 1) * M;
      // this lane's entries of S: k = 8 (l/16) + e (k-step 0), 32 + 8 (l/16) + e (k-step 1)
      double sv[2][8];
      double mx = 0.0;
#pragma unroll
      for (int kk = 0; kk < 2; kk++)
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const int k = 32 * kk + 8 * (lane >> 4) + e;
          sv[kk][e] = (live && k < M2) ? S[(k >= M ? M2 : 0) + (k >= M ? k - M : k)] : 0.0;
          mx = fmax(mx, fabs(sv[kk][e]));
        }
      mx = fmax(mx, __shfl_xor(mx, 16));
      mx = fmax(mx, __shfl_xor(mx, 32));
      double s_c = 0.0;
      if (mx > 0) {
        int ex;
        (void)frexp(mx, &ex);
        s_c = ldexp(1.0, ex);
      }
      // the column's slices as B fragments
      bf16x8 bq[2][kSlices];
#pragma unroll
      for (int kk = 0; kk < 2; kk++)
#pragma unroll
        for (int e = 0; e < 8; e++) {
          float qs[kSlices];
          cut_slices(s_c > 0 ? sv[kk][e] / s_c : 0.0, qs);
#pragma unroll
          for (int i = 0; i < kSlices; i++) bq[kk][i][e] = (__bf16)qs[i];
        }
      // one row tile at a time: anti-diagonal accumulators, A's slices
      // fetched per slice (L2-resident), the column's slices kept
      double *o = wsall + line * wsstride + oofs + (n & 1) * M;
      const double *scb = Asc + blk * 48;
#pragma unroll 1
      for (int t = 0; t < 3; t++) {
        f4 acc[kSlices];
#pragma unroll
        for (int d = 0; d < kSlices; d++) acc[d] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < kSlices; i++) {
          const bf16x8 a0 = Ablk[((t * 2 + 0) * kSlices + i) * 64 + lane];
          const bf16x8 a1 = Ablk[((t * 2 + 1) * kSlices + i) * 64 + lane];
#pragma unroll
          for (int j = 0; j + i < kSlices; j++) {
            acc[i + j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, bq[0][j], acc[i + j], 0, 0, 0);
            acc[i + j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bq[1][j], acc[i + j], 0, 0, 0);
          }
        }
        if (live) {
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int r = 16 * t + 4 * (lane >> 4) + v;
            if (r >= M2) continue;
            double val = 0.0;
#pragma unroll
            for (int d = kSlices - 1; d >= 0; d--) val += ldexp((double)acc[d][v], -8 * (d + 2));
            o[r < M ? r : M2 + r - M] = val * (scb[r] * s_c);
          }
        }
      }
    }
  }
}

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
// Software-pipelined: the LDS reads of trip b+1 (eight x values, eight head
// values, eight ring refills) are issued before the arithmetic of trip b, so
// LDS latency hides behind 64 FP64 operations.  Reads run at most one trip
// (eight columns) past len, which stays inside the zero-filled windows.
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
  double XA[8], HA[8], PA[8], XB[8], HB[8], PB[8];
#define L2C_TERM(am, ap, x, acc_)   \
  if (DIR == 0)                     \
    acc_ += ((am) * (ap)) * (x);    \
  else if (PRE)                     \
    acc_ -= ((am) * (ap)) * (x);    \
  else                              \
    acc_ -= ((am) * (x)) * (ap);
#define BAND_STEP(K, XV, HV, PV)                           \
  {                                                        \
    m[(K) & 3] = (HV);                                     \
    const double x_ = (XV);                                \
    L2C_TERM(m[(K) & 3], p[(K) & 3], x_, a[0])             \
    L2C_TERM(m[((K) + 3) & 3], p[((K) + 1) & 3], x_, a[1]) \
    L2C_TERM(m[((K) + 2) & 3], p[((K) + 2) & 3], x_, a[2]) \
    L2C_TERM(m[((K) + 1) & 3], p[((K) + 3) & 3], x_, a[3]) \
    p[(K) & 3] = (PV);                                     \
  }
  // trip t's data: x at xb + 10t, heads at hb + 8t, refills at pbp + 9t
#define LOAD_TRIP(X, H, P, T)                                      \
  _Pragma("unroll") for (int k_ = 0; k_ < 8; k_++) {               \
    X[k_] = xb[10 * (T) + OX[k_]];                                 \
    H[k_] = hb[8 * (T) + k_];                                      \
    P[k_] = pbp[9 * (T) + OP[k_ + 4]];                             \
  }
#define RUN_TRIP(X, H, P)                                                  \
  BAND_STEP(0, X[0], H[0], P[0]) BAND_STEP(1, X[1], H[1], P[1])            \
  BAND_STEP(2, X[2], H[2], P[2]) BAND_STEP(3, X[3], H[3], P[3])            \
  BAND_STEP(4, X[4], H[4], P[4]) BAND_STEP(5, X[5], H[5], P[5])            \
  BAND_STEP(6, X[6], H[6], P[6]) BAND_STEP(7, X[7], H[7], P[7])
  const int nb = len >> 3, rem = len & 7;
  LOAD_TRIP(XA, HA, PA, 0)
  int b = 0;
  for (; b + 2 <= nb; b += 2) {
    LOAD_TRIP(XB, HB, PB, 1)
    RUN_TRIP(XA, HA, PA)
    LOAD_TRIP(XA, HA, PA, 2)
    RUN_TRIP(XB, HB, PB)
    xb += 20;
    pbp += 18;
    hb += 16;
  }
  if (b < nb) {
    RUN_TRIP(XA, HA, PA)
    LOAD_TRIP(XA, HA, PA, 1)
  }
  if (rem) {
    BAND_STEP(0, XA[0], HA[0], PA[0])
    BAND_STEP(1, rem > 1 ? XA[1] : 0.0, HA[1], PA[1])
    BAND_STEP(2, rem > 2 ? XA[2] : 0.0, HA[2], PA[2])
    BAND_STEP(3, rem > 3 ? XA[3] : 0.0, HA[3], PA[3])
    BAND_STEP(4, rem > 4 ? XA[4] : 0.0, HA[4], PA[4])
    BAND_STEP(5, rem > 5 ? XA[5] : 0.0, HA[5], PA[5])
    BAND_STEP(6, rem > 6 ? XA[6] : 0.0, HA[6], PA[6])
  }
#undef RUN_TRIP
#undef LOAD_TRIP
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

// LDS plan of k_down (offsets in doubles).  The M2L vectors and the chain
// partial sums are dead once the band phase starts; the output staging
// reuses them.
struct DownLds {
  int xs, aw, cres, mres, qp, out, total;
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
  d.cres = d.aw + nas;
  d.mres = d.cres + (2 * G - 1) * M2;
  d.qp = d.mres + (g.ld + 2 * G - 1) * M2;
  d.out = d.mres;
  d.total = std::max(d.qp + std::max(g.ld, 1) * M2, d.out + 2 * G * s);
  return d;
}
size_t down8_shared_bytes(const Geo &g);
size_t down_shared_bytes(const Geo &g) {
  return g.fast ? down8_shared_bytes(g) : 8 * (size_t)down_lds(g).total;
}

// Subtree chunk slot cl (level ld + lam, position ci) of work-group Rr:
// slots are level-major, level ld + lam starting at 2^lam - 1.
__device__ __forceinline__ void sub_slot(int cl, int &lam, int &ci) {
  lam = 31 - __builtin_clz(cl + 1);
  ci = cl + 1 - (1 << lam);
}

// Address of M2L vector element e of the work-group's list [chain levels
// 0..ld-1 (ancestors of chunk Rr at level ld)][subtree slots], and whether the
// chunk has an M2L block (the last two chunks of every level have none).
__device__ __forceinline__ size_t m2l_elem(const Geo &g, int Rr, int e, int M2, bool &valid) {
  const int slot = e / M2, o = e - slot * M2;
  int lev;
  size_t r;
  if (slot < g.ld) {
    lev = slot;
    r = (size_t)Rr >> (g.ld - lev);
  } else {
    int lam, ci;
    sub_slot(slot - g.ld, lam, ci);
    lev = g.ld + lam;
    r = ((size_t)Rr << lam) + ci;
  }
  valid = r < ((size_t)4 << lev) - 2;
  return (slot_base(lev) + r) * M2 + o;
}

// Row k of the composite operator Q(ld - lev, Rr mod 2^(ld - lev)) for chain
// task (lev, k) = (task / M, task % M).
// Row k of the chain operator of task (lev, k).  The plan stores each Q
// matrix column-major, so the row's elements are M apart (qrow[l * M]) and
// the M threads of one level, rows k = 0..M-1, read consecutive words.
__device__ __forceinline__ const double *qrow_ptr(const double *Qg, const Geo &g, int Rr, int task,
                                                  int M) {
  const int lev = task / M, k = task - lev * M, kk = g.ld - lev;
  const size_t d = (size_t)Rr & (((size_t)1 << kk) - 1);
  return Qg + ((((size_t)1 << kk) - 2) + d) * (size_t)(M * M) + (size_t)k;
}

// Local expansions (L2L, src/C/leg2cheb.c:1437-1449) of the work-group's
// subtree into cres [slot][parity][M]: the root's c_ld = sum_lev Q(ld - lev,
// .) m2l_lev + m2l_ld replaces the chain c_l = D(r_l & 1) c_(l-1) + m2l_l
// (chunks without an M2L block, the last two of a level and all their
// ancestors, contribute zero exactly as in the chain); then levels
// ld+1..ld+gd: c = D(r & 1) c_parent + m2l.  mres: the M2L vectors in
// m2l_elem order; qp: ld x 2M scratch.  qr: row (tid / M, tid % M) of the
// chain operator when MT > 0 (else read from Qg).  DLDS: the subtree levels
// run on MFMA with D(p) from the LDS copy Dl ([p][k][M + 1], padded rows);
// else on VALU with D(p) row tid % M in d0r/d1r (MT > 0) or from Dg.  Every
// thread of the work-group must call this.
template <int MT, bool DLDS, class Sync>
__device__ __forceinline__ void subtree_locals(const Geo &g, int Rr, int tid, int NT,
                                               const double *mres, double *qp, double *cres,
                                               const double *qr, const double *d0r,
                                               const double *d1r, const double *Qg,
                                               const double *Dg, const double *Dl,
                                               Sync &&sync) {
  const int M = MT > 0 ? MT : g.M, M2 = 2 * M, MM = M * M;
  const int nq = g.ld * M;
  int task0 = tid;
  if (MT > 0 && tid < nq) {  // first task: the operator row is in registers
    const int lev = tid / M, k = tid - lev * M;
    const double *mv = mres + lev * M2;
    double o0 = 0, o1 = 0;
#pragma unroll
    for (int l = 0; l < (MT > 0 ? MT : 1); l++) {
      o0 += qr[l] * mv[l];
      o1 += qr[l] * mv[M + l];
    }
    qp[lev * M2 + k] = o0;
    qp[lev * M2 + M + k] = o1;
    task0 += NT;
  }
  for (int task = task0; task < nq; task += NT) {
    const int lev = task / M, k = task - lev * M;
    const double *mv = mres + lev * M2;
    const double *qrow = qrow_ptr(Qg, g, Rr, task, M);
    double o0 = 0, o1 = 0;
    for (int l = 0; l < M; l++) {
      o0 += qrow[l * M] * mv[l];
      o1 += qrow[l * M] * mv[M + l];
    }
    qp[lev * M2 + k] = o0;
    qp[lev * M2 + M + k] = o1;
  }
  sync();
  for (int o = tid; o < M2; o += NT) {
    // the chain terms in level order; four reads in flight per step
    double v = mres[g.ld * M2 + o];
    int lev = 0;
    for (; lev + 4 <= g.ld; lev += 4) {
      const double q0 = qp[lev * M2 + o], q1 = qp[(lev + 1) * M2 + o];
      const double q2 = qp[(lev + 2) * M2 + o], q3 = qp[(lev + 3) * M2 + o];
      v += q0;
      v += q1;
      v += q2;
      v += q3;
    }
    for (; lev < g.ld; lev++) v += qp[lev * M2 + o];
    cres[o] = v;
  }
  sync();
  if (DLDS) {
    // levels ld+1..ld+gd on v_mfma_f64_16x16x4f64: children of parity p are
    // C (M x n) = D(p) (M x M) C_parents (M x n), n = 2^lam children x 2
    // parities <= 16 columns: column j = child 2(j >> 1) + p, parity j & 1.
    // Work item (p, 16-row tile rt) per wave.
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nw = NT >> 6;
    const int nrt = (M + 15) >> 4, ksteps = (M + 3) >> 2;
    const int j = lane & 15;
    for (int lam = 1; lam <= g.gd; lam++) {
      const int n = 1 << lam, lev = g.ld + lam;
      const int b0 = n - 1, bp = (n >> 1) - 1;
      for (int item = wave; item < 2 * nrt; item += nw) {
        const int p = item / nrt, rt = item - p * nrt;
        const int ra = rt * 16 + j;  // A row
        d4 D = {0, 0, 0, 0};
        if (MT > 0) {  // all operand reads first, then the MFMA chain
          constexpr int QK = MT > 0 ? (MT + 3) / 4 : 1;
          double av[QK], bv[QK];
#pragma unroll
          for (int kk = 0; kk < QK; kk++) {
            const int k = kk * 4 + (lane >> 4);
            av[kk] = (ra < M && k < M) ? Dl[(p * M + ra) * (M + 1) + k] : 0.0;
            bv[kk] = (j < n && k < M) ? cres[(bp + (j >> 1)) * M2 + (j & 1) * M + k] : 0.0;
          }
#pragma unroll
          for (int kk = 0; kk < QK; kk++) D = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kk], bv[kk], D, 0, 0, 0);
        } else {
          for (int kk = 0; kk < ksteps; kk++) {
            const int k = kk * 4 + (lane >> 4);
            const double a = (ra < M && k < M) ? Dl[(p * M + ra) * (M + 1) + k] : 0.0;
            const double b = (j < n && k < M) ? cres[(bp + (j >> 1)) * M2 + (j & 1) * M + k] : 0.0;
            D = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, D, 0, 0, 0);
          }
        }
        const int c = 2 * (j >> 1) + p, sgc = j & 1;
        const bool valid = j < n && ((size_t)Rr << lam) + c < ((size_t)4 << lev) - 2;
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int row = rt * 16 + (lane >> 4) + 4 * v;
          if (row < M && j < n) {
            const int o = (b0 + c) * M2 + sgc * M + row;
            cres[o] = valid ? D[v] + mres[g.ld * M2 + o] : 0.0;
          }
        }
      }
      sync();
    }
    return;
  }
  const int kd = tid % M, ngrp = NT / M2;
  if (ngrp == 0) {  // fewer threads than 2M: one output per thread and pass, D from memory
    for (int lam = 1; lam <= g.gd; lam++) {
      const int n = 1 << lam, lev = g.ld + lam;
      const int b0 = n - 1, bp = (n >> 1) - 1;
      for (int idx = tid; idx < n * M2; idx += NT) {
        const int ci = idx / M2, oo = idx - ci * M2, sg = oo / M, k = oo - sg * M;
        const size_t r = ((size_t)Rr << lam) + ci;
        double v = 0;
        if (r < ((size_t)4 << lev) - 2) {
          const double *drow = Dg + (r & 1) * MM + (size_t)k * M;
          const double *par = cres + (bp + (ci >> 1)) * M2 + sg * M;
          for (int l = 0; l < M; l++) v += drow[l] * par[l];
          v += mres[(g.ld + b0 + ci) * M2 + oo];
        }
        cres[(b0 + ci) * M2 + oo] = v;
      }
      sync();
    }
    return;
  }
  auto l2l = [&](int p, const double *par) -> double {
    double a0 = 0, a1 = 0;
    if (MT > 0) {
#pragma unroll
      for (int l = 0; l < (MT > 0 ? MT : 1); l += 2) {
        a0 += (p ? d1r[l] : d0r[l]) * par[l];
        if (l + 1 < M) a1 += (p ? d1r[l + 1] : d0r[l + 1]) * par[l + 1];
      }
    } else {
      for (int l = 0; l < M; l++) a0 += Dg[p * MM + kd * M + l] * par[l];
    }
    return a0 + a1;
  };
  for (int lam = 1; lam <= g.gd; lam++) {
    const int n = 1 << lam, lev = g.ld + lam;
    const int b0 = n - 1, bp = (n >> 1) - 1;
    if (tid < ngrp * M2) {
      const int oo = tid % M2, sg = oo / M;
      for (int ci = tid / M2; ci < n; ci += ngrp) {
        const size_t r = ((size_t)Rr << lam) + ci;
        double v = 0;
        if (r < ((size_t)4 << lev) - 2)
          v = l2l((int)(r & 1), cres + (bp + (ci >> 1)) * M2 + sg * M) +
              mres[(g.ld + b0 + ci) * M2 + oo];
        cres[(b0 + ci) * M2 + oo] = v;
      }
    }
    sync();
  }
}

template <int DIR, int PRE, int MT>
__global__ void __launch_bounds__(kDownNT) __attribute__((amdgpu_waves_per_eu(2))) k_down(
    const double *__restrict__ in, double *__restrict__ out, LineMap lm, Geo g, int band_only,
    const double *__restrict__ wsall, size_t wsstride, const double *__restrict__ Dg,
    const double *__restrict__ Qg, const double *__restrict__ Tg, const double *__restrict__ tab,
    const double *__restrict__ htab) {
  extern __shared__ double sm[];
  const int M = MT > 0 ? MT : g.M;
  const int s = g.s, L = g.L, N = g.N, M2 = 2 * M, MM = M * M;
  const int G = 1 << g.gd;
  const int Rr = blockIdx.x;
  const int tid = threadIdx.x, NT = kDownNT;
  const size_t row0 = (size_t)Rr * G * 2 * s;
  if (row0 >= (size_t)N) return;
  const int nr = (int)std::min((size_t)G * 2 * s, (size_t)N - row0);
  const size_t off = line_off(lm, blockIdx.y);
  const double *src = in + off;
  double *dst = out + off;
  const double *ws = wsall + blockIdx.y * wsstride;
  const WsLayout lay = ws_layout(g);
  const double *m2l = ws + lay.m2l;
  const DownLds ld = down_lds(g);
  double *xs = sm + ld.xs;
  const int nxs = (ld.aw - ld.xs) / 2;
  double *aw = sm + ld.aw;
  double *cres = sm + ld.cres;
  double *mres = sm + ld.mres;
  double *qp = sm + ld.qp;
  double *ob = sm + ld.out;

  // ---- phase 0: every load issued before any store; addresses are clamped
  // into range and out-of-range values replaced afterwards, so the loads
  // issue as one straight-line batch
  const int nwin = G * 2 * s + 2 * s + 32;
  constexpr int QW = 10;
  double vx[QW], va[QW];
#pragma unroll
  for (int q = 0; q < QW; q++) {
    const size_t i = std::min(row0 + q * NT + tid, (size_t)N - 1);
    vx[q] = src[i * lm.es];
    va[q] = tab[i];
  }
  // M2L vectors: the ancestor chain (levels 0..ld-1) then the subtree slots
  const int nm = band_only ? 0 : (g.ld + 2 * G - 1) * M2;
  constexpr int QM = 8;
  double vm[QM];
  bool okm[QM];
#pragma unroll
  for (int q = 0; q < QM; q++) {
    const int e = q * NT + tid;
    const size_t a = m2l_elem(g, Rr, e < nm ? e : 0, M2, okm[q]);
    okm[q] = okm[q] && e < nm;
    vm[q] = m2l[okm[q] ? a : 0];
  }
  // ancestor chain in one step: task (lev, k) < ld*M takes row k of the
  // composite operator Q(ld - lev, Rr mod 2^(ld - lev)) from level lev to ld
  const int nq = band_only ? 0 : g.ld * M;
  double qr[MT > 0 ? MT : 1];
  if (MT > 0 && nq > 0) {
    const double *qrow = qrow_ptr(Qg, g, Rr, tid < nq ? tid : 0, M);
#pragma unroll
    for (int l = 0; l < (MT > 0 ? MT : 1); l++) qr[l] = qrow[l * M];
  }
#pragma unroll
  for (int q = 0; q < QW; q++) {
    const int e = q * NT + tid;
    if (e < nwin) {
      const size_t i = row0 + e;
      const double u = i < (size_t)N ? vx[q] : 0.0;
      xs[(e & 1) * nxs + sk4(e >> 1)] = (DIR == 1 && i > 0) ? u * (double)i : u;  // v_i = i u_i
      aw[sk8(e)] = i < (size_t)N ? va[q] : 0.0;
    }
  }
  for (int e = QW * NT + tid; e < nwin; e += NT) {
    const size_t i = row0 + e;
    double v = i < (size_t)N ? src[i * lm.es] : 0.0;
    if (DIR == 1 && i > 0) v *= (double)i;
    xs[(e & 1) * nxs + sk4(e >> 1)] = v;
    aw[sk8(e)] = i < (size_t)N ? tab[i] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < QM; q++) {
    const int e = q * NT + tid;
    if (e < nm) mres[e] = okm[q] ? vm[q] : 0.0;
  }
  for (int e = QM * NT + tid; e < nm; e += NT) {
    bool ok;
    const size_t a = m2l_elem(g, Rr, e, M2, ok);
    mres[e] = ok ? m2l[a] : 0.0;
  }
  // this thread's rows of the two L2L operators (k = tid % M), needed from
  // the subtree pass on; loaded after the window registers are released
  const int kd = tid % M;
  double d0r[MT > 0 ? MT : 1], d1r[MT > 0 ? MT : 1];
  if (MT > 0 && !band_only) {
#pragma unroll
    for (int l = 0; l < (MT > 0 ? MT : 1); l++) {
      d0r[l] = Dg[kd * M + l];
      d1r[l] = Dg[MM + kd * M + l];
    }
  }
  __syncthreads();

  // ---- phase 2: local expansions of the subtree
  if (!band_only)
    subtree_locals<MT, false>(g, Rr, tid, NT, mres, qp, cres, qr, d0r, d1r, Qg, Dg, nullptr,
                             LdsSync{});

  // ---- phase 3: band + leaf evaluation.  Task (chunk, group gam): rows
  // gam*4.. of parity 0, then rows (ngrp-1-gam)*4.. of parity 1, so every
  // thread runs the same number of columns, parity is uniform across a wave
  // in each half, and 16 consecutive lanes take 16 consecutive row groups.
  const int ngrp = (s + kR - 1) / kR;
  const int ntask = G * ngrp;
  const int nLm1 = g.nL - 1;
  const size_t nvalid = ((size_t)2 << L) - 2;
  const size_t c0 = (size_t)Rr * G * s;
  const int aligned = (s & 3) == 0;
  for (int task = tid; task < ntask; task += NT) {
    const int pp = task % ngrp, gch = task / ngrp;
    const size_t beta = (size_t)Rr * G + gch;
    for (int sg = 0; sg < 2; sg++) {
      const int gam = sg == 0 ? pp : ngrp - 1 - pp;
      const int cnt = (N - sg + 1) / 2;  // elements of parity sg
      const size_t cend = beta < (size_t)nLm1 ? (beta + 2) * s : (size_t)cnt;
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
      band_dispatch<DIR, PRE>(sg, aligned, xsg, aw, htab, rl0, len, acc);
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
// k_down8: the downward pass for leaf sizes s = 16, 32, 64.  grid
// (2^(ld+2), lines), 64 nw threads with nw = s/16 waves, G = 2^gd <= 16 leaf
// chunks per work-group.
//
// Near field in 8-row register blocks.  Rows of parity sg, chunk g and 8-row
// group gam = 2p + del run over column offsets t in [0, len), len = 2s - 8 gam
// (shorter in the last chunk).  Pair p's columns are cut at ts = s - 8p into
// segment A = [0, ts) and B = [ts, len); wave w runs segment A of pair w and
// segment B of pair nw-1-w, so every wave does 2s - 8(nw-1) columns (8 fewer
// for del = 1).  Lane = g + 16 sg + 32 del: all 64 lanes of a wave sit at the
// same t, so the Toeplitz factor Lambda(t - rho) (dn for C2L) is one uniform
// (scalar) load per column, and the LDS layouts make every lane address of an
// x or Lambda read distinct modulo 32 eight-byte slots:
//   x, per parity, half index h from the work-group start: h + h/s (one pad
//     per chunk; a trip of 8 columns never straddles one), parity 1 at +nxs
//     with nxs = 16 mod 32;
//   Lambda, per chunk g, local index q = 2(r - g s) + sg + t + rho
//     (< 3s + 16): g LW + q with LW = 2 mod 32.
// Far field: the subtree's local expansions evaluated at the leaf nodes on
// v_mfma_f64_16x16x4f64, F_par (s x G) = T_par (s x M) C_par (M x G), added to
// the band sums staged in LDS; then scaling and one coalesced store.
struct Down8Lds {
  int xs, nxs, lw, LW, hd, dl, cres, mres, qp, total;
};
__host__ __device__ inline Down8Lds down8_lds(const Geo &g) {
  const int G = 1 << g.gd, s = g.s, M2 = 2 * g.M;
  Down8Lds d;
  const int nh = (G + 1) * s + 8;  // half indices per parity, one trip of overrun
  const int nhp = nh + nh / s + 1;
  d.xs = 0;
  d.nxs = nhp + (48 - nhp % 32) % 32;  // = 16 mod 32
  d.LW = 3 * s + 16;
  d.LW += (34 - d.LW % 32) % 32;  // = 2 mod 32
  d.lw = d.xs + 2 * d.nxs;
  d.hd = d.lw + G * d.LW;  // band head, 2s + 16
  d.dl = d.hd + 2 * s + 16;  // D(0), D(1) with rows padded to M + 1
  d.cres = d.dl + g.M * M2 + M2;
  d.mres = d.cres + (2 * G - 1) * M2;
  d.qp = d.mres + (g.ld + 2 * G - 1) * M2;
  d.total = d.qp + std::max(g.ld, 1) * M2;
  return d;  // the output staging (2 G s doubles) reuses the x windows
}
size_t down8_shared_bytes(const Geo &g) { return 8 * (size_t)down8_lds(g).total; }
bool down8_eligible(const Geo &g) { return g.s == 16 || g.s == 32 || g.s == 64; }

// One band segment of 8 rows: columns t in [t0, t1) (t0 = 0 or a multiple of
// 8 shared by the wave), software-pipelined half a trip (4 columns) ahead.
// xsg: the lane's parity window; h0: half index of the first row from the
// work-group start; pw: the lane's Lambda window at q = t = 0 of row 0;
// head: the zero-padded band head in LDS (read at wave-uniform addresses:
// keeping it out of scalar memory leaves LDS reads the only users of the
// lgkm counter, so the half-trip prefetch is not drained by scalar waits).
// RAGGED: columns
// t >= len are masked (the last chunk, where len is not a multiple of 8).
template <int DIR, int PRE, bool RAGGED, bool LEAN = false>
__device__ __forceinline__ void band_seg8(const double *__restrict__ xsg,
                                          const double *__restrict__ pw,
                                          const double *__restrict__ head, int h0, int sh,
                                          int t0, int t1, int len, double acc[8]) {
  const int s = 1 << sh;
  const int ntrip = t1 > t0 ? (t1 - t0 + 7) >> 3 : 0;
  if (ntrip == 0) return;
  // ring state at t0: heads Lambda(t0 - 8 + k), refills p(t0 + k)
  double hr[8], pr[8];
  const double *pp = pw + t0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    hr[k] = t0 >= 8 ? head[t0 - 8 + k] : 0.0;
    pr[k] = pp[k];
  }
  // the trip being fetched: x pointer and position inside its chunk
  int hq = h0 + t0;
  const double *xq = xsg + hq + (hq >> sh);
  int mq = hq & (s - 1);
  const double *hp = head + t0;
  int tt = t0;
  double XA[4], PA[4], HA[4], XB[4], PB[4], HB[4];
  // half H (0: columns 0-3, 1: 4-7) of the current trip
#define FETCH(X, P, H, HALF)                                         \
  _Pragma("unroll") for (int k_ = 0; k_ < 4; k_++) {                 \
    X[k_] = xq[4 * (HALF) + k_];                                     \
    if (RAGGED) X[k_] = tt + 4 * (HALF) + k_ < len ? X[k_] : 0.0;    \
    P[k_] = pp[8 + 4 * (HALF) + k_];                                 \
    H[k_] = hp[4 * (HALF) + k_];                                     \
  }
#define NEXT_TRIP                   \
  {                                 \
    const int pad_ = mq + 8 == s;   \
    mq = pad_ ? 0 : mq + 8;         \
    xq += 8 + pad_;                 \
    pp += 8;                        \
    hp += 8;                        \
    tt += 8;                        \
  }
#define TERM(hv, pv, xv, a)          \
  if (DIR == 0)                      \
    a += ((hv) * (pv)) * (xv);       \
  else if (PRE)                      \
    a -= ((hv) * (pv)) * (xv);       \
  else                               \
    a -= ((hv) * (xv)) * (pv);
  // column t (ring slot K = t mod 8): head Lambda(t) enters slot K, row rho
  // uses Lambda(t - rho) and p(t + rho), then p(t + 8) replaces p(t)
#define STEP(K, X, P, H, J)                                \
  {                                                        \
    hr[K] = H[J];                                          \
    const double x_ = X[J];                                \
    TERM(hr[(K) & 7], pr[(K) & 7], x_, acc[0])             \
    TERM(hr[((K) + 7) & 7], pr[((K) + 1) & 7], x_, acc[1]) \
    TERM(hr[((K) + 6) & 7], pr[((K) + 2) & 7], x_, acc[2]) \
    TERM(hr[((K) + 5) & 7], pr[((K) + 3) & 7], x_, acc[3]) \
    TERM(hr[((K) + 4) & 7], pr[((K) + 4) & 7], x_, acc[4]) \
    TERM(hr[((K) + 3) & 7], pr[((K) + 5) & 7], x_, acc[5]) \
    TERM(hr[((K) + 2) & 7], pr[((K) + 6) & 7], x_, acc[6]) \
    TERM(hr[((K) + 1) & 7], pr[((K) + 7) & 7], x_, acc[7]) \
    pr[K] = P[J];                                          \
    if (LEAN) __builtin_amdgcn_sched_barrier(0);           \
  }
  FETCH(XA, PA, HA, 0)
  for (int b = 0; b < ntrip; b++) {
    FETCH(XB, PB, HB, 1)
    STEP(0, XA, PA, HA, 0) STEP(1, XA, PA, HA, 1) STEP(2, XA, PA, HA, 2) STEP(3, XA, PA, HA, 3)
    NEXT_TRIP
    FETCH(XA, PA, HA, 0)
    STEP(4, XB, PB, HB, 0) STEP(5, XB, PB, HB, 1) STEP(6, XB, PB, HB, 2) STEP(7, XB, PB, HB, 3)
  }
#undef STEP
#undef TERM
#undef NEXT_TRIP
#undef FETCH
}

// Modes: K8_FULL band and far field; K8_BAND the near field alone, scaled
// (directM).
enum { K8_FULL = 0, K8_BAND = 1 };

template <int DIR, int PRE, int MT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_down8(
    const double *__restrict__ in, double *__restrict__ out, LineMap lm, Geo g, int mode,
    const double *__restrict__ wsall, size_t wsstride, const double *__restrict__ Dg,
    const double *__restrict__ Qg, const double *__restrict__ Tg, const double *__restrict__ tab,
    const double *__restrict__ head) {
  extern __shared__ double sm[];
  const int M = MT > 0 ? MT : g.M;
  const int s = g.s, N = g.N, M2 = 2 * M, MM = M * M;
  // N: the plan geometry; Nl: this line's length (inputs past it read as zero)
  const int Nl = line_len(lm, blockIdx.y, N);
  const int G = 1 << g.gd, sh = 31 - __builtin_clz(s);
  const int nw = s >> 4, NT = 64 * nw;
  const int Rr = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  // the wave index is wave-uniform; saying so lets the band's column offsets
  // and head loads live in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool do_band = true, do_far = mode == K8_FULL;
  const size_t row0 = (size_t)Rr * G * 2 * s;
  if (row0 >= (size_t)Nl) return;  // the whole work-group
  const int nr = (int)std::min((size_t)G * 2 * s, (size_t)Nl - row0);
  const size_t hc0 = (size_t)Rr * G * s;  // first half index of the work-group
  const size_t off = line_off(lm, blockIdx.y);
  const double *src = in + off;
  double *dst = out + off;
  const double *m2l = wsall + blockIdx.y * wsstride + ws_layout(g).m2l;
  const Down8Lds ly = down8_lds(g);
  double *xs = sm + ly.xs, *lw = sm + ly.lw, *cres = sm + ly.cres;
  double *mres = sm + ly.mres, *qp = sm + ly.qp, *hd = sm + ly.hd, *dl = sm + ly.dl;
  // output staging after the band: segment A sums in ob (the x windows),
  // segment B sums in ob2 (the Lambda windows); slot of output e: e + e / 2s
  // (one pad per chunk spreads the lanes of every writer over distinct banks)
  double *ob = xs, *ob2 = lw;
  auto slot = [&](int e) { return e + (e >> (sh + 1)); };

  // ---- phase 0: every global load of the kernel issued up front (clamped
  // addresses: one straight-line batch).  The tree operands (M2L vectors, the
  // chain row of Q, D) are needed first; the band windows (x, the Lambda span,
  // the head) stay in flight through the tree pass, whose barriers are
  // LDS-only.
  const int nm = do_far ? (g.ld + 2 * G - 1) * M2 : 0;
  constexpr int QM = 6;
  double vm[QM];
  bool okm[QM];
  const int nq = do_far ? g.ld * M : 0;
  double qr[MT > 0 ? MT : 1];
  constexpr int QD = MT > 0 ? (2 * MT * MT + 63) / 64 : 1;  // D(0), D(1) for the LDS copy
  double vd[QD];
  if (do_far) {
#pragma unroll
    for (int q = 0; q < QM; q++) {
      const int e = q * NT + tid;
      const size_t a = m2l_elem(g, Rr, e < nm ? e : 0, M2, okm[q]);
      okm[q] = okm[q] && e < nm;
      vm[q] = m2l[okm[q] ? a : 0];
    }
    if (MT > 0 && nq > 0) {
      const double *qrow = qrow_ptr(Qg, g, Rr, tid < nq ? tid : 0, M);
#pragma unroll
      for (int l = 0; l < (MT > 0 ? MT : 1); l++) qr[l] = qrow[l * M];
    }
#pragma unroll
    for (int q = 0; q < QD; q++) {
      const int e = q * NT + tid;
      vd[q] = Dg[e < 2 * MM ? e : 0];
    }
  }
  // keep the tree operands ahead of the windows in the memory pipeline: the
  // stores below then wait for them alone
  __builtin_amdgcn_sched_barrier(0);
  // x: input elements row0 + e, e < nX (both parities interleaved)
  const int nX = do_band ? 2 * ((G + 1) * s + 8) : 0;
  constexpr int QX = 9;  // ceil(nX / NT) for G = 16
  double vx[QX];
  // Lambda: the contiguous span 2 hc0 + e, e < 2Gs + s + 16; element e lies
  // in chunk window e / 2s and, when e mod 2s < s + 16, also in the one before
  const int QL = 3 * s + 16, nLs = do_band ? 2 * G * s + s + 16 : 0;
  constexpr int QLS = 9;  // ceil(nLs / NT) for G = 16
  double vl[QLS];
  const int nhd = 2 * s + 16;  // <= NT
  double vh = 0.0;
  if (do_band) {
#pragma unroll
    for (int q = 0; q < QX; q++)
      vx[q] = src[std::min(row0 + q * NT + tid, (size_t)Nl - 1) * lm.es];
#pragma unroll
    for (int q = 0; q < QLS; q++)
      vl[q] = tab[std::min(2 * hc0 + q * NT + tid, (size_t)N - 1)];
    vh = head[tid < nhd ? tid : 0];
  }
  __builtin_amdgcn_sched_barrier(0);
  if (do_far) {
#pragma unroll
    for (int q = 0; q < QM; q++) {
      const int e = q * NT + tid;
      if (e < nm) mres[e] = okm[q] ? vm[q] : 0.0;
    }
  }
  for (int e = QM * NT + tid; e < nm; e += NT) {
    bool ok;
    const size_t a = m2l_elem(g, Rr, e, M2, ok);
    mres[e] = ok ? m2l[a] : 0.0;
  }
  if (do_far) {
#pragma unroll
    for (int q = 0; q < QD; q++) {
      const int e = q * NT + tid;
      if (MT > 0 && e < 2 * MM) dl[(e / M) * (M + 1) + e % M] = vd[q];
    }
    for (int e = (MT > 0 ? QD * NT : 0) + tid; e < 2 * MM; e += NT)
      dl[(e / M) * (M + 1) + e % M] = Dg[e];
  }
  lds_sync();

  // ---- phase 2: local expansions of the subtree
  if (do_far)
    subtree_locals<MT, true>(g, Rr, tid, NT, mres, qp, cres, qr, nullptr, nullptr, Qg, Dg, dl,
                            LdsSync{});

  // A operand (T) of this wave's two far-field tiles, in flight during the band
  const int ntm = s >> 4, ksteps = (M + 3) >> 2;
  constexpr int QK = MT > 0 ? (MT + 3) / 4 : 1;
  double ta[2][QK];
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int tile = wave + j * nw, sgf = tile / ntm, rt = tile - sgf * ntm;
    const int m = rt * 16 + (lane & 15);
#pragma unroll
    for (int kk = 0; kk < QK; kk++) {
      const int k = kk * 4 + (lane >> 4);
      ta[j][kk] = (MT > 0 && do_far && k < M) ? Tg[((size_t)sgf * s + m) * M + k] : 0.0;
    }
  }

  if (do_band) {
    // ---- the band windows
#pragma unroll
    for (int q = 0; q < QX; q++) {
      const int e = q * NT + tid;
      if (e < nX) {
        const size_t i = row0 + e;
        const double u = i < (size_t)Nl ? vx[q] : 0.0;
        const int h = e >> 1;
        xs[(e & 1) * ly.nxs + h + (h >> sh)] = (DIR == 1 && i > 0) ? u * (double)i : u;  // v_i = i u_i
      }
    }
    for (int e = QX * NT + tid; e < nX; e += NT) {
      const size_t i = row0 + e;
      double u = i < (size_t)Nl ? src[i * lm.es] : 0.0;
      if (DIR == 1 && i > 0) u *= (double)i;
      const int h = e >> 1;
      xs[(e & 1) * ly.nxs + h + (h >> sh)] = u;
    }
    auto put_lambda = [&](int e, double v) {
      const int c = e >> (sh + 1), q = e - (c << (sh + 1));
      if (2 * hc0 + e >= (size_t)N) v = 0.0;
      if (c < G) lw[c * ly.LW + q] = v;
      if (c > 0 && q + 2 * s < QL) lw[(c - 1) * ly.LW + q + 2 * s] = v;
    };
#pragma unroll
    for (int q = 0; q < QLS; q++) {
      const int e = q * NT + tid;
      if (e < nLs) put_lambda(e, vl[q]);
    }
    for (int e = QLS * NT + tid; e < nLs; e += NT)
      put_lambda(e, tab[std::min(2 * hc0 + e, (size_t)N - 1)]);
    if (tid < nhd) hd[tid] = vh;
    lds_sync();

    // ---- phase 3: near field.  Lane (chunk gl, parity sg, del); wave:
    // segment A of pair wave, segment B of pair nw-1-wave
    const int gl = lane & 15, sg = (lane >> 4) & 1, del = lane >> 5;
    const int cnt = (N - sg + 1) / 2;  // elements of parity sg
    const size_t beta = (size_t)Rr * G + gl;
    const size_t nLm1 = (size_t)g.nL - 1;
    const size_t cend = beta < nLm1 ? (beta + 2) * s : (size_t)cnt;
    const bool ragged = (size_t)(Rr + 1) * G > nLm1;  // work-group holds the last chunk
    const double *xsg = xs + sg * ly.nxs;
    double accA[8], accB[8];
    int rA = -1, rB = -1;  // first row (half index) of each segment, -1: none
    {
      const int p = wave, gam = 2 * p + del, ts = s - 8 * p;
      const size_t r0 = beta * s + 8 * gam;
      const bool live = gl < G && r0 < (size_t)cnt;
      const int len = live ? (int)(cend - r0) : 0;
      const int h0 = (int)(r0 - hc0);
      const double *pw = lw + gl * ly.LW + 16 * gam + sg;
#pragma unroll
      for (int k = 0; k < 8; k++) accA[k] = 0.0;
      if (live) {
        rA = h0;
        if (DIR == 1) {  // C2L diagonal: sqrt(pi) an_i v_i (src/C/leg2cheb.c:416)
          const double *xr = xsg + h0 + (h0 >> sh);
#pragma unroll
          for (int k = 0; k < 8; k++) accA[k] = kSqrtPi * xr[k] * pw[2 * k];
        }
      }
      if (ragged)
        band_seg8<DIR, PRE, true>(xsg, pw, hd, h0, sh, 0, std::min(ts, len), len, accA);
      else
        band_seg8<DIR, PRE, false>(xsg, pw, hd, h0, sh, 0, ts, len, accA);
    }
    {
      const int p = nw - 1 - wave, gam = 2 * p + del, ts = s - 8 * p;
      const size_t r0 = beta * s + 8 * gam;
      const bool live = gl < G && r0 < (size_t)cnt;
      const int len = live ? (int)(cend - r0) : 0;
      const int h0 = (int)(r0 - hc0);
      const double *pw = lw + gl * ly.LW + 16 * gam + sg;
#pragma unroll
      for (int k = 0; k < 8; k++) accB[k] = 0.0;
      if (live) rB = h0;
      if (ragged)
        band_seg8<DIR, PRE, true>(xsg, pw, hd, h0, sh, ts, len, len, accB);
      else
        band_seg8<DIR, PRE, false>(xsg, pw, hd, h0, sh, ts, len, len, accB);
    }
    lds_sync();  // windows dead: the staging takes their place

    // ---- phase 4: segment A sums to ob, segment B sums to ob2 (each row
    // once in each)
    auto stage = [&](int rl, const double *acc, double *dstg) {
      if (rl < 0) return;
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (hc0 + rl + k < (size_t)cnt) dstg[slot(2 * (rl + k) + sg)] = acc[k];
    };
    stage(rA, accA, ob);
    stage(rB, accB, ob2);
    lds_sync();
  }
  if (do_far) {
    // F_par^T (G x s) = C_par^T (G x M) T_par^T (M x s), tile (parity sgf,
    // rows 16 rt..): A[n][k] = c_k of leaf n, B[k][m] = T_par[16 rt + m][k];
    // lane l then holds leaf l/16 + 4v, row 16 rt + l%16; added into ob.
    const int n = lane & 15;
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int tile = wave + j * nw, sgf = tile / ntm, rt = tile - sgf * ntm;
      d4 D = {0, 0, 0, 0};
      if (MT > 0) {
        double av[QK];
#pragma unroll
        for (int kk = 0; kk < QK; kk++) {
          const int k = kk * 4 + (lane >> 4);
          av[kk] = (n < G && k < M) ? cres[(G - 1 + n) * M2 + sgf * M + k] : 0.0;
        }
#pragma unroll
        for (int kk = 0; kk < QK; kk++) D = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kk], ta[j][kk], D, 0, 0, 0);
      } else {
        for (int kk = 0; kk < ksteps; kk++) {
          const int k = kk * 4 + (lane >> 4);
          const double bv = k < M ? Tg[((size_t)sgf * s + rt * 16 + n) * M + k] : 0.0;
          const double av = (n < G && k < M) ? cres[(G - 1 + n) * M2 + sgf * M + k] : 0.0;
          D = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, D, 0, 0, 0);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int leaf = (lane >> 4) + 4 * v, m = rt * 16 + n;
        const int e = 2 * (leaf * s + m) + sgf;
        if (leaf < G && e < nr) ob[slot(e)] += D[v];
      }
    }
    lds_sync();
  }
  // ---- store: the scaled sum
  for (int e = tid; e < nr; e += NT) {
    const size_t i = row0 + e;
    double v = ob[slot(e)] + ob2[slot(e)];
    if (DIR == 0) {
      v *= k2Pi;
      if (i == 0) v *= 0.5;
    } else {
      v *= (double)i + 0.5;
    }
    dst[i * lm.es] = v;
  }
}

// ---------------------------------------------------------------------------
// k_down8r: k_down8's work (downward pass, near-field band, leaf evaluation;
// src/C/leg2cheb.c:1437-1483 and the band of :236-535) with wave roles.  The
// work-group has 2 s/16 waves: the first s/16 (the tree role) load the M2L
// vectors and operators, run the ancestor chain and the subtree's local
// expansions and evaluate them at the leaf nodes on MFMA (as k_far); the
// others (the band role) load the input and Lambda windows and sum the band,
// one row pair of 8 rows over all its columns per wave.  The roles share no
// LDS and synchronise among themselves with RoleSync; the one hardware
// barrier between them is the join before the scaled sum is stored.  So the
// band's FP64 VALU work overlaps the tree role's loads, LDS round trips and
// dependent MFMA steps inside each work-group, at 128 VGPRs and 78.5 KB of
// LDS (two work-groups, 16 waves, per CU at s = 64).
struct Down8rLds {
  int xs, nxs, lw, LW, hd, cres, dl, mres, qp, obf, ctr, total;
};
__host__ __device__ inline Down8rLds down8r_lds(const Geo &g) {
  const Down8Lds d8 = down8_lds(g);
  const int G = 1 << g.gd, s = g.s, M2 = 2 * g.M;
  Down8rLds d;
  d.xs = d8.xs;
  d.nxs = d8.nxs;
  d.lw = d8.lw;
  d.LW = d8.LW;
  d.hd = d8.hd;
  d.cres = d.hd + 2 * s + 16;
  d.dl = d.cres + (2 * G - 1) * M2;  // D(0), D(1) with rows padded to M + 1
  d.mres = d.dl + g.M * M2 + M2;
  d.qp = d.mres + (g.ld + 2 * G - 1) * M2;
  const int tree_end = d.qp + std::max(g.ld, 1) * M2;
  d.obf = d.dl;  // far-field staging (2 G s + G) over the tree's dead operands
  d.ctr = std::max(tree_end, d.obf + 2 * G * s + G);
  d.total = d.ctr + 1;  // the two roles' barrier counters
  return d;  // the band staging (2 G s + G) reuses the x windows
}
size_t down8r_shared_bytes(const Geo &g) { return 8 * (size_t)down8r_lds(g).total; }

template <int DIR, int PRE, int MT>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) k_down8r(
    const double *__restrict__ in, double *__restrict__ out, LineMap lm, Geo g,
    const double *__restrict__ wsall, size_t wsstride, const double *__restrict__ Dg,
    const double *__restrict__ Qg, const double *__restrict__ Tg, const double *__restrict__ tab,
    const double *__restrict__ head, int shift) {
  extern __shared__ double sm[];
  const int M = MT > 0 ? MT : g.M;
  const int s = g.s, N = g.N, M2 = 2 * M, MM = M * M;
  const int Nl = line_len(lm, blockIdx.y, N);
  const int G = 1 << g.gd, sh = 31 - __builtin_clz(s);
  const int nw = s >> 4, NT = 64 * nw;  // waves and threads of each role
  const int Rr = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool tree = wave < nw;
  const int tid = tree ? (int)threadIdx.x : (int)threadIdx.x - NT;  // index in the role
  const size_t row0 = (size_t)Rr * G * 2 * s;
  if (row0 >= (size_t)Nl) return;  // the whole work-group
  const int nr = (int)std::min((size_t)G * 2 * s, (size_t)Nl - row0);
  const size_t hc0 = (size_t)Rr * G * s;
  const size_t off = line_off(lm, blockIdx.y);
  const double *src = in + off;
  double *dst = out + off;
  // the layout is work-group uniform; its divisions run on the VALU, so its
  // fields are moved to scalar registers rather than held (and spilled) in
  // vector registers through the band loops
  Down8rLds ly = down8r_lds(g);
  ly.nxs = __builtin_amdgcn_readfirstlane(ly.nxs);
  ly.lw = __builtin_amdgcn_readfirstlane(ly.lw);
  ly.LW = __builtin_amdgcn_readfirstlane(ly.LW);
  ly.hd = __builtin_amdgcn_readfirstlane(ly.hd);
  ly.cres = __builtin_amdgcn_readfirstlane(ly.cres);
  ly.dl = __builtin_amdgcn_readfirstlane(ly.dl);
  ly.mres = __builtin_amdgcn_readfirstlane(ly.mres);
  ly.qp = __builtin_amdgcn_readfirstlane(ly.qp);
  ly.obf = __builtin_amdgcn_readfirstlane(ly.obf);
  ly.ctr = __builtin_amdgcn_readfirstlane(ly.ctr);
  double *xs = sm + ly.xs, *lw = sm + ly.lw, *hd = sm + ly.hd;
  double *cres = sm + ly.cres, *dl = sm + ly.dl, *mres = sm + ly.mres, *qp = sm + ly.qp;
  double *ob = xs, *obf = sm + ly.obf;
  int *ctr = reinterpret_cast<int *>(sm + ly.ctr);
  auto slot = [&](int e) { return e + (e >> (sh + 1)); };
  // band lanes (both roles): chunk gl = lane % 16, parity sg = lane / 16 % 2,
  // del = lane / 32; row group gam = 2 p + del of pair p holds rows 8 gam ..
  // 8 gam + 7 of its chunk.  Computed where used, so that nothing of it is
  // live through the tree role's phases.
  // columns [t0, t1) of pair p's row group (t1 < 0: to its end) added to acc
  auto band = [&](int p, int t0, int t1, double acc[8]) {
    const int gl = lane & 15, sg = (lane >> 4) & 1, gam = 2 * p + (lane >> 5);
    const int cnt = (N - sg + 1) / 2;
    const size_t beta = (size_t)Rr * G + gl, nLm1 = (size_t)g.nL - 1;
    const size_t cend = beta < nLm1 ? (beta + 2) * s : (size_t)cnt;
    const size_t r0 = beta * s + 8 * gam;
    const bool live = gl < G && r0 < (size_t)cnt;
    const int len = live ? (int)(cend - r0) : 0;
    const int h0 = (int)(r0 - hc0);
    const double *pw = lw + gl * ly.LW + 16 * gam + sg;
    const double *xsg = xs + sg * ly.nxs;
    const int a0 = std::min(t0, len), a1 = t1 < 0 ? len : std::min(t1, len);
    if ((size_t)(Rr + 1) * G > nLm1)  // the work-group holds the last chunk
      band_seg8<DIR, PRE, true, true>(xsg, pw, hd, h0, sh, a0, a1, len, acc);
    else
      band_seg8<DIR, PRE, false, true>(xsg, pw, hd, h0, sh, a0, a1, len, acc);
  };
  // the sums of pair p's row group of this lane into staging st (= or +=)
  auto stage_group = [&](int p, const double acc[8], double *st, bool add) {
    // a fresh copy of the lane index (the empty asm hides that it equals
    // the one the band used), so that the band's row indices are recomputed
    // here instead of being held in registers through the band
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const int gl = ln & 15, sg = (ln >> 4) & 1;
    const int cnt = (N - sg + 1) / 2;
    const size_t r0 = ((size_t)Rr * G + gl) * s + 8 * (2 * p + (ln >> 5));
    if (gl >= G || r0 >= (size_t)cnt) return;
    const int rl = (int)(r0 - hc0);
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (r0 + q < (size_t)cnt) {
        double &o = st[slot(2 * (rl + q) + sg)];
        o = add ? o + acc[q] : acc[q];
      }
  };
  if (threadIdx.x == 0) {
    ctr[0] = 0;
    ctr[1] = 0;
  }
  lds_sync();  // hardware barrier 1 of 2

  if (tree) {
    // the tree role's chain of dependent LDS and MFMA steps is the critical
    // path: its waves issue ahead of the band role's FMA stream
    __builtin_amdgcn_s_setprio(3);
    RoleSync rs{ctr, nw, 0};
    const double *m2l = wsall + blockIdx.y * wsstride + ws_layout(g).m2l;
    const int nm = (g.ld + 2 * G - 1) * M2;
    constexpr int QM = 6;
    double vm[QM];
    bool okm[QM];
    const int nq = g.ld * M;
    double qr[MT > 0 ? MT : 1];
    constexpr int QD = MT > 0 ? (2 * MT * MT + 63) / 64 : 1;
    double vd[QD];
#pragma unroll
    for (int q = 0; q < QM; q++) {
      const int e = q * NT + tid;
      const size_t a = m2l_elem(g, Rr, e < nm ? e : 0, M2, okm[q]);
      okm[q] = okm[q] && e < nm;
      vm[q] = m2l[okm[q] ? a : 0];
    }
    if (MT > 0 && nq > 0) {
      const double *qrow = qrow_ptr(Qg, g, Rr, tid < nq ? tid : 0, M);
#pragma unroll
      for (int l = 0; l < (MT > 0 ? MT : 1); l++) qr[l] = qrow[l * M];
    }
#pragma unroll
    for (int q = 0; q < QD; q++) {
      const int e = q * NT + tid;
      vd[q] = Dg[e < 2 * MM ? e : 0];
    }
#pragma unroll
    for (int q = 0; q < QM; q++) {
      const int e = q * NT + tid;
      if (e < nm) mres[e] = okm[q] ? vm[q] : 0.0;
    }
    for (int e = QM * NT + tid; e < nm; e += NT) {
      bool ok;
      const size_t a = m2l_elem(g, Rr, e, M2, ok);
      mres[e] = ok ? m2l[a] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < QD; q++) {
      const int e = q * NT + tid;
      if (MT > 0 && e < 2 * MM) dl[(e / M) * (M + 1) + e % M] = vd[q];
    }
    for (int e = (MT > 0 ? QD * NT : 0) + tid; e < 2 * MM; e += NT)
      dl[(e / M) * (M + 1) + e % M] = Dg[e];
    rs();
    subtree_locals<MT, true>(g, Rr, tid, NT, mres, qp, cres, qr, nullptr, nullptr, Qg, Dg, dl,
                             rs);
    // far field: F_par^T (G x s) = C_par^T (G x M) T_par^T (M x s) on
    // v_mfma_f64_16x16x4f64, two tiles (parity, 16 rows) per wave, written
    // to obf (the chain and subtree operands under it are dead: the last
    // role barrier of subtree_locals is behind every wave)
    const int ntm = s >> 4, ksteps = (M + 3) >> 2;
    constexpr int QK = MT > 0 ? (MT + 3) / 4 : 1;
    const int n = lane & 15;
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int tile = wave + j * nw, sgf = tile / ntm, rt = tile - sgf * ntm;
      d4 D = {0, 0, 0, 0};
      if (MT > 0) {
        double av[QK], bv[QK];
#pragma unroll
        for (int kk = 0; kk < QK; kk++) {
          const int k = kk * 4 + (lane >> 4);
          bv[kk] = k < M ? Tg[((size_t)sgf * s + rt * 16 + n) * M + k] : 0.0;
          av[kk] = (n < G && k < M) ? cres[(G - 1 + n) * M2 + sgf * M + k] : 0.0;
        }
#pragma unroll
        for (int kk = 0; kk < QK; kk++) D = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kk], bv[kk], D, 0, 0, 0);
      } else {
        for (int kk = 0; kk < ksteps; kk++) {
          const int k = kk * 4 + (lane >> 4);
          const double bv = k < M ? Tg[((size_t)sgf * s + rt * 16 + n) * M + k] : 0.0;
          const double av = (n < G && k < M) ? cres[(G - 1 + n) * M2 + sgf * M + k] : 0.0;
          D = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, D, 0, 0, 0);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int leaf = (lane >> 4) + 4 * v, m = rt * 16 + n;
        const int e = 2 * (leaf * s + m) + sgf;
        if (leaf < G && e < nr) obf[slot(e)] = D[v];
      }
    }
    rs();  // far field staged
    // segment B of pair nw - 1 - wave (columns s - 8p - shift onward,
    // k_down8's split moved by shift, a multiple of 8 in [-s, 8 nw]), once
    // the band role has the windows in LDS; added into obf
    const int p = nw - 1 - wave;
    double acc[8];
#pragma unroll
    for (int q = 0; q < 8; q++) acc[q] = 0.0;
    role_wait(ctr + 1, nw);
    band(p, s - 8 * p - shift, -1, acc);
    stage_group(p, acc, obf, true);
    role_arrive(ctr + 1);  // done with the windows
  } else {
    RoleSync rs{ctr + 1, nw, 0};
    // the window phase is address arithmetic and loads: at the default
    // priority its VALU instructions lose issue to the older work-group's
    // band FMAs on the same CU, and the second work-group of a CU gets its
    // windows only when the first one's band is done
    __builtin_amdgcn_s_setprio(2);
    // ---- the windows: x (both parities), the Lambda span, the head
    const int nX = 2 * ((G + 1) * s + 8);
    constexpr int QX = 9;  // ceil(nX / NT) for G = 16, s = 64
    double vx[QX];
    const int QL = 3 * s + 16, nLs = 2 * G * s + s + 16;
    constexpr int QLS = 9;
    double vl[QLS];
    const int nhd = 2 * s + 16;
#pragma unroll
    for (int q = 0; q < QX; q++) vx[q] = src[std::min(row0 + q * NT + tid, (size_t)Nl - 1) * lm.es];
#pragma unroll
    for (int q = 0; q < QLS; q++) vl[q] = tab[std::min(2 * hc0 + q * NT + tid, (size_t)N - 1)];
    double vh[3];
#pragma unroll
    for (int q = 0; q < 3; q++) vh[q] = head[q * NT + tid < nhd ? q * NT + tid : 0];
#pragma unroll
    for (int q = 0; q < QX; q++) {
      const int e = q * NT + tid;
      if (e < nX) {
        const size_t i = row0 + e;
        const double u = i < (size_t)Nl ? vx[q] : 0.0;
        const int h = e >> 1;
        xs[(e & 1) * ly.nxs + h + (h >> sh)] = (DIR == 1 && i > 0) ? u * (double)i : u;
      }
    }
    for (int e = QX * NT + tid; e < nX; e += NT) {
      const size_t i = row0 + e;
      double u = i < (size_t)Nl ? src[i * lm.es] : 0.0;
      if (DIR == 1 && i > 0) u *= (double)i;
      const int h = e >> 1;
      xs[(e & 1) * ly.nxs + h + (h >> sh)] = u;
    }
    auto put_lambda = [&](int e, double v) {
      const int c = e >> (sh + 1), q = e - (c << (sh + 1));
      if (2 * hc0 + e >= (size_t)N) v = 0.0;
      if (c < G) lw[c * ly.LW + q] = v;
      if (c > 0 && q + 2 * s < QL) lw[(c - 1) * ly.LW + q + 2 * s] = v;
    };
#pragma unroll
    for (int q = 0; q < QLS; q++) {
      const int e = q * NT + tid;
      if (e < nLs) put_lambda(e, vl[q]);
    }
    for (int e = QLS * NT + tid; e < nLs; e += NT)
      put_lambda(e, tab[std::min(2 * hc0 + e, (size_t)N - 1)]);
#pragma unroll
    for (int q = 0; q < 3; q++)
      if (q * NT + tid < nhd) hd[q * NT + tid] = vh[q];
    for (int e = 3 * NT + tid; e < nhd; e += NT) hd[e] = head[e];
    rs();

    // ---- segment A of pair wave - nw (columns [0, s - 8p - shift)), then the
    // C2L diagonal (added after the band: zero-started sums keep a single
    // set of accumulators live through the band's two code paths)
    __builtin_amdgcn_s_setprio(0);
    const int p = wave - nw;
    double acc[8];
#pragma unroll
    for (int q = 0; q < 8; q++) acc[q] = 0.0;
    band(p, 0, s - 8 * p - shift, acc);
    if (DIR == 1) {  // C2L diagonal: sqrt(pi) an_i v_i (src/C/leg2cheb.c:416)
      int ln = lane;  // a fresh copy (see stage_group)
      asm volatile("" : "+v"(ln));
      const int gl = ln & 15, sg = (ln >> 4) & 1, gam = 2 * p + (ln >> 5);
      const size_t r0 = ((size_t)Rr * G + gl) * s + 8 * gam;
      if (gl < G && r0 < (size_t)((N - sg + 1) / 2)) {
        const int rl = (int)(r0 - hc0);
        const double *xr = xs + sg * ly.nxs + rl + (rl >> sh);
        const double *pw = lw + gl * ly.LW + 16 * gam + sg;
#pragma unroll
        for (int q = 0; q < 8; q++) acc[q] += kSqrtPi * xr[q] * pw[2 * q];
      }
    }
    // windows dead once both roles' band waves are through: the staging
    // takes their place
    role_arrive(ctr + 1);
    role_wait(ctr + 1, 3 * nw);
    stage_group(p, acc, ob, false);
  }
  lds_sync();  // hardware barrier 2 of 2: the join
  __builtin_amdgcn_s_setprio(0);

  // ---- store: band plus far field, scaled
  for (int e = threadIdx.x; e < nr; e += 2 * NT) {
    const size_t i = row0 + e;
    double v = ob[slot(e)] + obf[slot(e)];
    if (DIR == 0) {
      v *= k2Pi;
      if (i == 0) v *= 0.5;
    } else {
      v *= (double)i + 0.5;
    }
    dst[i * lm.es] = v;
  }
}

// ---------------------------------------------------------------------------
// Split downward pass: k_band (the near field) and k_far (L2L and the leaf
// evaluation), each lean enough that several work-groups share a CU, so one
// work-group's loads and barriers overlap another's arithmetic.  k_band has
// no input from the tree: it runs first and writes the raw band sums (the
// C2L diagonal included) over the output; k_far then adds the far field and
// applies the final scaling (src/C/leg2cheb.c:1467-1483 and the band scaling
// of :236-535).  Both use k_down8's work-group partition (2^gd leaf chunks,
// s/16 waves), band lane layout and LDS windows.
struct BandLds {
  int xs, nxs, lw, LW, hd, total;
};
__host__ __device__ inline BandLds band_lds(const Geo &g) {
  const Down8Lds d = down8_lds(g);
  return {d.xs, d.nxs, d.lw, d.LW, d.hd, d.hd + 2 * g.s + 16};
}
size_t band_shared_bytes(const Geo &g) { return 8 * (size_t)band_lds(g).total; }

// Band sums of the work-group's rows, written over the output: raw for
// k_far, or with the final scaling when scale != 0 (the band alone,
// directM).  2 s/16 waves: wave w < s/16 runs segment A of row pair w, wave
// s/16 + w segment B of row pair s/16 - 1 - w (k_down8's split of each row
// group's columns, one segment per wave), so no wave holds two sets of sums
// and the kernel fits 128 VGPRs.
template <int DIR, int PRE>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) k_band(
    const double *__restrict__ in, double *__restrict__ out, LineMap lm, Geo g, int scale,
    const double *__restrict__ tab, const double *__restrict__ head) {
  extern __shared__ double sm[];
  const int s = g.s, N = g.N;
  const int Nl = line_len(lm, blockIdx.y, N);
  const int G = 1 << g.gd, sh = 31 - __builtin_clz(s);
  const int nw = s >> 4, NT = 128 * nw;
  const int Rr = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const size_t row0 = (size_t)Rr * G * 2 * s;
  if (row0 >= (size_t)Nl) return;  // the whole work-group
  const int nr = (int)std::min((size_t)G * 2 * s, (size_t)Nl - row0);
  const size_t hc0 = (size_t)Rr * G * s;
  const size_t off = line_off(lm, blockIdx.y);
  const double *src = in + off;
  double *dst = out + off;
  const BandLds ly = band_lds(g);
  double *xs = sm + ly.xs, *lw = sm + ly.lw, *hd = sm + ly.hd;
  double *ob = xs, *ob2 = lw;  // staging after the band (see k_down8)
  auto slot = [&](int e) { return e + (e >> (sh + 1)); };

  // ---- loads: the x window, the Lambda span, the head (one batch), at
  // raised priority: their address arithmetic would otherwise wait behind
  // the band FMAs of an older work-group on the same CU
  __builtin_amdgcn_s_setprio(2);
  const int nX = 2 * ((G + 1) * s + 8);
  constexpr int QX = 5;  // ceil(nX / NT) for G = 16
  double vx[QX];
  const int QL = 3 * s + 16, nLs = 2 * G * s + s + 16;
  constexpr int QLS = 5;
  double vl[QLS];
  const int nhd = 2 * s + 16;
#pragma unroll
  for (int q = 0; q < QX; q++) vx[q] = src[std::min(row0 + q * NT + tid, (size_t)Nl - 1) * lm.es];
#pragma unroll
  for (int q = 0; q < QLS; q++) vl[q] = tab[std::min(2 * hc0 + q * NT + tid, (size_t)N - 1)];
  const double vh = head[tid < nhd ? tid : 0];
#pragma unroll
  for (int q = 0; q < QX; q++) {
    const int e = q * NT + tid;
    if (e < nX) {
      const size_t i = row0 + e;
      const double u = i < (size_t)Nl ? vx[q] : 0.0;
      const int h = e >> 1;
      xs[(e & 1) * ly.nxs + h + (h >> sh)] = (DIR == 1 && i > 0) ? u * (double)i : u;
    }
  }
  for (int e = QX * NT + tid; e < nX; e += NT) {
    const size_t i = row0 + e;
    double u = i < (size_t)Nl ? src[i * lm.es] : 0.0;
    if (DIR == 1 && i > 0) u *= (double)i;
    const int h = e >> 1;
    xs[(e & 1) * ly.nxs + h + (h >> sh)] = u;
  }
  auto put_lambda = [&](int e, double v) {
    const int c = e >> (sh + 1), q = e - (c << (sh + 1));
    if (2 * hc0 + e >= (size_t)N) v = 0.0;
    if (c < G) lw[c * ly.LW + q] = v;
    if (c > 0 && q + 2 * s < QL) lw[(c - 1) * ly.LW + q + 2 * s] = v;
  };
#pragma unroll
  for (int q = 0; q < QLS; q++) {
    const int e = q * NT + tid;
    if (e < nLs) put_lambda(e, vl[q]);
  }
  for (int e = QLS * NT + tid; e < nLs; e += NT)
    put_lambda(e, tab[std::min(2 * hc0 + e, (size_t)N - 1)]);
  if (tid < nhd) hd[tid] = vh;
  lds_sync();
  __builtin_amdgcn_s_setprio(0);

  // ---- the band: lane (chunk gl, parity sg, del), one segment per wave
  const int gl = lane & 15, sg = (lane >> 4) & 1, del = lane >> 5;
  const int cnt = (N - sg + 1) / 2;
  const size_t beta = (size_t)Rr * G + gl;
  const size_t nLm1 = (size_t)g.nL - 1;
  const size_t cend = beta < nLm1 ? (beta + 2) * s : (size_t)cnt;
  const bool ragged = (size_t)(Rr + 1) * G > nLm1;
  const double *xsg = xs + sg * ly.nxs;
  const bool segB = wave >= nw;
  const int p = segB ? 2 * nw - 1 - wave : wave, gam = 2 * p + del, ts = s - 8 * p;
  const size_t r0 = beta * s + 8 * gam;
  const bool live = gl < G && r0 < (size_t)cnt;
  const int len = live ? (int)(cend - r0) : 0;
  const int h0 = (int)(r0 - hc0);
  const double *pw = lw + gl * ly.LW + 16 * gam + sg;
  double acc[8];
#pragma unroll
  for (int k = 0; k < 8; k++) acc[k] = 0.0;
  if (!segB) {
    if (DIR == 1 && live) {  // C2L diagonal: sqrt(pi) an_i v_i (src/C/leg2cheb.c:416)
      const double *xr = xsg + h0 + (h0 >> sh);
#pragma unroll
      for (int k = 0; k < 8; k++) acc[k] = kSqrtPi * xr[k] * pw[2 * k];
    }
    if (ragged)
      band_seg8<DIR, PRE, true, true>(xsg, pw, hd, h0, sh, 0, std::min(ts, len), len, acc);
    else
      band_seg8<DIR, PRE, false, true>(xsg, pw, hd, h0, sh, 0, ts, len, acc);
  } else {
    if (ragged)
      band_seg8<DIR, PRE, true, true>(xsg, pw, hd, h0, sh, ts, len, len, acc);
    else
      band_seg8<DIR, PRE, false, true>(xsg, pw, hd, h0, sh, ts, len, len, acc);
  }
  lds_sync();  // windows dead: the staging takes their place
  if (live) {
    double *dstg = segB ? ob2 : ob;
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (hc0 + h0 + k < (size_t)cnt) dstg[slot(2 * (h0 + k) + sg)] = acc[k];
  }
  lds_sync();
  for (int e = tid; e < nr; e += NT) {
    const size_t i = row0 + e;
    double v = ob[slot(e)] + ob2[slot(e)];
    if (scale) {
      if (DIR == 0) {
        v *= k2Pi;
        if (i == 0) v *= 0.5;
      } else {
        v *= (double)i + 0.5;
      }
    }
    dst[i * lm.es] = v;
  }
}

// k_far: L2L down the ancestor chain and the subtree (as k_down8), the
// subtree's local expansions at the leaf nodes on MFMA, added to the band
// sums k_band left in the output, scaled and stored.
struct FarLds {
  int dl, cres, mres, qp, ob, total;
};
__host__ __device__ inline FarLds far_lds(const Geo &g) {
  const int G = 1 << g.gd, s = g.s, M2 = 2 * g.M;
  FarLds d;
  d.dl = 0;  // D(0), D(1) with rows padded to M + 1
  d.cres = d.dl + g.M * M2 + M2;
  d.mres = d.cres + (2 * G - 1) * M2;
  d.qp = d.mres + (g.ld + 2 * G - 1) * M2;
  d.ob = d.qp + std::max(g.ld, 1) * M2;
  d.total = d.ob + 2 * G * s + G;  // far field, slot(e) = e + e / 2s
  return d;
}
size_t far_shared_bytes(const Geo &g) { return 8 * (size_t)far_lds(g).total; }

template <int DIR, int MT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_far(
    double *__restrict__ out, LineMap lm, Geo g, const double *__restrict__ wsall,
    size_t wsstride, const double *__restrict__ Dg, const double *__restrict__ Qg,
    const double *__restrict__ Tg) {
  extern __shared__ double sm[];
  const int M = MT > 0 ? MT : g.M;
  const int s = g.s, N = g.N, M2 = 2 * M, MM = M * M;
  const int Nl = line_len(lm, blockIdx.y, N);
  const int G = 1 << g.gd, sh = 31 - __builtin_clz(s);
  const int nw = s >> 4, NT = 64 * nw;
  const int Rr = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const size_t row0 = (size_t)Rr * G * 2 * s;
  if (row0 >= (size_t)Nl) return;
  const int nr = (int)std::min((size_t)G * 2 * s, (size_t)Nl - row0);
  const size_t off = line_off(lm, blockIdx.y);
  double *dst = out + off;
  const double *m2l = wsall + blockIdx.y * wsstride + ws_layout(g).m2l;
  const FarLds ly = far_lds(g);
  double *dl = sm + ly.dl, *cres = sm + ly.cres, *mres = sm + ly.mres, *qp = sm + ly.qp;
  double *ob = sm + ly.ob;
  auto slot = [&](int e) { return e + (e >> (sh + 1)); };

  // ---- loads: the tree operands (M2L vectors, the chain row of Q, D), the
  // band sums to add to (used last), T for the leaf evaluation
  const int nm = (g.ld + 2 * G - 1) * M2;
  constexpr int QM = 6;
  double vm[QM];
  bool okm[QM];
  const int nq = g.ld * M;
  double qr[MT > 0 ? MT : 1];
  constexpr int QD = MT > 0 ? (2 * MT * MT + 63) / 64 : 1;
  double vd[QD];
#pragma unroll
  for (int q = 0; q < QM; q++) {
    const int e = q * NT + tid;
    const size_t a = m2l_elem(g, Rr, e < nm ? e : 0, M2, okm[q]);
    okm[q] = okm[q] && e < nm;
    vm[q] = m2l[okm[q] ? a : 0];
  }
  if (MT > 0 && nq > 0) {
    const double *qrow = qrow_ptr(Qg, g, Rr, tid < nq ? tid : 0, M);
#pragma unroll
    for (int l = 0; l < (MT > 0 ? MT : 1); l++) qr[l] = qrow[l * M];
  }
#pragma unroll
  for (int q = 0; q < QD; q++) {
    const int e = q * NT + tid;
    vd[q] = Dg[e < 2 * MM ? e : 0];
  }
  __builtin_amdgcn_sched_barrier(0);
  // the band sums this work-group adds to, in flight through the tree pass
  // (nr <= 2 G s <= 8 NT)
  constexpr int QB = 8;
  double vb[QB];
#pragma unroll
  for (int q = 0; q < QB; q++) {
    const int e = q * NT + tid;
    vb[q] = dst[(row0 + (e < nr ? e : 0)) * lm.es];
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int q = 0; q < QM; q++) {
    const int e = q * NT + tid;
    if (e < nm) mres[e] = okm[q] ? vm[q] : 0.0;
  }
  for (int e = QM * NT + tid; e < nm; e += NT) {
    bool ok;
    const size_t a = m2l_elem(g, Rr, e, M2, ok);
    mres[e] = ok ? m2l[a] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < QD; q++) {
    const int e = q * NT + tid;
    if (MT > 0 && e < 2 * MM) dl[(e / M) * (M + 1) + e % M] = vd[q];
  }
  for (int e = (MT > 0 ? QD * NT : 0) + tid; e < 2 * MM; e += NT)
    dl[(e / M) * (M + 1) + e % M] = Dg[e];
  lds_sync();

  subtree_locals<MT, true>(g, Rr, tid, NT, mres, qp, cres, qr, nullptr, nullptr, Qg, Dg, dl,
                           LdsSync{});

  // ---- far field: F_par^T (G x s) = C_par^T (G x M) T_par^T (M x s) on
  // v_mfma_f64_16x16x4f64, two tiles (parity, 16 rows) per wave
  const int ntm = s >> 4, ksteps = (M + 3) >> 2;
  constexpr int QK = MT > 0 ? (MT + 3) / 4 : 1;
  const int n = lane & 15;
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int tile = wave + j * nw, sgf = tile / ntm, rt = tile - sgf * ntm;
    d4 D = {0, 0, 0, 0};
    if (MT > 0) {
      double av[QK], bv[QK];
#pragma unroll
      for (int kk = 0; kk < QK; kk++) {
        const int k = kk * 4 + (lane >> 4);
        bv[kk] = k < M ? Tg[((size_t)sgf * s + rt * 16 + n) * M + k] : 0.0;
        av[kk] = (n < G && k < M) ? cres[(G - 1 + n) * M2 + sgf * M + k] : 0.0;
      }
#pragma unroll
      for (int kk = 0; kk < QK; kk++) D = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kk], bv[kk], D, 0, 0, 0);
    } else {
      for (int kk = 0; kk < ksteps; kk++) {
        const int k = kk * 4 + (lane >> 4);
        const double bv = k < M ? Tg[((size_t)sgf * s + rt * 16 + n) * M + k] : 0.0;
        const double av = (n < G && k < M) ? cres[(G - 1 + n) * M2 + sgf * M + k] : 0.0;
        D = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, D, 0, 0, 0);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int leaf = (lane >> 4) + 4 * v, m = rt * 16 + n;
      const int e = 2 * (leaf * s + m) + sgf;
      if (leaf < G && e < nr) ob[slot(e)] = D[v];
    }
  }
  lds_sync();
  // ---- the band sums plus the far field, scaled
#pragma unroll
  for (int q = 0; q < QB; q++) {
    const int e = q * NT + tid;
    if (e >= nr) continue;
    const size_t i = row0 + e;
    double v = vb[q] + ob[slot(e)];
    if (DIR == 0) {
      v *= k2Pi;
      if (i == 0) v *= 0.5;
    } else {
      v *= (double)i + 0.5;
    }
    dst[i * lm.es] = v;
  }
}

// ---------------------------------------------------------------------------
// direct O(N^2) sums, one row per thread (src/C/leg2cheb.c:152-234).  Grid
// (row blocks, ly, lz): line = blockIdx.z * ly + blockIdx.y, so any number of
// lines runs in one launch (ly <= 65535, the grid's y limit).
__global__ void __launch_bounds__(256) k_direct(const double *__restrict__ in,
                                                double *__restrict__ out, LineMap lm, int N,
                                                int dir, const double *__restrict__ a,
                                                const double *__restrict__ an,
                                                const double *__restrict__ dn, size_t nlines) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t line = (size_t)blockIdx.z * gridDim.y + blockIdx.y;
  if (line >= nlines) return;
  N = line_len(lm, line, N);  // this line's length (tables cover the plan's N)
  if (i >= N) return;
  const size_t off = line_off(lm, line);
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

// k_direct with one wave per output row, for single transforms of larger N
// (one thread per row leaves a few dozen waves, each a serial loop of N/2
// terms): the wave's lanes take every 64th term of the row, in the same
// term order per lane, and add their partial sums across the wave.  Grid
// ((N + 3) / 4, lines): four rows per work-group, line blockIdx.y of lm
// (unpacked maps).
__global__ void __launch_bounds__(256) k_direct_rows(const double *__restrict__ in,
                                                     double *__restrict__ out, LineMap lm, int N,
                                                     int dir, const double *__restrict__ a,
                                                     const double *__restrict__ an,
                                                     const double *__restrict__ dn) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= N) return;  // whole waves: the row index is wave-uniform
  const size_t off = line_off(lm, blockIdx.y);
  const double *u = in + off;
  const size_t es = lm.es;
  // (four independent sums per lane, so that a trip's loads are in flight
  // together, measured no faster: 10.3 against 9.6 us at N = 4096; the
  // kernel is bound by its three loads per term, not by their latency)
  double acc = 0;
  if (dir == L2C) {
    for (int n = 2 * lane; n + i < N; n += 128)
      acc += a[n / 2] * k2Pi * a[n / 2 + i] * u[(size_t)(n + i) * es];
  } else {
    for (int n = 2 + 2 * lane; n + i < N; n += 128)
      acc -= dn[n / 2] * an[n / 2 + i] * (u[(size_t)(n + i) * es] * (double)(n + i));
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
  if (lane == 0) {
    if (dir == L2C) {
      if (i == 0) acc /= 2;
    } else {
      const double vi = i == 0 ? u[0] : u[(size_t)i * es] * (double)i;
      acc = (kSqrtPi * vi * an[i] + acc) * ((double)i + 0.5);
    }
    out[off + (size_t)i * es] = acc;
  }
}

// ---------------------------------------------------------------------------
// Direct sums of single transforms (and of line calls of a few lines, one
// line per grid.z) of N above 3072 with the band's register rings
// (src/C/leg2cheb.c:152-234).  Per parity sg the rows r and columns m
// (element 2m + sg) run over [0, R), R = (N - sg + 1) / 2, and
//   y_r = sum_{m >= r} H(m - r) P(m + r + sg) X(m)
// with H = P = Lambda, X = u (L2C), or H = dn (H(0) = 0: the diagonal is
// added after the sum), P = an, X(m) = u_n n (C2L).  A work-group of four
// waves owns 8 G rows of one parity (G groups of 8 rows, 4 / G waves each)
// and stages X, P and H for the columns its rows touch in LDS, in chunks of
// kDbChunk columns (one chunk up to N = 8192); per chunk, lane j of
// the wave for part c of group g takes the columns t = m - r0 in
// [(64c + j) T, (64c + j + 1) T) of the group's 8 rows, T a multiple of 8
// with T/8 odd; the 64 lanes' sums are added across the wave, then the
// parts through LDS.  Every window stores index e at e + e/8, so the 64
// lanes' addresses (64 segments of T/8 padded blocks apart) fall in distinct
// banks.  G = 1 (8 rows, four parts) while two work-groups' windows fit a
// CU, else G = 4 (32 rows, one part).  Kernel trace (us), G = 1 / 2 / 4: N =
// 3500 6.4 / 6.4 / 7.0, 4096 7.1 / 7.4 / 7.7, 6000 15.6 / 10.2 / 9.8, 8192
// 20.9 / 12.7 / 11.0 (from 6000 on one work-group per CU: G = 1 then takes
// two rounds, and more rows per work-group load the windows fewer times).
constexpr int kDirectBandMinN = 3072;
// columns per LDS window load: a work-group whose rows touch more columns
// runs them in chunks of this many, the windows reloaded for each
constexpr int kDbChunk = 4096;
__host__ __device__ inline int dband_win(int cols) { return cols + 128; }
__host__ __device__ inline int dband_pad(int e) { return e + (e >> 3); }
size_t dband_shared_bytes(int N) {
  const int w = dband_pad(dband_win(std::min((N + 1) / 2, kDbChunk))) + 1;
  return 8 * (size_t)(3 * w + 24);  // windows, the other parts' row sums (G = 1)
}
// G = 1 while two work-groups' windows fit one CU's 160 KB of LDS, else 4
inline int dband_groups(int N) { return dband_shared_bytes(N) <= 80 * 1024 ? 1 : 4; }

// Columns [t0, t1) of the lane's 8 rows, t0 a multiple of 8; xq, pq, hq: the
// padded blocks of X(r0 + t0), P(2 r0 + sg + t0), H(t0), the block before
// hq holding H(t0 - 8 .. t0 - 1) (zero below 0).  Columns t >= len read X as
// zero.  Rings as in band_seg8: slot t mod 8 holds H(t - 8 .. t)
// and P(t .. t + 7); row rho of column t takes H(t - rho) P(t + rho) X(t).
__device__ __forceinline__ void direct_seg8(const double *__restrict__ xq,
                                            const double *__restrict__ pq,
                                            const double *__restrict__ hq, int t0, int t1,
                                            int len, double acc[8]) {
  const int ntrip = t1 > t0 ? (t1 - t0 + 7) >> 3 : 0;
  if (ntrip == 0) return;
  double hr[8], pr[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    hr[k] = hq[k - 9];
    pr[k] = pq[k];
  }
  int tt = t0;
  double XA[4], PA[4], HA[4], XB[4], PB[4], HB[4];
#define DFETCH(X, P, H, HALF)                                       \
  _Pragma("unroll") for (int k_ = 0; k_ < 4; k_++) {                \
    X[k_] = tt + 4 * (HALF) + k_ < len ? xq[4 * (HALF) + k_] : 0.0; \
    P[k_] = pq[9 + 4 * (HALF) + k_];                                \
    H[k_] = hq[4 * (HALF) + k_];                                    \
  }
#define DSTEP(K, X, P, H, J)                                          \
  {                                                                   \
    hr[K] = H[J];                                                     \
    const double x_ = X[J];                                           \
    _Pragma("unroll") for (int r_ = 0; r_ < 8; r_++)                  \
      acc[r_] += (hr[((K) - r_) & 7] * pr[((K) + r_) & 7]) * x_;      \
    pr[K] = P[J];                                                     \
  }
  DFETCH(XA, PA, HA, 0)
  for (int b = 0; b < ntrip; b++) {
    DFETCH(XB, PB, HB, 1)
    DSTEP(0, XA, PA, HA, 0) DSTEP(1, XA, PA, HA, 1) DSTEP(2, XA, PA, HA, 2) DSTEP(3, XA, PA, HA, 3)
    xq += 9;
    pq += 9;
    hq += 9;
    tt += 8;
    DFETCH(XA, PA, HA, 0)
    DSTEP(4, XB, PB, HB, 0) DSTEP(5, XB, PB, HB, 1) DSTEP(6, XB, PB, HB, 2) DSTEP(7, XB, PB, HB, 3)
  }
#undef DSTEP
#undef DFETCH
}

// Grid ((R0 + 8 G - 1) / (8 G), 2, lines) with R0 = (N + 1) / 2, 256
// threads, dband_shared_bytes(N) of LDS; line blockIdx.z of lm (unpacked
// maps).
template <int DIR, int G>
__global__ void __launch_bounds__(256) k_direct_band(const double *__restrict__ in,
                                                     double *__restrict__ out, LineMap lm, int N,
                                                     const double *__restrict__ a,
                                                     const double *__restrict__ an,
                                                     const double *__restrict__ dn) {
  extern __shared__ double sm[];
  const int sg = blockIdx.y;
  const int R = (N - sg + 1) / 2;
  static_assert(G == 1 || G == 4, "row groups per work-group");
  constexpr int NP = 4 / G;  // parts of each group's columns
  const int Rb = blockIdx.x * 8 * G;
  if (Rb >= R) return;  // the whole work-group
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int span = R - Rb;  // columns Rb .. R - 1
  const int nch = (span + kDbChunk - 1) / kDbChunk;  // column chunks
  const int W = dband_win(std::min(span, kDbChunk)), WP = dband_pad(W) + 1;
  double *xw = sm, *pw = sm + WP, *hw = sm + 2 * WP;
  const size_t off = line_off(lm, blockIdx.z);
  const double *u = in + off;
  const size_t es = lm.es;
  const double *Ht = DIR == 0 ? a : dn;
  const double *Pt = DIR == 0 ? a : an;
  const int nH = DIR == 0 ? N : (N + 1) / 2;  // entries of the H table
  const int g = wave % G, c = wave / G;  // row group, column part
  const int r0 = Rb + 8 * g;
  const bool live = r0 < R;  // (every wave reaches the barriers below)
  const int len = live ? R - r0 : 0;
  double acc[8];
#pragma unroll
  for (int k = 0; k < 8; k++) acc[k] = 0.0;
  for (int ch = 0; ch < nch; ch++) {
    // columns t = m - r0 in [tb, tb + kDbChunk): windows X(Rb + tb + e),
    // P(2 Rb + sg + tb + e), H(tb - 8 + e), e < W, in batches of QF
    // entries per thread, the batch's loads issued before its stores (one
    // entry per trip: 8.6 against 8.2 us at N = 4096; QF = 16 takes 174
    // VGPRs and was slower still)
    const int tb = ch * kDbChunk;
    constexpr int QF = 8;
    for (int e0 = tid; e0 < W; e0 += QF * 256) {
      double vx[QF], vp[QF], vh[QF];
#pragma unroll
      for (int k = 0; k < QF; k++) {
        const int e = e0 + k * 256;
        const int m = Rb + tb + e, qi = 2 * Rb + sg + tb + e, hk = tb - 8 + e;
        vx[k] = m < R ? u[(size_t)(2 * m + sg) * es] : 0.0;
        vp[k] = qi < N ? Pt[qi] : 0.0;
        vh[k] = (hk >= 0 && hk < span && hk < nH && !(DIR == 1 && hk == 0)) ? Ht[hk] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < QF; k++) {
        const int e = e0 + k * 256;
        if (e < W) {
          const int pe = dband_pad(e);
          xw[pe] = DIR == 1 ? vx[k] * (double)(2 * (Rb + tb + e) + sg) : vx[k];
          pw[pe] = vp[k];
          hw[pe] = vh[k];
        }
      }
    }
    __syncthreads();
    const int cw = std::min(kDbChunk, span - tb);
    int T8 = (cw + 512 * NP - 1) / (512 * NP);
    T8 |= 1;
    const int T = 8 * T8;
    const int t0 = tb + (64 * c + lane) * T;
    const int t1 = std::min(std::min(t0 + T, tb + kDbChunk), len);
    if (t0 < t1)  // false for every lane of a wave past the last row
      direct_seg8(xw + dband_pad(8 * g + t0 - tb), pw + dband_pad(16 * g + t0 - tb),
                  hw + dband_pad(t0 - tb + 8), t0, t1, len, acc);
    if (ch + 1 < nch) __syncthreads();  // the next chunk's loads overwrite the windows
  }
  // the 64 lanes' sums of the 8 rows, halving the rows at each of the
  // first three exchanges (lanes with bit 5 set keep rows 4-7, bit 4 the
  // upper two of those four, bit 3 the upper one): 10 exchanges instead of
  // 48, and lane l ends with row l >> 3
  const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
  double k4[4], k2[2];
#pragma unroll
  for (int j = 0; j < 4; j++)
    k4[j] = (b5 ? acc[j + 4] : acc[j]) + __shfl_xor(b5 ? acc[j] : acc[j + 4], 32, 64);
#pragma unroll
  for (int j = 0; j < 2; j++)
    k2[j] = (b4 ? k4[j + 2] : k4[j]) + __shfl_xor(b4 ? k4[j] : k4[j + 2], 16, 64);
  double v = (b3 ? k2[1] : k2[0]) + __shfl_xor(b3 ? k2[0] : k2[1], 8, 64);
#pragma unroll
  for (int d = 4; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if constexpr (NP > 1) {
    double *red = sm + 3 * WP;  // [part - 1][group][row]: the other parts' sums
    if (c > 0 && (lane & 7) == 0) red[(c - 1) * 8 * G + 8 * g + (lane >> 3)] = v;
    __syncthreads();
    if (c == 0 && (lane & 7) == 0) {
#pragma unroll
      for (int q = 1; q < NP; q++) v += red[(q - 1) * 8 * G + 8 * g + (lane >> 3)];
    }
  }
  if (c != 0 || !live) return;
  const int r = r0 + (lane >> 3);
  if ((lane & 7) == 0 && r < R) {
    const int i = 2 * r + sg;
    if (DIR == 0) {
      v *= k2Pi;
      if (i == 0) v *= 0.5;
    } else {
      const double vi = i == 0 ? u[0] : u[(size_t)i * es] * (double)i;
      v = (kSqrtPi * vi * an[i] - v) * ((double)i + 0.5);
    }
    out[off + (size_t)i * es] = v;
  }
}

// ---------------------------------------------------------------------------
// Batched lines of the direct path (N <= kDenseMaxN) as dense products on FP64
// MFMA: per parity p, OUT_p (NPP x LT) = K_p (NPP x NPP) IN_p (NPP x LT) for a
// tile of LT lines, K_p upper triangular (the direct sums of
// src/C/leg2cheb.c:152-234 with scaling and C2L weights folded in).
//
// The two parities run one after the other through one LDS tile [c][line]
// (row stride kDenseLTP: 36 KB at NPP = 128) and the kernel fits 128 VGPRs,
// so four work-groups share a CU and one's loads overlap another's products.
// v_mfma_f64_16x16x4f64 with A =
// K_p from global memory (the plan's table, L2-resident, in fragment order,
// streamed kDenseKG k-steps ahead) and B = the input rows from LDS.  Wave w
// owns row tiles w and NR-1-w, so every wave runs the same number of k-steps;
// results go back through the LDS tile so that the stores keep the loads'
// coalescing.
//
// CONTIG (lines with unit element stride): lanes run along element pairs
// (2c, 2c+1), loaded and stored as one 16-byte access; the odd parity's
// inputs wait in registers during the even parity's products, and the even
// parity's outputs wait in registers during the odd one's.  Strided lines
// (lanes along the lines) load and store one parity at a time.
constexpr int kDenseLT = 32, kDenseLTP = 36, kDenseNT = 256, kDenseKG = 4;
template <bool CONTIG>
__global__ void __launch_bounds__(kDenseNT) __attribute__((amdgpu_waves_per_eu(4)))
    k_lines_dense(const double *__restrict__ in, double *__restrict__ out, LineMap lm,
                  size_t nlines, int N, int NPP, const double *__restrict__ K) {
  extern __shared__ double sm[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const size_t l0 = (size_t)blockIdx.x * kDenseLT;
  const int LTP = kDenseLTP;
  constexpr bool contig = CONTIG;
  // start and length of the tile's lines, computed once (line_off divides)
  size_t *toff = (size_t *)(sm + NPP * LTP);
  int *tlen = (int *)(toff + kDenseLT);
  if (tid < kDenseLT) {
    const size_t line = l0 + tid;
    toff[tid] = line < nlines ? line_off(lm, line) : 0;
    tlen[tid] = line < nlines ? line_len(lm, line, N) : 0;
  }
  __syncthreads();
  // element q < 16 of this thread in one parity's tile (c < NPP <= 128, 32
  // lines): lanes run along c when the lines are contiguous, along the lines
  // (two rows c per wave) otherwise, so neighbouring lanes touch neighbouring
  // words of one parity
  auto tile_c = [&](int q) { return contig ? lane + 64 * (q & 1) : 2 * wave + (lane >> 5) + 8 * q; };
  auto tile_l = [&](int q) { return contig ? wave + 4 * (q >> 1) : (lane & 31); };
  // strided lines: one line per thread, rows c0 + 8 q, so one base address
  // and a constant step (16 elements) serve all 16 elements
  const int c0 = 2 * wave + (lane >> 5), ls = lane & 31;
  auto load = [&](int p, double *vt) {
    const double *src = in + toff[ls] + (size_t)(2 * c0 + p) * lm.es;
    const size_t step = 16 * lm.es;
    const int len = tlen[ls];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int c = c0 + 8 * q;
      vt[q] = c < NPP && 2 * c + p < len ? src[q * step] : 0.0;
    }
  };
  const int NR = NPP >> 4, npair = (NR + 1) >> 1;
  const bool active = wave < npair;
  const int R0 = wave, R1 = NR - 1 - wave;
  const bool two = active && R1 != R0;
  const int nk0 = active ? (NPP - 16 * R0) >> 2 : 0, nk1 = two ? (NPP - 16 * R1) >> 2 : 0;
  const double *Bp = sm + (lane >> 4) * LTP + (lane & 15);
  // contiguous lines: element pair (2c, 2c+1) of a line as one 16-byte
  // access when it is aligned and whole, else two guarded words
  auto pair_at = [&](int q, double &e0, double &e1) {
    const int c = tile_c(q), l = tile_l(q), len = tlen[l];
    const size_t a = toff[l] + 2 * (size_t)c;
    e0 = e1 = 0.0;
    if (c >= NPP) return;
    if (2 * c + 1 < len && !(a & 1)) {
      const double2 t = *reinterpret_cast<const double2 *>(in + a);
      e0 = t.x;
      e1 = t.y;
    } else {
      if (2 * c < len) e0 = in[a];
      if (2 * c + 1 < len) e1 = in[a + 1];
    }
  };
  double vt[16], hold[16];
  if (CONTIG) {
#pragma unroll
    for (int q = 0; q < 16; q++) pair_at(q, hold[q], vt[q]);  // hold: parity 0 for now
  } else {
    load(0, vt);
  }
  for (int p = 0; p < 2; p++) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int c = tile_c(q);
      if (c < NPP) sm[c * LTP + tile_l(q)] = CONTIG && p == 0 ? hold[q] : vt[q];
    }
    __syncthreads();
    d4 D00 = {0.0, 0.0, 0.0, 0.0}, D01 = D00, D10 = D00, D11 = D00;
    if (active) {
      // A fragments of row tile R from its diagonal k-step on: 64 doubles each
      const double *K0 = K + (((size_t)p * NR + R0) * (NPP >> 2) + 4 * R0) * 64 + lane;
      const double *K1 = K + (((size_t)p * NR + R1) * (NPP >> 2) + 4 * R1) * 64 + lane;
      constexpr int G = kDenseKG, NG = 32 / G;  // nk0 <= 32, nk1 <= 16
      double a0[2][G], a1[2][G];
#pragma unroll
      for (int kk = 0; kk < G; kk++) {
        a0[0][kk] = kk < nk0 ? K0[64 * kk] : 0.0;
        a1[0][kk] = kk < nk1 ? K1[64 * kk] : 0.0;
      }
#pragma unroll
      for (int gi = 0; gi < NG; gi++) {
        const int cur = gi & 1;
        if (gi + 1 < NG) {
#pragma unroll
          for (int kk = 0; kk < G; kk++) {
            const int k = G * (gi + 1) + kk;
            a0[cur ^ 1][kk] = k < nk0 ? K0[64 * k] : 0.0;
            if (k < 16) a1[cur ^ 1][kk] = k < nk1 ? K1[64 * k] : 0.0;
          }
        }
#pragma unroll
        for (int kk = 0; kk < G; kk++) {
          const int k = G * gi + kk;
          if (k < nk0) {
            const double *b = Bp + (16 * R0 + 4 * k) * LTP;
            D00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[cur][kk], b[0], D00, 0, 0, 0);
            D01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[cur][kk], b[16], D01, 0, 0, 0);
          }
          if (k < 16 && k < nk1) {
            const double *b = Bp + (16 * R1 + 4 * k) * LTP;
            D10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[cur][kk], b[0], D10, 0, 0, 0);
            D11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[cur][kk], b[16], D11, 0, 0, 0);
          }
        }
      }
    }
    __syncthreads();  // every wave is done with the input tile
    if (active) {
      double *o = sm + (16 * R0 + (lane >> 4)) * LTP + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; v++) {
        o[4 * v * LTP] = D00[v];
        o[4 * v * LTP + 16] = D01[v];
      }
      if (two) {
        o = sm + (16 * R1 + (lane >> 4)) * LTP + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; v++) {
          o[4 * v * LTP] = D10[v];
          o[4 * v * LTP + 16] = D11[v];
        }
      }
    }
    __syncthreads();
    if (CONTIG) {
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int c = tile_c(q), l = tile_l(q);
        const double v = c < NPP ? sm[c * LTP + l] : 0.0;
        if (p == 0) {
          hold[q] = v;  // even outputs, stored with the odd ones
          continue;
        }
        const int len = tlen[l];
        const size_t a = toff[l] + 2 * (size_t)c;
        if (c >= NPP || 2 * c >= len) continue;
        if (2 * c + 1 < len && !(a & 1)) {
          *reinterpret_cast<double2 *>(out + a) = make_double2(hold[q], v);
        } else {
          out[a] = hold[q];
          if (2 * c + 1 < len) out[a + 1] = v;
        }
      }
    } else {
      double *dst = out + toff[ls] + (size_t)(2 * c0 + p) * lm.es;
      const size_t step = 16 * lm.es;
      const int len = tlen[ls];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int c = c0 + 8 * q;
        if (c < NPP && 2 * c + p < len) dst[q * step] = sm[c * LTP + ls];
      }
    }
    __syncthreads();  // the tile is free for the next parity
    if (!CONTIG && p == 0) load(1, vt);
  }
}

// the tile [NPP][LTP], then the lines' offsets (size_t) and lengths (int)
static size_t dense_shared_bytes(int NPP) {
  return 8 * (size_t)NPP * kDenseLTP + kDenseLT * (sizeof(size_t) + sizeof(int));
}

// Batched lines of the direct path with kDenseMaxN < N <= kGemmMaxN: the same
// parity blocks K_p (NPP up to 2048) as a tiled product on FP64 MFMA.  A
// work-group owns 64 pair-rows (row tiles 4 rt + wave, one per wave) of one
// parity p for 32 lines and walks the columns c >= 64 rt in chunks of 64:
// the chunk's inputs 2c + p of its 32 lines are staged in LDS [c][line], then
// each wave runs v_mfma_f64_16x16x4f64 with A = K_p fragments (global
// memory, L2/MALL-resident) and B = the staged rows, two 16-line column
// tiles.  The next chunk's inputs and fragments load into registers during
// the current chunk's products.  The results go out through the same LDS
// tile.  The two parities are separate work-groups: twice as many, each
// with half the registers and LDS of one holding both, and row tile 0 (which
// walks all NPP / 64 chunks) splits its chain in two.  Grid (2 x line
// tiles, row tiles), the parity in the lowest bit of blockIdx.x:
// the long row tiles of both parities are dispatched first, and consecutive
// work-groups share a row tile's K rows.
// LDS row stride 33 doubles: the staging writes (lanes along c) and the
// result reads step 66 dwords, two banks apart, while each 16-lane row of an
// MFMA operand read stays contiguous
constexpr int kGemmRT = 64, kGemmLT = 32, kGemmKC = 64, kGemmNT = 256, kGemmLTP = 33;
constexpr size_t kGemmFewLines = 16384;
// The work of tile (bx, by) of the tiled dense product (grid (2 x line tiles,
// row tiles)); k_lines_gemm runs one grid of them, k_lines_gemm_multi the
// grids of several size classes in one launch.
template <bool CONTIG>
__device__ __forceinline__ void lines_gemm_tile(const double *__restrict__ in,
                                                double *__restrict__ out, const LineMap &lm,
                                                size_t nlines, int N, int NPP,
                                                const double *__restrict__ K, unsigned bx,
                                                unsigned by, double *sm) {
  constexpr int LTP = kGemmLTP, KC = kGemmKC, LT = kGemmLT;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int rt = (int)by, p = bx & 1;
  const size_t l0 = (size_t)(bx >> 1) * LT;
  size_t *toff = (size_t *)(sm + KC * LTP);
  int *tlen = (int *)(toff + LT);
  if (tid < LT) {
    const size_t line = l0 + tid;
    toff[tid] = line < nlines ? line_off(lm, line) : 0;
    tlen[tid] = line < nlines ? line_len(lm, line, N) : 0;
  }
  __syncthreads();
  const int NR = NPP >> 4, nkt = NPP >> 2, R = 4 * rt + wave;
  const bool active = R < NR;  // the last row tile may be partly past NPP
  const size_t es = CONTIG ? 1 : lm.es;
  const d4 z = {0.0, 0.0, 0.0, 0.0};
  d4 D0 = z, D1 = z;  // 16-line halves
  const double *Bp = sm + (lane >> 4) * LTP + (lane & 15);
  const double *Kp = K + ((size_t)(p * NR + R) * nkt) * 64 + lane;
  constexpr int QS = KC * LT / kGemmNT;  // staged inputs per thread per chunk
  static_assert(KC == 64 && kGemmRT == 64 && kGemmNT == 256, "thread q-th line is 4 q + wave");
  // this thread's lines in every chunk and in the results: 4 q + wave (q <
  // QS), column c0 + lane (row 64 rt + lane); their starts and lengths are
  // wave-uniform and kept in scalar registers
  const double *src[QS];
  double *dst[QS];
  int len[QS];
#pragma unroll
  for (int q = 0; q < QS; q++) {
    const size_t o = toff[4 * q + wave];
    const size_t ou = (size_t)(unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)o) |
                      ((size_t)(unsigned)__builtin_amdgcn_readfirstlane((int)(o >> 32)) << 32);
    len[q] = __builtin_amdgcn_readfirstlane(tlen[4 * q + wave]);
    src[q] = in + ou + p * es;
    dst[q] = out + ou + p * es;
  }
  // chunk c0's inputs of this thread and this wave's fragments, into registers
  double e[QS], an[KC / 4];
  auto fetch = [&](int c0) {
    const int ks0 = c0 == kGemmRT * rt ? 4 * wave : 0, nks = min(KC, NPP - c0) >> 2;
    if (active && ks0 == 0 && nks == KC / 4) {
#pragma unroll
      for (int ks = 0; ks < KC / 4; ks++) an[ks] = Kp[((c0 >> 2) + ks) * 64];
    } else {
#pragma unroll
      for (int ks = 0; ks < KC / 4; ks++)
        an[ks] = active && ks >= ks0 && ks < nks ? Kp[((c0 >> 2) + ks) * 64] : 0.0;
    }
    const int c = c0 + lane;
#pragma unroll
    for (int q = 0; q < QS; q++) e[q] = 2 * c + p < len[q] ? src[q][(size_t)(2 * c) * es] : 0.0;
  };
  fetch(kGemmRT * rt);
  for (int c0 = kGemmRT * rt; c0 < NPP; c0 += KC) {
    // k-steps of this chunk for this wave: from its diagonal in the first
    // chunk (K_p is zero below it), up to NPP in the last
    const int ks0 = c0 == kGemmRT * rt ? 4 * wave : 0;
    const int nks = min(KC, NPP - c0) >> 2;
    double a[KC / 4];
#pragma unroll
    for (int q = 0; q < QS; q++) sm[lane * LTP + 4 * q + wave] = e[q];
#pragma unroll
    for (int ks = 0; ks < KC / 4; ks++) a[ks] = an[ks];
    __syncthreads();
    // the next chunk's loads run during this chunk's products
    if (c0 + KC < NPP) fetch(c0 + KC);
    if (active && ks0 == 0 && nks == KC / 4) {
#pragma unroll
      for (int ks = 0; ks < KC / 4; ks++) {
        const double *b = Bp + 4 * ks * LTP;
        D0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ks], b[0], D0, 0, 0, 0);
        D1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ks], b[16], D1, 0, 0, 0);
      }
    } else if (active) {
#pragma unroll
      for (int ks = 0; ks < KC / 4; ks++) {
        if (ks < ks0 || ks >= nks) continue;
        const double *b = Bp + 4 * ks * LTP;
        D0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ks], b[0], D0, 0, 0, 0);
        D1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ks], b[16], D1, 0, 0, 0);
      }
    }
    __syncthreads();  // the tile is free for the next chunk
  }
  // results: row 16 wave + lane/16 + 4 v of the tile, line 16 h + lane % 16
  {
    double *o = sm + (16 * wave + (lane >> 4)) * LTP + (lane & 15);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      o[4 * v * LTP] = D0[v];
      o[4 * v * LTP + 16] = D1[v];
    }
  }
  __syncthreads();
  const int r = kGemmRT * rt + lane;
#pragma unroll
  for (int q = 0; q < QS; q++)
    if (r < NPP && 2 * r + p < len[q]) dst[q][(size_t)(2 * r) * es] = sm[lane * LTP + 4 * q + wave];
}

template <bool CONTIG>
__global__ void __launch_bounds__(kGemmNT) __attribute__((amdgpu_waves_per_eu(3)))
    k_lines_gemm(const double *__restrict__ in, double *__restrict__ out, LineMap lm,
                 size_t nlines, int N, int NPP, const double *__restrict__ K) {
  extern __shared__ double sm[];
  lines_gemm_tile<CONTIG>(in, out, lm, nlines, N, NPP, K, blockIdx.x, blockIdx.y, sm);
}

// Several size classes of an uneven batch (packed, unit-stride lines) on the
// tiled dense product in one launch: a flat grid over every class's tiles,
// classes in the order given, each class's tiles row tile by row tile (the
// long row tiles first, as in k_lines_gemm's grid).
struct DenseClass {
  const double *K;
  const size_t *off;
  const int *n;
  size_t nlines;
  int N, NPP;
  unsigned bxs, tiles;  // 2 x line tiles; bxs x row tiles
};
constexpr int kMaxDenseClasses = 8;
struct DenseClasses {
  DenseClass c[kMaxDenseClasses];
  int count;
};
__global__ void __launch_bounds__(kGemmNT) __attribute__((amdgpu_waves_per_eu(3)))
    k_lines_gemm_multi(const double *__restrict__ in, double *__restrict__ out, DenseClasses dc) {
  extern __shared__ double sm[];
  unsigned t = blockIdx.x;
  int c = 0;
  while (c + 1 < dc.count && t >= dc.c[c].tiles) t -= dc.c[c++].tiles;
  const DenseClass &d = dc.c[c];
  if (t >= d.tiles) return;
  LineMap lm{1, 0, 0, 1, d.off, d.n};
  lines_gemm_tile<true>(in, out, lm, d.nlines, d.N, d.NPP, d.K, t % d.bxs, t / d.bxs, sm);
}

size_t gemm_shared_bytes() {
  return 8 * (size_t)kGemmKC * kGemmLTP + kGemmLT * (sizeof(size_t) + sizeof(int));
}

// One transform by direct sums from the Lambda tables a, an, dn (those of a
// direct plan of size N, or of an FMM plan): k_direct_band for N above
// kDirectBandMinN unless band is false (the plan's direct_band), else
// k_direct_rows.  Kernel trace, L2C, k_direct_band /
// k_direct_rows: N = 1024 5.8 / 3.8 us, 4096 8.2 / 9.7 us, 8192 12.2 /
// 27.8 us (below about 3000 the band kernel's window loads and cross-lane
// sums cost more than the row kernel's loads save).
static void launch_direct_one(Prof *prof, int N, size_t dir, const double *a, const double *an,
                              const double *dn, const double *d_in, double *d_out,
                              const LineMap &lm, hipStream_t st, bool band, unsigned nlines = 1) {
  if (N > kDirectBandMinN && band) {
    const int G = dband_groups(N), rows = 8 * G;
    const dim3 grid((unsigned)(((N + 1) / 2 + rows - 1) / rows), 2, nlines);
    const size_t shm = dband_shared_bytes(N);
    if (G == 1) {
      if (dir == L2C)
        launch_k(prof, k_direct_band<0, 1>, grid, 256, shm, st, d_in, d_out, lm, N, a, an, dn);
      else
        launch_k(prof, k_direct_band<1, 1>, grid, 256, shm, st, d_in, d_out, lm, N, a, an, dn);
    } else {
      if (dir == L2C)
        launch_k(prof, k_direct_band<0, 4>, grid, 256, shm, st, d_in, d_out, lm, N, a, an, dn);
      else
        launch_k(prof, k_direct_band<1, 4>, grid, 256, shm, st, d_in, d_out, lm, N, a, an, dn);
    }
    return;
  }
  launch_k(prof, k_direct_rows, dim3((unsigned)((N + 3) / 4), nlines), 256, 0, st, d_in, d_out,
           lm, N, (int)dir, a, an, dn);
}

// k_direct over nlines lines of map lm (one launch for any number of lines);
// single transforms from kDirectRowsMinN on take launch_direct_one
constexpr int kDirectRowsMinN = 1024;
static void launch_direct(const direct_plan_s *dp, const double *d_in, double *d_out, size_t dir,
                          const LineMap &lm, size_t nlines, hipStream_t st) {
  const int N = (int)dp->N;
  if (nlines == 1 && N >= kDirectRowsMinN) {
    launch_direct_one(dp->prof, N, dir, dp->d_a, dp->d_an, dp->d_dn, d_in, d_out, lm, st,
                      dp->direct_band);
    return;
  }
  const unsigned gy = (unsigned)std::min<size_t>(nlines, dp->lines_y);
  const unsigned gz = (unsigned)((nlines + gy - 1) / gy);
  launch_k(dp->prof, k_direct, dim3((N + 255) / 256, gy, gz), 256, 0, st, d_in, d_out, lm, N,
      (int)dir, dp->d_a, dp->d_an, dp->d_dn, nlines);
}

// Whether a call of nlines lines through direct plan dp takes the tiled dense
// product (k_lines_gemm): the one-tile kernel has the lower cost per line,
// the tiled one far more work-groups for a given call, so above N = 32 the
// tiled kernel takes calls of up to kGemmFewLines lines (N = 256: 600 lines
// in 7 us against 18 us; 65,536 lines in 119 us against 90 us) unless the
// plan's dense_tile says otherwise.
bool dense_tiled(const direct_plan_s *dp, size_t dir, size_t nlines) {
  const int N = (int)dp->N;
  return dp->d_K[dir] && nlines >= 4 &&
         (N > kDenseMaxN || (dp->dense_tile == 2 && N > 32) ||
          (dp->dense_tile == 0 && N > 32 && nlines <= kGemmFewLines));
}

// Packed unit-stride lines of several direct plans in one tiled dense launch.
int run_dense_classes(const DenseItem *items, int count, const double *d_in, double *d_out,
                      size_t dir, hipStream_t st) {
  DenseClasses dc{};
  size_t total = 0;
  for (int i = 0; i < count; i++) {
    if (i >= kMaxDenseClasses) {
      set_error("run_dense_classes: too many classes for one launch");
      return -1;
    }
    const direct_plan_s *dp = items[i].dp;
    DenseClass &c = dc.c[i];
    c.K = dp->d_K[dir];
    c.off = items[i].off;
    c.n = items[i].n;
    c.nlines = items[i].nlines;
    c.N = (int)dp->N;
    c.NPP = dp->NPP;
    c.bxs = (unsigned)(2 * ((c.nlines + kGemmLT - 1) / kGemmLT));
    c.tiles = c.bxs * (unsigned)((dp->NPP + kGemmRT - 1) / kGemmRT);
    total += c.tiles;
  }
  dc.count = count;
  if (total == 0) return 0;
  k_lines_gemm_multi<<<(unsigned)total, kGemmNT, gemm_shared_bytes(), st>>>(d_in, d_out, dc);
  L2C_HIP(hipGetLastError(), -1);
  return 0;
}

// Batched lines through a direct plan: dense MFMA products when the plan has
// them and the batch is large enough, else one row per thread.
static int launch_direct_lines(const direct_plan_s *dp, const double *d_in, double *d_out,
                               size_t dir, const LineMap &lm, size_t nlines, hipStream_t st) {
  const int N = (int)dp->N;
  if (dense_tiled(dp, dir, nlines)) {
    const size_t nb = (nlines + kGemmLT - 1) / kGemmLT;
    const dim3 grid((unsigned)(2 * nb), (unsigned)((dp->NPP + kGemmRT - 1) / kGemmRT));
    if (lm.es == 1)
      launch_k(dp->prof, k_lines_gemm<true>, grid, kGemmNT, gemm_shared_bytes(), st, d_in, d_out,
          lm, nlines, N, dp->NPP, dp->d_K[dir]);
    else
      launch_k(dp->prof, k_lines_gemm<false>, grid, kGemmNT, gemm_shared_bytes(), st, d_in, d_out,
          lm, nlines, N, dp->NPP, dp->d_K[dir]);
  } else if (dp->d_K[dir] && nlines >= 4) {
    const size_t nb = (nlines + kDenseLT - 1) / kDenseLT;
    if (lm.es == 1)
      launch_k(dp->prof, k_lines_dense<true>, dim3((unsigned)nb), kDenseNT,
          dense_shared_bytes(dp->NPP), st, d_in, d_out, lm, nlines, N, dp->NPP, dp->d_K[dir]);
    else
      launch_k(dp->prof, k_lines_dense<false>, dim3((unsigned)nb), kDenseNT,
          dense_shared_bytes(dp->NPP), st, d_in, d_out, lm, nlines, N, dp->NPP, dp->d_K[dir]);
  } else {
    launch_direct(dp, d_in, d_out, dir, lm, nlines, st);
  }
  L2C_HIP(hipGetLastError(), -1);
  return 0;
}

// ---------------------------------------------------------------------------
// host orchestration

// Whether a call of nlines lines runs its M2L below lu on MFMA (precision
// form 2, or FP64 MFMA for line calls from m2l_mma_lines lines on).
static bool m2l_on_mfma(const fmm_plan_s *p, size_t nlines) {
  return p->geo.M == 18 && (p->precision == 2 ||
                            (p->m2l_mma_lines > 0 && nlines >= (size_t)p->m2l_mma_lines));
}
// The work-group geometry of a call: single-line calls take geo1, line
// calls geo (the MFMA M2L forms, whose form-2 tables are cut at geo's lu,
// always take geo).
static const Geo &call_geo(const fmm_plan_s *p, size_t nlines) {
  return nlines == 1 && !m2l_on_mfma(p, nlines) ? p->geo1 : p->geo;
}

// Workspace for a batch of lines: grown on demand (at least doubling), owned
// by the plan.  Earlier calls may still be running on a caller's stream, so
// the store being replaced is retired, not freed: the plan releases it when
// it is itself freed, and nothing waits here on work of other streams.
static int ensure_ws(fmm_plan_s *p, size_t nlines, const Geo &g, double **w) {
  const size_t need = std::max<size_t>(nlines, 1) * ws_doubles(g);
  if (p->wl_cap < need) {
    const size_t cap = std::max(need, 2 * p->wl_cap);
    double *nw = nullptr;
    L2C_HIP(hipMalloc(&nw, cap * sizeof(double)), -1);
    if (p->d_wl) p->wl_retired.push_back(p->d_wl);
    p->d_wl = nw;
    p->wl_cap = cap;
  }
  *w = p->d_wl;
  return 0;
}

// k_down8 launch in the given mode on stream st
template <int MT>
static void launch_down8(fmm_plan_s *p, const Geo &g, const double *d_in, double *d_out, size_t dir,
                         const LineMap &lm, size_t nlines, hipStream_t st, int mode, double *ws) {
  const size_t wsstride = ws_doubles(g);
  dim3 gd((unsigned)((size_t)4 << g.ld), (unsigned)nlines);
  const size_t shm = down_shared_bytes(g);
  const int nt = 64 * (g.s >> 4);
  const double *tab = dir == L2C ? p->t.a : p->t.an;
  const double *htab = p->t.head[dir];
  prof_begin(p->prof, K_DOWN_BAND, st);
  if (mode == K8_FULL && p->down_roles) {
    const size_t shr = down8r_shared_bytes(g);
    const int rsh = std::max(-g.s, std::min(p->roles_shift, 8 * (g.s >> 4))) & ~7;
    if (dir == L2C)
      launch_k(p->prof, k_down8r<0, 0, MT>, gd, 2 * nt, shr, st, d_in, d_out, lm, g, ws, wsstride,
          p->t.D, p->t.Q, p->t.T, tab, htab, rsh);
    else if (p->precompute)
      launch_k(p->prof, k_down8r<1, 1, MT>, gd, 2 * nt, shr, st, d_in, d_out, lm, g, ws, wsstride,
          p->t.D, p->t.Q, p->t.T, tab, htab, rsh);
    else
      launch_k(p->prof, k_down8r<1, 0, MT>, gd, 2 * nt, shr, st, d_in, d_out, lm, g, ws, wsstride,
          p->t.D, p->t.Q, p->t.T, tab, htab, rsh);
  } else if (dir == L2C)
    launch_k(p->prof, k_down8<0, 0, MT>, gd, nt, shm, st, d_in, d_out, lm, g, mode, ws, wsstride,
        p->t.D, p->t.Q, p->t.T, tab, htab);
  else if (p->precompute)
    launch_k(p->prof, k_down8<1, 1, MT>, gd, nt, shm, st, d_in, d_out, lm, g, mode, ws, wsstride,
        p->t.D, p->t.Q, p->t.T, tab, htab);
  else
    launch_k(p->prof, k_down8<1, 0, MT>, gd, nt, shm, st, d_in, d_out, lm, g, mode, ws, wsstride,
        p->t.D, p->t.Q, p->t.T, tab, htab);
  prof_end(p->prof, st);
}

// The split downward pass on st: k_band (before the tree kernels, which it
// does not need), then k_up and k_m2l, then k_far.
template <int MT>
static void launch_band(fmm_plan_s *p, const Geo &g, const double *d_in, double *d_out, size_t dir,
                        const LineMap &lm, size_t nlines, hipStream_t st, int scale) {
  dim3 gd((unsigned)((size_t)4 << g.ld), (unsigned)nlines);
  const size_t shm = band_shared_bytes(g);
  const int nt = 128 * (g.s >> 4);
  const double *tab = dir == L2C ? p->t.a : p->t.an;
  const double *htab = p->t.head[dir];
  prof_begin(p->prof, K_BAND, st);
  if (dir == L2C)
    launch_k(p->prof, k_band<0, 0>, gd, nt, shm, st, d_in, d_out, lm, g, scale, tab, htab);
  else if (p->precompute)
    launch_k(p->prof, k_band<1, 1>, gd, nt, shm, st, d_in, d_out, lm, g, scale, tab, htab);
  else
    launch_k(p->prof, k_band<1, 0>, gd, nt, shm, st, d_in, d_out, lm, g, scale, tab, htab);
  prof_end(p->prof, st);
}

template <int MT>
static void launch_far(fmm_plan_s *p, const Geo &g, double *d_out, size_t dir, const LineMap &lm,
                       size_t nlines, hipStream_t st, double *ws) {
  dim3 gd((unsigned)((size_t)4 << g.ld), (unsigned)nlines);
  const int nt = 64 * (g.s >> 4);
  prof_begin(p->prof, K_FAR, st);
  if (dir == L2C)
    launch_k(p->prof, k_far<0, MT>, gd, nt, far_shared_bytes(g), st, d_out, lm, g, ws,
        ws_doubles(g), p->t.D, p->t.Q, p->t.T);
  else
    launch_k(p->prof, k_far<1, MT>, gd, nt, far_shared_bytes(g), st, d_out, lm, g, ws,
        ws_doubles(g), p->t.D, p->t.Q, p->t.T);
  prof_end(p->prof, st);
}

// Whether this call takes the split downward pass (k_band + k_far): always
// (down_split 1), never (0, the default), or (2) for calls of several lines
// of deep trees (L >= 10, N >= 2^18 at s = 64).  Against k_down8 the split
// pass wins on such line calls (8 lines of 2^20 in 293 us against 315 us,
// 32 lines in 1017 against 1116 us) and loses on single lines (58 against
// 56 us at 2^20); k_down8r beats both (8 lines 286 us, 32 lines 1187
// against 1230 us for the split pass in the same run, one line 48.6 us).
static bool use_split(const fmm_plan_s *p, size_t nlines) {
  return p->geo.fast &&
         (p->down_split == 1 || (p->down_split == 2 && nlines > 1 && p->geo.L >= 10));
}

// One execute (or a batch of lines): upward pass, M2L, downward pass, all on st.
template <int MT>
static int launch_fmm_m(fmm_plan_s *p, const Geo &g, const double *d_in, double *d_out,
                        size_t dir, const LineMap &lm, size_t nlines, hipStream_t st,
                        int band_only, double *ws) {
  const size_t wsstride = ws_doubles(g);
  if (lm.n && !g.fast) {  // per-line lengths: k_up and k_down8 only
    set_error("lines of uneven length need a plan with leaf size s = 16, 32 or 64");
    return -1;
  }
  const bool split = use_split(p, nlines);
  if (split) {
    launch_band<MT>(p, g, d_in, d_out, dir, lm, nlines, st, band_only);
    if (band_only) return 0;
  }
  if (!band_only) {
    dim3 gu((unsigned)((size_t)4 << g.lu), (unsigned)nlines);
    prof_begin(p->prof, K_UP_SUBTREE, st);
    // subtree levels on MFMA for deep subtrees of calls of up to 8 lines
    // (latency-bound: 2^20 47.86 -> 47.31 us per transform, 8 lines of 2^20
    // 248.5 -> 246.0 us); at 2^2 chunks (2^16) and for calls of 32 lines and
    // more (throughput-bound: 32 lines of 2^18 234.9 -> 236.8 us) the VALU
    // levels are as fast or faster
    if (MT > 0 && g.gu >= 3 && nlines <= 8)
      launch_k(p->prof, k_up<MT, MT != 0>, gu, kUpNT, up_shared_bytes(g), st, d_in, lm, g, (int)dir,
               p->t.T, p->t.U, p->t.P, ws, wsstride);
    else
      launch_k(p->prof, k_up<MT, false>, gu, kUpNT, up_shared_bytes(g), st, d_in, lm, g, (int)dir,
               p->t.T, p->t.U, p->t.P, ws, wsstride);
    prof_end(p->prof, st);
    const int ntop = (int)total_blocks(g.lu), ntot = (int)total_blocks(g.L);
    const int bpw = m2l_bpw(g);
    const unsigned gx = (unsigned)(ntop + (ntot - ntop + bpw - 1) / bpw);
    // lines per work-group: share each staged block over up to 8 lines while
    // the grid keeps at least ~2048 work-groups
    int lpb = 1;
    while (lpb < 8 && (size_t)gx * ((nlines + 2 * lpb - 1) / (2 * lpb)) >= 2048) lpb *= 2;
    dim3 gm(gx, (unsigned)((nlines + lpb - 1) / lpb));
    prof_begin(p->prof, K_M2L, st);
    const DevTables::Split &sp = p->t.sp[dir];
    const SplitView sv{p->precision == 1 ? sp.rec : nullptr, sp.pos, sp.n64, sp.n32, sp.n16,
                       sp.rec_bytes};
    const bool mma = MT == 18 && m2l_on_mfma(p, nlines);
    if (!mma && MT == 18 && p->t.At[dir] && !sv.rec && bpw == 4) {
      // the same grid; the blocks below lu from At, with the LDS of one
      // staged block (the top blocks') per work-group instead of four
      launch_k(p->prof, k_m2l_t<MT == 18 ? 18 : 1>, gm, kM2lNT,
          8 * (size_t)(m2l_wave_lds(g.M) + kM2lNT), st, g, p->t.A[dir], p->t.At[dir], ws,
          wsstride, nlines, lpb);
    } else if (!mma) {
      launch_k(p->prof, k_m2l<MT>, gm, kM2lNT, m2l_shared_bytes(g), st, g, p->t.A[dir], ws,
          wsstride, bpw, nlines, lpb, sv);
    } else {
      // the blocks above lu (multipoles summed from the k_up contributions)
      // as before; the rest on MFMA
      if (ntop > 0)
        launch_k(p->prof, k_m2l<MT>, dim3((unsigned)ntop, gm.y), kM2lNT, m2l_shared_bytes(g), st, g,
            p->t.A[dir], ws, wsstride, bpw, nlines, lpb, sv);
      const int nblk = ntot - ntop;
      const unsigned mx = (unsigned)((nblk + 3) / 4);
      const size_t ngroups = (nlines + 7) / 8;
      // line groups per work-group: one while the grid stays small, more so
      // that about 8192 work-groups share the lines
      const size_t gpw = std::max<size_t>(1, (mx * ngroups + 8191) / 8192);
      const dim3 gmm(mx, (unsigned)((ngroups + gpw - 1) / gpw));
      if (p->precision == 2) {
        const DevTables::Slices &sl = p->t.sl[dir];
        launch_k(p->prof, k_m2l_mma<2>, gmm, 256, 0, st, g, p->t.A[dir], (const bf16x8 *)sl.q,
            sl.sc, ws, wsstride, nlines, (int)gpw);
      } else {
        launch_k(p->prof, k_m2l_mma<0>, gmm, 256, 4 * 3 * 18 * 19 * sizeof(double), st, g,
            p->t.A[dir], nullptr, nullptr, ws, wsstride, nlines, (int)gpw);
      }
    }
    prof_end(p->prof, st);
  }
  if (split) {
    launch_far<MT>(p, g, d_out, dir, lm, nlines, st, ws);
    return 0;
  }
  if (g.fast) {
    launch_down8<MT>(p, g, d_in, d_out, dir, lm, nlines, st, band_only ? K8_BAND : K8_FULL, ws);
    return 0;
  }
  dim3 gd((unsigned)((size_t)4 << g.ld), (unsigned)nlines);
  const size_t shm = down_shared_bytes(g);
  const double *tab = dir == L2C ? p->t.a : p->t.an;
  const double *htab = p->t.head[dir];
  prof_begin(p->prof, K_DOWN_BAND, st);
  if (dir == L2C)
    launch_k(p->prof, k_down<0, 0, MT>, gd, kDownNT, shm, st, d_in, d_out, lm, g, band_only, ws,
        wsstride, p->t.D, p->t.Q, p->t.T, tab, htab);
  else if (p->precompute)
    launch_k(p->prof, k_down<1, 1, MT>, gd, kDownNT, shm, st, d_in, d_out, lm, g, band_only, ws,
        wsstride, p->t.D, p->t.Q, p->t.T, tab, htab);
  else
    launch_k(p->prof, k_down<1, 0, MT>, gd, kDownNT, shm, st, d_in, d_out, lm, g, band_only, ws,
        wsstride, p->t.D, p->t.Q, p->t.T, tab, htab);
  prof_end(p->prof, st);
  return 0;
}

static int launch_fmm(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir,
                      const LineMap &lm, size_t nlines, hipStream_t st, int band_only) {
  double *ws = nullptr;
  const Geo &g = call_geo(p, nlines);
  if (ensure_ws(p, nlines, g, &ws)) return -1;
  const int rc = g.M == 18 ? launch_fmm_m<18>(p, g, d_in, d_out, dir, lm, nlines, st, band_only, ws)
                           : launch_fmm_m<0>(p, g, d_in, d_out, dir, lm, nlines, st, band_only, ws);
  if (rc) return rc;
  L2C_HIP(hipGetLastError(), -1);
  return 0;
}
int run_fmm_launch(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir, size_t stride,
                   hipStream_t st, int band_only) {
  LineMap lm{1, 0, 0, stride};
  if (!band_only && p->N <= (size_t)p->direct_max_n) {
    // small N: the direct sums in one launch (the plan's tables a, an, dn
    // are those of a direct plan of the same N)
    prof_begin(p->prof, K_DIRECT, st);
    launch_direct_one(p->prof, (int)p->N, dir, p->t.a, p->t.an, p->t.dn, d_in, d_out, lm, st,
                      p->direct_band);
    prof_end(p->prof, st);
    L2C_HIP(hipGetLastError(), -1);
    return 0;
  }
  return launch_fmm(p, d_in, d_out, dir, lm, 1, st, band_only);
}
int run_fmm(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir, size_t stride,
            hipStream_t st, int band_only) {
  prof_step(p->prof);
  return run_fmm_launch(p, d_in, d_out, dir, stride, st, band_only);
}

int run_direct(const direct_plan_s *dp, const double *d_in, double *d_out, size_t dir,
               size_t stride, hipStream_t st, size_t precompute) {
  (void)precompute;
  prof_step(dp->prof);
  LineMap lm{1, 0, 0, stride};
  prof_begin(dp->prof, K_DIRECT, st);
  launch_direct(dp, d_in, d_out, dir, lm, 1, st);
  prof_end(dp->prof, st);
  L2C_HIP(hipGetLastError(), -1);
  return 0;
}

// Lines of uneven length packed back to back: line l (l < nlines) starts at
// d_off[l] and has d_n[l] <= plan N elements (unit stride).  FMM plans
// (leaf size 16, 32 or 64) or direct plans (L = 0).
int run_packed(fmm_plan_s *p, const direct_plan_s *dp, const double *d_in, double *d_out,
               size_t dir, size_t nlines, const size_t *d_off, const int *d_n, hipStream_t st) {
  for (size_t l0 = 0; l0 < nlines; l0 += 65535) {
    const size_t nl = std::min<size_t>(65535, nlines - l0);
    LineMap lm{1, 0, 0, 1, d_off + l0, d_n + l0};
    if (dp) {
      if (launch_direct_lines(dp, d_in, d_out, dir, lm, nl, st)) return -1;
    } else {
      if (p->wl_cap < nl * ws_doubles(call_geo(p, nl))) {
        set_error("run_packed: workspace not reserved for this many lines");
        return -1;
      }
      if (launch_fmm(p, d_in, d_out, dir, lm, nl, st, 0)) return -1;
    }
  }
  return 0;
}

// Precision form 2 tables of direction dir: the BF16 slices and row scales of
// every block at levels >= lu (k_slice_A).
int build_slices(fmm_plan_s *p, int dir) {
  const Geo &g = p->geo;
  DevTables::Slices &sl = p->t.sl[dir];
  if (sl.q) return 0;
  const size_t nblk = total_blocks(g.L) - total_blocks(g.lu);
  if (nblk == 0) return 0;
  const size_t nq = nblk * 3 * 2 * kSlices * 64 * 8;
  L2C_HIP(hipMalloc(&sl.q, nq * sizeof(unsigned short)), -1);
  L2C_HIP(hipMalloc(&sl.sc, nblk * 48 * sizeof(double)), -1);
  p->hbm_bytes += nq * sizeof(unsigned short) + nblk * 48 * sizeof(double);
  k_slice_A<<<(unsigned)nblk, 64, 0, p->stream>>>(p->t.A[dir], g.M, g.L, g.lu, sl.q, sl.sc);
  L2C_HIP(hipGetLastError(), -1);
  L2C_HIP(hipStreamSynchronize(p->stream), -1);
  return 0;
}

// Grow the plan's line workspace to nlines lines now (so that later launches,
// graph capture included, never allocate).
int reserve_lines(fmm_plan_s *p, size_t nlines) {
  double *w = nullptr;
  const Geo &g = ws_doubles(p->geo1) > ws_doubles(p->geo) ? p->geo1 : p->geo;
  return ensure_ws(p, std::min<size_t>(nlines, 65535), g, &w);
}

// Whether a line call of nlines lines of FMM plan p runs as direct sums in
// one launch: the plan's small-N route (direct_max_n) while the call holds at
// most direct_line_elems elements, where the FMM's fixed cost of about 22 us
// per call is still the larger.  Per call, FMM / direct (us): N = 2048 with
// 8 lines 21.7 / 16.6 (16 lines 21.8 / 29.0), 4096 with 4 lines 23.1 / 14.3
// (8 lines 23.2 / 23.9), 8192 with 2 lines 24.5 / 19.1 (3 lines 24.6 / 25.9).
static bool direct_lines(const fmm_plan_s *p, size_t nlines) {
  return p->N <= (size_t)p->direct_max_n && nlines * p->N <= (size_t)p->direct_line_elems;
}

int run_lines(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir, size_t nlines,
              size_t n0, size_t ls0, size_t ls1, size_t es, hipStream_t st) {
  prof_step(p->prof);
  LineMap lm{n0 ? n0 : 1, ls0, ls1, es};
  if (p->L == 0) {
    // dense MFMA products when the plan has them (N <= 4096) and there are
    // at least four lines, else one row per thread; either way one launch
    const direct_plan_s *dp = p->dplan;
    prof_begin(dp->prof, K_DIRECT, st);
    const int rc = launch_direct_lines(dp, d_in, d_out, dir, lm, nlines, st);
    prof_end(dp->prof, st);
    return rc;
  }
  if (direct_lines(p, nlines)) {
    prof_begin(p->prof, K_DIRECT, st);
    launch_direct_one(p->prof, (int)p->N, dir, p->t.a, p->t.an, p->t.dn, d_in, d_out, lm, st,
                      p->direct_band, (unsigned)nlines);
    prof_end(p->prof, st);
    L2C_HIP(hipGetLastError(), -1);
    return 0;
  }
  // lines go in grid.y: at most 65535 per launch, each chunk reusing the
  // front of the plan's line workspace
  for (size_t l0 = 0; l0 < nlines; l0 += 65535) {
    lm.l0 = l0;
    if (launch_fmm(p, d_in, d_out, dir, lm, std::min<size_t>(65535, nlines - l0), st, 0))
      return -1;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Workspace of l2c_execute_lines_device_noalloc: a fixed device array, one
// copy per GPU, held until the process ends.  Host code never names it in a
// launch or copy: it passes the per-device address from hipGetSymbolAddress.
constexpr size_t kScratchDoubles = (size_t)4 << 20;  // 32 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices] = {};
static bool g_scratch_done[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// The scratch address on the current device, looked up once per device (with
// that device current, as the calling thread has it).  nullptr on failure.
static double *scratch_ptr() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("l2c_execute_lines_device_noalloc: no current device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_done[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return nullptr;
    }
    g_scratch[dev] = (double *)ptr;
    g_scratch_done[dev] = true;
  }
  return g_scratch[dev];
}

// Lines through an FMM plan with the fixed scratch as the workspace: as many
// lines per launch as the scratch holds (and at most 65,535), launches in
// order on st, each reusing the scratch front.  Nothing here allocates, and
// nothing in the scratch needs clearing (every element a launch reads was
// written earlier in the same launch sequence).
static int run_lines_scratch(fmm_plan_s *p, double *scr, const double *d_in, double *d_out,
                             size_t dir, size_t nlines, size_t n0, size_t ls0, size_t ls1,
                             size_t es, hipStream_t st) {
  LineMap lm{n0 ? n0 : 1, ls0, ls1, es};
  if (p->L == 0) return launch_direct_lines(p->dplan, d_in, d_out, dir, lm, nlines, st);
  if (direct_lines(p, nlines)) {  // as l2c_execute_lines_device: no workspace
    launch_direct_one(p->prof, (int)p->N, dir, p->t.a, p->t.an, p->t.dn, d_in, d_out, lm, st,
                      p->direct_band, (unsigned)nlines);
    L2C_HIP(hipGetLastError(), -1);
    return 0;
  }
  const size_t wsd = std::max(ws_doubles(p->geo), ws_doubles(p->geo1));
  const size_t per = std::min<size_t>(65535, kScratchDoubles / wsd);
  if (per == 0) {
    set_error("l2c_execute_lines_device_noalloc: one line of this plan needs " +
              std::to_string(wsd * sizeof(double)) + " bytes of workspace, more than the " +
              std::to_string(kScratchDoubles * sizeof(double)) + "-byte scratch");
    return -1;
  }
  for (size_t l0 = 0; l0 < nlines; l0 += per) {
    lm.l0 = l0;
    const size_t nl = std::min(per, nlines - l0);
    const Geo &g = call_geo(p, nl);
    const int rc = g.M == 18 ? launch_fmm_m<18>(p, g, d_in, d_out, dir, lm, nl, st, 0, scr)
                             : launch_fmm_m<0>(p, g, d_in, d_out, dir, lm, nl, st, 0, scr);
    if (rc) return rc;
    L2C_HIP(hipGetLastError(), -1);
  }
  return 0;
}

}  // namespace l2c

using namespace l2c;

extern "C" {

int l2c_execute_lines_device_noalloc(fmm_plan *p, const double *d_in, double *d_out,
                                     size_t direction, size_t nlines, size_t n0, size_t ls0,
                                     size_t ls1, size_t es, void *stream) {
  clear_error();
  if (!p || !d_in || !d_out || direction > 1 || es == 0 || nlines == 0) {
    set_error("l2c_execute_lines_device_noalloc: invalid argument");
    return -1;
  }
  const bool planned = p->L == 0 ? (p->dplan->direction == BOTH || p->dplan->direction == direction)
                                 : p->t.A[direction] != nullptr;
  if (!planned) {
    set_error("l2c_execute_lines_device_noalloc: direction was not planned");
    return -1;
  }
  double *scr = scratch_ptr();
  if (!scr) return -1;
  // no timing events here (their pool would grow by allocation)
  const int cur = p->prof ? p->prof->cur : 0;
  if (p->prof) p->prof->cur = 0;
  const int rc = run_lines_scratch(p, scr, d_in, d_out, direction, nlines, n0, ls0, ls1, es,
                                   stream ? (hipStream_t)stream : p->stream);
  if (p->prof) p->prof->cur = cur;
  return rc;
}

}  // extern "C"
