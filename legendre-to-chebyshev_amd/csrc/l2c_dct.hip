// l2c_dct.hip -- length-N DCT-II and DCT-III of real FP64 vectors on gfx950:
// the steps between function values at the Chebyshev-Gauss points and
// Chebyshev coefficients, on either side of the L2C / C2L transforms.
//
// Conventions are FFTW's (unnormalised):
//   REDFT10 (kind 2):  y_k = 2 sum_{n<N} x_n cos(pi k (2n+1) / (2N))
//   REDFT01 (kind 3):  y_k = x_0 + 2 sum_{1<=n<N} x_n cos(pi n (2k+1) / (2N))
// The reference times FFTW's REDFT10 as its yardstick for the transform
// (src/bin/l2c.c:414-463, published times src/results/figure4.sh:124-141).
//
// Method (J. Makhoul, "A fast cosine transform in one and two dimensions",
// IEEE Trans. ASSP 28 (1980) 27-34): a length-N DCT of real data is one
// complex FFT of length M = N/2.  With w = e^{-2 pi i / (4N)}:
//   DCT-II:  v_j = x_{2j}, v_{N-1-j} = x_{2j+1} (j < M); z_m = v_{2m} + i v_{2m+1};
//            Z = FFT_M(z); for k < M with Zc = conj(Z_{(M-k) mod M}):
//            y_k - i y_{N-k} = w^k [ (Z_k + Zc) + w^{4k} (-i)(Z_k - Zc) ]
//            (k = 0 gives y_0 and y_M).  x[4m..4m+3] feeds z_m and z_{M-1-m}.
//   DCT-III: V_k = conj(w^k) (Y_k - i Y_{N-k})  (Y_N = 0),
//            Z'_k = (V_k + V_{k+M}) + i conj(w^{4k}) (V_k - V_{k+M}),
//            z' = unnormalised inverse FFT_M(Z'), and
//            x[4m..4m+3] = (z'_m.re, z'_{M-1-m}.im, z'_m.im, z'_{M-1-m}.re).
//
// Kernels.  M <= 2048 (N <= 4096): one work-group holds whole vectors in LDS
// (k_dct2_1p / k_dct3_1p).  Larger M uses the four-step FFT over HBM: M =
// Cf * Lf, and
//   first pass: columns c < Cf (index m = c + Cf n), FFT of length Lf over n,
//               times e^{-+2 pi i c k / M}, rows of the workspace W[c][k];
//   last pass:  the length-Cf DFTs over c, output index Cq k1 + column.
// For M < 2^23 the last pass is one FFT of length Lq = Cf on Cq = Lf columns
// (two passes, lengths up to 2048).  From M = 2^23 on, Cf = L1 L2 is split
// once more (three passes of length <= 256): k_fft_mid runs the FFTs of
// length L2 and the last pass those of length Lq = L1 on Cq = L2 Lf columns.
// Each work-group of the first and last pass owns a chunk of columns together
// with the chunk that the DCT's index symmetry pairs with it, so every pre-
// and post-processing step finds both of its operands in LDS.  DCT-II pairs
// first-pass columns c <-> Cf-1-c (input permutation) and last-pass columns
// <-> (Cq - column) mod Cq (conjugate pairs); DCT-III the other way round.
// FFTs are radix-8 Stockham stages through LDS.
//
// Latency, not bandwidth, bounds a work-group here: every global load a
// work-group needs (its data and its twiddles) is issued at the start, the
// in-FFT twiddles come from a quarter-period table in LDS, and each thread
// owns 8 elements (radix-8 stages) or, in the passes of small jobs, 4
// (radix-4 stages, twice the waves per work-group).  All arithmetic is
// FP64; every twiddle is a correctly rounded table entry (built on the host
// in extended precision) or the product of two of them.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "l2c_internal.h"

namespace l2c {
namespace {

constexpr int kTwLog = 12;        // in-FFT roots e^{-2 pi i u / 4096}
constexpr int kOnePassLogM = 11;  // M <= 2048 in one pass
constexpr int kThreePassLogM = 23;  // M >= 2^23 in three passes, else two
constexpr int kMinPassLog = 5, kMaxPassLog = 11;  // pass lengths 32 ... 2048
constexpr int kMaxMidLog = 8;  // three-pass plans split 2^23, 2^24 as 7/8/8, 8/8/8
constexpr int kMaxLogM = 24;        // three passes of at most 256 each
constexpr int kElems = 2048;      // elements per work-group (at least two columns)
constexpr size_t kMaxGridY = 65535;  // vectors per launch of the multi-pass kernels

// Device twiddle tables of a plan; w = e^{-2 pi i / (4N)}.
struct Tabs {
  const double2 *g1, *g4;  // w^j, w^{4j}: j < M (one pass) or j < max(Cf, Cq)
  const double2 *h1, *h4;  // w^{Cq j}, w^{4 Cq j}, j < Lq (rows of the last pass)
  const double2 *e1, *e4;  // w^{Cf j}, w^{4 Cf j}, j < Lf (rows of the first pass)
  const double2 *ph, *pl;  // e^{-2 pi i t / M} = ph[t >> b] * pl[t & (2^b - 1)]
  const double2 *tw;       // e^{-2 pi i u / 4096}, u < 1024
  int b;
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// a * (S i), S = +-1
template <int S>
__device__ __forceinline__ double2 mul_si(double2 a) {
  return S > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// in-register DFTs, X_k = sum_n v_n e^{S 2 pi i n k / R}
template <int R, int S>
__device__ __forceinline__ void dft(double2 *v) {
  if constexpr (R == 2) {
    const double2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
  } else if constexpr (R == 4) {
    const double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const double2 t2 = cadd(v[1], v[3]), t3 = mul_si<S>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  } else {
    static_assert(R == 8, "radix");
    constexpr double c = 0.70710678118654752440;  // sqrt(2)/2
    double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft<4, S>(e);
    dft<4, S>(o);
    // o_k *= e^{S 2 pi i k / 8}
    const double2 o1 = make_double2(c * (o[1].x - S * o[1].y), c * (o[1].y + S * o[1].x));
    const double2 o2 = mul_si<S>(o[2]);
    const double2 o3 = make_double2(c * (-o[3].x - S * o[3].y), c * (S * o[3].x - o[3].y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
}

// Quarter-period root table of an FFT of length L in LDS:
// ts[r] = e^{-2 pi i r / L}, r < L/4 (empty for L <= 8, whose single stage
// needs no twiddles).  Loaded by the first L/4 threads.
template <int LOGL, int T>
__device__ __forceinline__ void load_quarter(double2 *ts, const double2 *__restrict__ tw) {
  if constexpr (LOGL > 3) {
    constexpr int Q4 = 1 << (LOGL - 2);
#pragma unroll
    for (int it = 0; it < (Q4 + T - 1) / T; ++it) {
      const int r = threadIdx.x + it * T;
      if (Q4 % T == 0 || r < Q4) ts[r] = tw[r << (kTwLog - LOGL)];
    }
  }
}

// e^{-2 pi i u / L} for 0 <= u < L from the quarter table: the quadrant
// u / (L/4) multiplies the first-quadrant root by (-i)^q.
template <int LOGL>
__device__ __forceinline__ double2 quarter_root(const double2 *ts, int u) {
  const double2 w = ts[u & ((1 << (LOGL - 2)) - 1)];
  const int q = u >> (LOGL - 2);
  double2 r = (q & 1) ? make_double2(w.y, -w.x) : w;
  if (q & 2) r = make_double2(-r.x, -r.y);
  return r;
}

// One Stockham stage of radix R on NC columns of length L held in LDS with
// column stride L + 1.  Sub-transform size 2^LOGNS before the stage.  The
// reads of every butterfly go to registers, then the stage writes in place.
template <int LOGL, int NC, int T, int R, int LOGNS, int S>
__device__ __forceinline__ void fft_stage(double2 *buf, const double2 *ts) {
  constexpr int L = 1 << LOGL, LD = L + 1, Q = L / R, TOT = NC * Q;
  constexpr int NB = (TOT + T - 1) / T, NS = 1 << LOGNS;
  constexpr int STEP = L / (NS * R);
  double2 v[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int id = threadIdx.x + b * T;
    if (TOT % T == 0 || id < TOT) {
      const int col = id / Q, j = id % Q;
      const double2 *src = buf + col * LD + j;
#pragma unroll
      for (int t = 0; t < R; ++t) v[b][t] = src[t * Q];
      if constexpr (NS > 1) {
        const int k = j & (NS - 1);
#pragma unroll
        for (int t = 1; t < R; ++t) {
          double2 w = quarter_root<LOGL>(ts, k * t * STEP);
          if constexpr (S > 0) w.y = -w.y;
          v[b][t] = cmul(v[b][t], w);
        }
      }
      dft<R, S>(v[b]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int id = threadIdx.x + b * T;
    if (TOT % T == 0 || id < TOT) {
      const int col = id / Q, j = id % Q, k = j & (NS - 1);
      double2 *dst = buf + col * LD + (j - k) * R + k;
#pragma unroll
      for (int t = 0; t < R; ++t) dst[t * NS] = v[b][t];
    }
  }
  __syncthreads();
}

// FFT (S = -1) or unnormalised inverse FFT (S = +1) of NC LDS columns of
// length 2^LOGL in stages of radix 2^LOGRM (8 or 4), a smaller stage first
// when LOGL is not a multiple of LOGRM.  The caller synchronises before the
// call.
template <int LOGL, int NC, int T, int S, int LOGRM = 3, int LOGNS = 0>
__device__ __forceinline__ void fft_cols(double2 *buf, const double2 *ts) {
  if constexpr (LOGNS < LOGL) {
    constexpr int LOGR = (LOGNS == 0 && LOGL % LOGRM) ? LOGL % LOGRM : LOGRM;
    fft_stage<LOGL, NC, T, 1 << LOGR, LOGNS, S>(buf, ts);
    fft_cols<LOGL, NC, T, S, LOGRM, LOGNS + LOGR>(buf, ts);
  }
}

// Work-group index, XCD-aware: the hardware deals consecutive work-groups
// round robin over the 8 XCDs, so this mapping hands consecutive column
// chunks (which share cache lines) to work-groups on the same XCD and L2.
__device__ __forceinline__ int chunk_index() {
  const int g = gridDim.x, b = blockIdx.x;
  return (g & 7) ? b : (b & 7) * (g >> 3) + (b >> 3);
}

// Column i of a work-group with chunk base a (C = NC/2 chunk columns):
// reflect pairing c <-> L-1-c, and negate pairing c <-> (L-c) mod L where the
// two self-paired columns 0 and L/2 both belong to chunk 0.
template <int C>
__device__ __forceinline__ int col_reflect(int i, int a, int L) {
  return i < C ? a + i : L - 1 - a - (i - C);
}
template <int C>
__device__ __forceinline__ int col_negate(int i, int a, int L) {
  if (i < C) return a + i;
  const int c = a + i - C;
  return c ? L - c : L / 2;
}
template <int C>
__device__ __forceinline__ int partner(int i) {
  return i < C ? i + C : i - C;
}

// e^{-+2 pi i t / M} from the two-level table
__device__ __forceinline__ void pass_root_load(const Tabs &tb, size_t t, double2 &hi, double2 &lo) {
  hi = tb.ph[t >> tb.b];
  lo = tb.pl[t & ((size_t(1) << tb.b) - 1)];
}

// ---- DCT-II (REDFT10) --------------------------------------------------
// Post-processing of output index k (0 <= k < M) from Z_k, Zc =
// conj(Z_{(M-k) mod M}), wk = w^k and w4 = w^{4k}: writes y[k] and y[N-k]
// (y[0] and y[M] for k = 0), elements es apart.
__device__ __forceinline__ void dct2_post(double *y, size_t k, size_t M, double2 Zk, double2 Zc,
                                          double2 wk, double2 w4, size_t es = 1) {
  if (k == 0) {
    y[0] = 2 * (Zk.x + Zk.y);
    y[M * es] = 1.41421356237309504880 * (Zk.x - Zk.y);
    return;
  }
  const double2 e = cadd(Zk, Zc);
  const double2 d = csub(Zk, Zc);
  const double2 o = make_double2(d.y, -d.x);  // -i (Z_k - Zc)
  const double2 W = cmul(wk, cadd(e, cmul(w4, o)));
  y[k * es] = W.x;
  y[(2 * M - k) * es] = -W.y;
}

// Item id of a one-pass work-group -> (vector vi, index i < W).  Unit-stride
// vectors run i fastest (neighbouring lanes, neighbouring words; pairs of
// words as 16-byte accesses); strided lines run the vector fastest
// (neighbouring lanes, neighbouring lines).
template <int W, int NV, bool STRIDED>
__device__ __forceinline__ void item(int id, int &vi, int &i) {
  if (STRIDED) {
    vi = id % NV;
    i = id / NV;
  } else {
    vi = id / W;
    i = id % W;
  }
}

// The start of each of the work-group's lines (line_off divides), in LDS.
template <int NV, int T>
__device__ __forceinline__ void line_starts(size_t *loff, const LineMap &lm, size_t v0,
                                            size_t count) {
#pragma unroll
  for (int it = 0; it < (NV + T - 1) / T; ++it) {
    const int v = threadIdx.x + it * T;
    if (v < NV) loff[v] = v0 + v < count ? line_off(lm, v0 + v) : 0;
  }
}

// Layouts of the one-pass kernels' vectors: contiguous (vector v at v N),
// lines with unit element stride at even offsets, and strided lines.
enum { kVecs = 0, kLinesUnit = 1, kLinesStrided = 2 };

// One pass for M = 2^LOGM <= 2048: NV = 2048 / M vectors per work-group.
// Vector v (v < count) starts at v N (kVecs) or line_off(lm, v) (lines), its
// elements lm.es apart (kLinesStrided) or 1 apart.
template <int LOGM, int NV, int LAY>
__global__ void __launch_bounds__(kElems / 8)
    k_dct2_1p(const double *__restrict__ x, double *__restrict__ y, size_t count, LineMap lm,
              Tabs tb) {
  constexpr int M = 1 << LOGM, LD = M + 1, H = M / 2, T = kElems / 8;
  constexpr int NI = NV * M / T, NL = NV * H / T;  // items per thread: post, load
  extern __shared__ double2 z[];
  double2 *ts = z + NV * LD;
  size_t *loff = reinterpret_cast<size_t *>(ts + (LOGM > 3 ? M / 4 : 0));
  constexpr bool STRIDED = LAY == kLinesStrided;
  const size_t v0 = size_t(blockIdx.x) * NV, es = STRIDED ? lm.es : 1;
  if constexpr (LAY != kVecs) {
    line_starts<NV, T>(loff, lm, v0, count);
    lds_sync();  // line starts; then every global load is issued before the first LDS store
  }
  auto start = [&](int vi) -> size_t { return LAY == kVecs ? (v0 + vi) * (2 * size_t(M)) : loff[vi]; };
  double2 t1[NI], t4[NI], p0[NL], p1[NL];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    int vi, k;
    item<M, NV, STRIDED>(threadIdx.x + it * T, vi, k);
    t1[it] = tb.g1[k];
    t4[it] = tb.g4[k];
  }
#pragma unroll
  for (int it = 0; it < NL; ++it) {
    int vi, m;
    item<H, NV, STRIDED>(threadIdx.x + it * T, vi, m);
    if (v0 + vi < count) {
      const double *src = x + start(vi) + 4 * size_t(m) * es;
      if (STRIDED) {
        p0[it] = make_double2(src[0], src[es]);
        p1[it] = make_double2(src[2 * es], src[3 * es]);
      } else {
        p0[it] = *reinterpret_cast<const double2 *>(src);
        p1[it] = *reinterpret_cast<const double2 *>(src + 2);
      }
    }
  }
  load_quarter<LOGM, T>(ts, tb.tw);
#pragma unroll
  for (int it = 0; it < NL; ++it) {
    int vi, m;
    item<H, NV, STRIDED>(threadIdx.x + it * T, vi, m);
    if (v0 + vi < count) {
      z[vi * LD + m] = make_double2(p0[it].x, p1[it].x);
      z[vi * LD + (M - 1 - m)] = make_double2(p1[it].y, p0[it].y);
    }
  }
  __syncthreads();
  fft_cols<LOGM, NV, T, -1>(z, ts);
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    int vi, k;
    item<M, NV, STRIDED>(threadIdx.x + it * T, vi, k);
    if (v0 + vi < count)
      dct2_post(y + start(vi), k, M, z[vi * LD + k], conj2(z[vi * LD + ((M - k) & (M - 1))]),
                t1[it], t4[it], es);
  }
}

// Pass 1: FFT length L = M2 over n2; NC columns n1 (reflect pairing).
template <int LOGL, int NC, int PT>
__global__ void __launch_bounds__(NC * (1 << LOGL) / PT)
    k_dct2_p1(const double *__restrict__ x, double2 *__restrict__ W, int logM1, Tabs tb) {
  constexpr int L = 1 << LOGL, LD = L + 1, C = NC / 2, T = NC * L / PT;
  extern __shared__ double2 z[];
  double2 *ts = z + NC * LD;
  const int M1 = 1 << logM1;
  const size_t M = size_t(M1) << LOGL;
  x += size_t(blockIdx.y) * 2 * M;
  W += size_t(blockIdx.y) * M;
  const int a = chunk_index() * C;
  // store items: i = id / L, k2 = id % L; twiddle e^{-2 pi i n1 k2 / M}
  double2 th[PT], tl[PT], p0[PT / 2], p1[PT / 2];
#pragma unroll
  for (int it = 0; it < PT; ++it) {
    const int id = threadIdx.x + it * T;
    const size_t n1 = size_t(col_reflect<C>(id / L, a, M1));
    pass_root_load(tb, (n1 * size_t(id % L)) & (M - 1), th[it], tl[it]);
  }
  // load items: i = id % NC, n2 = id / NC < L/2
#pragma unroll
  for (int it = 0; it < PT / 2; ++it) {
    const int id = threadIdx.x + it * T, i = id % NC, n2 = id / NC;
    const size_t m = size_t(col_reflect<C>(i, a, M1)) + (size_t(n2) << logM1);
    p0[it] = *reinterpret_cast<const double2 *>(x + 4 * m);
    p1[it] = *reinterpret_cast<const double2 *>(x + 4 * m + 2);
  }
  load_quarter<LOGL, T>(ts, tb.tw);
#pragma unroll
  for (int it = 0; it < PT / 2; ++it) {
    const int id = threadIdx.x + it * T, i = id % NC, n2 = id / NC;
    z[i * LD + n2] = make_double2(p0[it].x, p1[it].x);
    z[partner<C>(i) * LD + (L - 1 - n2)] = make_double2(p1[it].y, p0[it].y);
  }
  __syncthreads();
  fft_cols<LOGL, NC, T, -1, PT == 8 ? 3 : 2>(z, ts);
#pragma unroll
  for (int it = 0; it < PT; ++it) {
    const int id = threadIdx.x + it * T, i = id / L, k2 = id % L;
    const size_t n1 = size_t(col_reflect<C>(i, a, M1));
    W[(n1 << LOGL) + k2] = cmul(z[i * LD + k2], cmul(th[it], tl[it]));
  }
}

// Pass 2: FFT length L = M1 over n1; NC columns k2 (negate pairing).
template <int LOGL, int NC, int PT>
__global__ void __launch_bounds__(NC * (1 << LOGL) / PT)
    k_dct2_p2(const double2 *__restrict__ W, double *__restrict__ y, int logM2, Tabs tb) {
  constexpr int L = 1 << LOGL, LD = L + 1, C = NC / 2, T = NC * L / PT;
  extern __shared__ double2 z[];
  double2 *ts = z + NC * LD;
  const int M2 = 1 << logM2;
  const size_t M = size_t(M2) << LOGL;
  W += size_t(blockIdx.y) * M;
  y += size_t(blockIdx.y) * 2 * M;
  const int a = chunk_index() * C;
  // items: i = id % NC (fixed per thread, T % NC == 0), k1 = n1 = id / NC
  const int i = threadIdx.x % NC;
  const int k2 = col_negate<C>(i, a, M2);
  const double2 c1 = tb.g1[k2], c4 = tb.g4[k2];
  double2 r1[PT], r4[PT], d[PT];
#pragma unroll
  for (int it = 0; it < PT; ++it) {
    const int k1 = (threadIdx.x + it * T) / NC;
    r1[it] = tb.h1[k1];
    r4[it] = tb.h4[k1];
  }
#pragma unroll
  for (int it = 0; it < PT; ++it) {
    const int n1 = (threadIdx.x + it * T) / NC;
    d[it] = W[size_t(n1) * M2 + k2];
  }
  load_quarter<LOGL, T>(ts, tb.tw);
#pragma unroll
  for (int it = 0; it < PT; ++it) z[i * LD + (threadIdx.x + it * T) / NC] = d[it];
  __syncthreads();
  fft_cols<LOGL, NC, T, -1, PT == 8 ? 3 : 2>(z, ts);
  // the partner of column k2: itself for k2 = 0 (row (L - k1) mod L) and for
  // k2 = M2/2, else column partner(i); rows run backwards (L - 1 - k1)
  const int ip = (k2 == M2 / 2 || k2 == 0) ? i : partner<C>(i);
#pragma unroll
  for (int it = 0; it < PT; ++it) {
    const int k1 = (threadIdx.x + it * T) / NC;
    const int row = k2 == 0 ? ((L - k1) & (L - 1)) : L - 1 - k1;
    dct2_post(y, size_t(k1) * M2 + k2, M, z[i * LD + k1], conj2(z[ip * LD + row]),
              cmul(r1[it], c1), cmul(r4[it], c4));
  }
}

// ---- DCT-III (REDFT01) -------------------------------------------------
// Z'_k from a = Y_k, b = Y_{N-k}, c = Y_{k+M}, d = Y_{M-k}, wk = w^k and
// w4 = w^{4k}.
__device__ __forceinline__ double2 dct3_pre(double a, double b, double c, double d, double2 wk,
                                            double2 w4) {
  constexpr double h = 0.70710678118654752440;  // e^{i pi/4} = h (1 + i)
  const double2 cw = conj2(wk);
  const double2 vk = cmul(cw, make_double2(a, -b));
  const double2 vm = cmul(cw, make_double2(h * (c + d), h * (c - d)));  // e^{i pi/4}(c - i d)
  const double2 s = cadd(vk, vm);
  const double2 t = cmul(conj2(w4), csub(vk, vm));
  return make_double2(s.x - t.y, s.y + t.x);
}

template <int LOGM, int NV, int LAY>
__global__ void __launch_bounds__(kElems / 8)
    k_dct3_1p(const double *__restrict__ Y, double *__restrict__ x, size_t count, LineMap lm,
              Tabs tb) {
  constexpr int M = 1 << LOGM, LD = M + 1, H = M / 2, N = 2 * M, T = kElems / 8;
  constexpr int NI = NV * M / T, NL = NV * N / T, NS = NV * H / T;
  extern __shared__ double2 z[];
  double2 *ts = z + NV * LD;
  size_t *loff = reinterpret_cast<size_t *>(ts + (LOGM > 3 ? M / 4 : 0));
  double *yr = reinterpret_cast<double *>(z);  // column stride 2 LD doubles
  constexpr bool STRIDED = LAY == kLinesStrided;
  const size_t v0 = size_t(blockIdx.x) * NV, es = STRIDED ? lm.es : 1;
  if constexpr (LAY != kVecs) {
    line_starts<NV, T>(loff, lm, v0, count);
    lds_sync();  // line starts; then every global load is issued before the first LDS store
  }
  auto start = [&](int vi) -> size_t { return LAY == kVecs ? (v0 + vi) * (2 * size_t(M)) : loff[vi]; };
  double2 t1[NI], t4[NI];
  double q[NL];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    int vi, k;
    item<M, NV, STRIDED>(threadIdx.x + it * T, vi, k);
    t1[it] = tb.g1[k];
    t4[it] = tb.g4[k];
  }
#pragma unroll
  for (int it = 0; it < NL; ++it) {
    int vi, r;
    item<N, NV, STRIDED>(threadIdx.x + it * T, vi, r);
    if (v0 + vi < count) q[it] = Y[start(vi) + size_t(r) * es];
  }
  load_quarter<LOGM, T>(ts, tb.tw);
#pragma unroll
  for (int it = 0; it < NL; ++it) {
    int vi, r;
    item<N, NV, STRIDED>(threadIdx.x + it * T, vi, r);
    if (v0 + vi < count) yr[vi * 2 * LD + r] = q[it];
  }
  __syncthreads();
  double2 zp[NI];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    int vi, k;
    item<M, NV, STRIDED>(threadIdx.x + it * T, vi, k);
    const double *c = yr + vi * 2 * LD;
    zp[it] = dct3_pre(c[k], k ? c[N - k] : 0.0, c[k + M], c[M - k], t1[it], t4[it]);
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    int vi, k;
    item<M, NV, STRIDED>(threadIdx.x + it * T, vi, k);
    z[vi * LD + k] = zp[it];
  }
  __syncthreads();
  fft_cols<LOGM, NV, T, 1>(z, ts);
#pragma unroll
  for (int it = 0; it < NS; ++it) {
    int vi, m;
    item<H, NV, STRIDED>(threadIdx.x + it * T, vi, m);
    if (v0 + vi < count) {
      const double2 A = z[vi * LD + m], B = z[vi * LD + (M - 1 - m)];
      double *dst = x + start(vi) + 4 * size_t(m) * es;
      if (STRIDED) {
        dst[0] = A.x;
        dst[es] = B.y;
        dst[2 * es] = A.y;
        dst[3 * es] = B.x;
      } else {
        *reinterpret_cast<double2 *>(dst) = make_double2(A.x, B.y);
        *reinterpret_cast<double2 *>(dst + 2) = make_double2(A.y, B.x);
      }
    }
  }
}

// Pass 1: inverse FFT length L = M2 over n2; NC columns n1 (negate pairing).
// Y is read as a 2 M2 x M1 row-major matrix.
template <int LOGL, int NC, int PT>
__global__ void __launch_bounds__(NC * (1 << LOGL) / PT)
    k_dct3_p1(const double *__restrict__ Y, double2 *__restrict__ W, int logM1, Tabs tb) {
  constexpr int L = 1 << LOGL, LD = L + 1, C = NC / 2, T = NC * L / PT;
  extern __shared__ double2 z[];
  double2 *ts = z + NC * LD;
  double *yr = reinterpret_cast<double *>(z);
  const int M1 = 1 << logM1;
  const size_t M = size_t(M1) << LOGL;
  Y += size_t(blockIdx.y) * 2 * M;
  W += size_t(blockIdx.y) * M;
  const int a = chunk_index() * C;
  // items: i = id % NC (fixed per thread), n2 = id / NC; k = n1 + M1 n2
  const int i = threadIdx.x % NC;
  const int n1 = col_negate<C>(i, a, M1);
  const double2 c1 = tb.g1[n1], c4 = tb.g4[n1];
  double2 r1[PT], r4[PT];
  double q[2 * PT];
#pragma unroll
  for (int it = 0; it < PT; ++it) {
    const int n2 = (threadIdx.x + it * T) / NC;
    r1[it] = tb.e1[n2];
    r4[it] = tb.e4[n2];
  }
#pragma unroll
  for (int it = 0; it < 2 * PT; ++it) {
    const int r = (threadIdx.x + it * T) / NC;  // rows 0 .. 2L-1
    q[it] = Y[size_t(r) * M1 + n1];
  }
  load_quarter<LOGL, T>(ts, tb.tw);
#pragma unroll
  for (int it = 0; it < 2 * PT; ++it) yr[i * 2 * LD + (threadIdx.x + it * T) / NC] = q[it];
  __syncthreads();
  const double *c = yr + i * 2 * LD;
  const double *p = yr + (n1 == M1 / 2 || n1 == 0 ? i : partner<C>(i)) * 2 * LD;
  double2 zp[PT];
#pragma unroll
  for (int it = 0; it < PT; ++it) {
    const int n2 = (threadIdx.x + it * T) / NC;
    double bb, dd;
    if (n1 == 0) {
      bb = n2 ? c[2 * L - n2] : 0.0;
      dd = c[L - n2];
    } else {
      bb = p[2 * L - 1 - n2];
      dd = p[L - 1 - n2];
    }
    zp[it] = dct3_pre(c[n2], bb, c[n2 + L], dd, cmul(c1, r1[it]), cmul(c4, r4[it]));
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < PT; ++it) z[i * LD + (threadIdx.x + it * T) / NC] = zp[it];
  __syncthreads();
  // store items: i = id / L, k2 = id % L; twiddle e^{+2 pi i n1 k2 / M}
  double2 th[PT], tl[PT];
#pragma unroll
  for (int it = 0; it < PT; ++it) {
    const int id = threadIdx.x + it * T;
    const size_t m1 = size_t(col_negate<C>(id / L, a, M1));
    pass_root_load(tb, (m1 * size_t(id % L)) & (M - 1), th[it], tl[it]);
  }
  fft_cols<LOGL, NC, T, 1, PT == 8 ? 3 : 2>(z, ts);
#pragma unroll
  for (int it = 0; it < PT; ++it) {
    const int id = threadIdx.x + it * T, j = id / L, k2 = id % L;
    const size_t m1 = size_t(col_negate<C>(j, a, M1));
    W[(m1 << LOGL) + k2] = cmul(z[j * LD + k2], conj2(cmul(th[it], tl[it])));
  }
}

// Pass 2: inverse FFT length L = M1 over n1; NC columns k2 (reflect pairing).
template <int LOGL, int NC, int PT>
__global__ void __launch_bounds__(NC * (1 << LOGL) / PT)
    k_dct3_p2(const double2 *__restrict__ W, double *__restrict__ x, int logM2, Tabs tb) {
  constexpr int L = 1 << LOGL, LD = L + 1, C = NC / 2, T = NC * L / PT;
  extern __shared__ double2 z[];
  double2 *ts = z + NC * LD;
  const int M2 = 1 << logM2;
  const size_t M = size_t(M2) << LOGL;
  W += size_t(blockIdx.y) * M;
  x += size_t(blockIdx.y) * 2 * M;
  const int a = chunk_index() * C;
  const int i = threadIdx.x % NC;
  const int k2 = col_reflect<C>(i, a, M2);
  double2 d[PT];
#pragma unroll
  for (int it = 0; it < PT; ++it) d[it] = W[size_t((threadIdx.x + it * T) / NC) * M2 + k2];
  load_quarter<LOGL, T>(ts, tb.tw);
#pragma unroll
  for (int it = 0; it < PT; ++it) z[i * LD + (threadIdx.x + it * T) / NC] = d[it];
  __syncthreads();
  fft_cols<LOGL, NC, T, 1, PT == 8 ? 3 : 2>(z, ts);
  const int ip = partner<C>(i);
#pragma unroll
  for (int it = 0; it < PT / 2; ++it) {
    const int k1 = (threadIdx.x + it * T) / NC;  // < L/2
    const size_t m = size_t(k1) * M2 + k2;
    const double2 A = z[i * LD + k1], B = z[ip * LD + (L - 1 - k1)];
    double *dst = x + 4 * m;
    *reinterpret_cast<double2 *>(dst) = make_double2(A.x, B.y);
    *reinterpret_cast<double2 *>(dst + 2) = make_double2(A.y, B.x);
  }
}

// ---- middle pass (three-pass form) ---------------------------------------
// With M = L1 L2 L3 the first pass leaves W[c][k3] (c = n1 + L1 n2 < L1 L2,
// k3 < L3); the length-L1L2 DFT over c that remains for every k3 is itself
// split: this pass runs the FFTs of length L = L2 over n2 of W[n1 + L1 n2][k3],
// multiplies by e^{S 2 pi i n1 k2 / (L1 L2)} and writes W2[n1][k3 + L3 k2],
// which the last pass reads as L1 rows of L2 L3 columns.  A work-group owns
// one n1 and NC consecutive k3, so every load and store is a run of NC
// complex values.
template <int LOGL, int NC, int S, int PT>
__global__ void __launch_bounds__(NC * (1 << LOGL) / PT)
    k_fft_mid(const double2 *__restrict__ W, double2 *__restrict__ W2, int logL1, int logL3,
              Tabs tb) {
  constexpr int L = 1 << LOGL, LD = L + 1, T = NC * L / PT;
  extern __shared__ double2 z[];
  double2 *ts = z + NC * LD;
  const size_t M = size_t(1) << (logL1 + LOGL + logL3);
  W += size_t(blockIdx.y) * M;
  W2 += size_t(blockIdx.y) * M;
  const int per = (1 << logL3) / NC;  // k3 chunks per n1
  const size_t n1 = blockIdx.x / per;
  const int i = threadIdx.x % NC;
  const size_t k3 = size_t(blockIdx.x % per) * NC + i;
  double2 th[PT], tl[PT], d[PT];
#pragma unroll
  for (int it = 0; it < PT; ++it) {
    const size_t k2 = (threadIdx.x + it * T) / NC;
    pass_root_load(tb, (n1 * k2) << logL3, th[it], tl[it]);
  }
#pragma unroll
  for (int it = 0; it < PT; ++it) {
    const size_t n2 = (threadIdx.x + it * T) / NC;
    d[it] = W[((n1 + (n2 << logL1)) << logL3) + k3];
  }
  load_quarter<LOGL, T>(ts, tb.tw);
#pragma unroll
  for (int it = 0; it < PT; ++it) z[i * LD + (threadIdx.x + it * T) / NC] = d[it];
  __syncthreads();
  fft_cols<LOGL, NC, T, S, PT == 8 ? 3 : 2>(z, ts);
#pragma unroll
  for (int it = 0; it < PT; ++it) {
    const size_t k2 = (threadIdx.x + it * T) / NC;
    double2 w = cmul(th[it], tl[it]);
    if constexpr (S > 0) w.y = -w.y;
    W2[(((n1 << LOGL) + k2) << logL3) + k3] = cmul(z[i * LD + k2], w);
  }
}

// N = 1 and N = 2, one vector per thread.
__global__ void k_dct_tiny(const double *__restrict__ in, double *__restrict__ out, size_t count,
                           LineMap lm, int N, int kind) {
  const size_t v = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= count) return;
  constexpr double r2 = 1.41421356237309504880;
  const size_t o = line_off(lm, v), es = lm.es;
  if (N == 1) {
    out[o] = kind == 2 ? 2 * in[o] : in[o];
    return;
  }
  const double a = in[o], b = in[o + es];
  if (kind == 2) {
    out[o] = 2 * (a + b);
    out[o + es] = r2 * (a - b);
  } else {
    out[o] = a + r2 * b;
    out[o + es] = a - r2 * b;
  }
}

// ---- host ----------------------------------------------------------------
// e^{-2 pi i u / R} (R a multiple of 8), the angle reduced to [0, pi/4] and
// its cosine and sine evaluated in long double.
double2 host_root(size_t u, size_t R) {
  u %= R;
  bool neg_s = false, neg_c = false, swap = false;
  if (u > R / 2) {  // sin(2pi - a) = -sin a
    u = R - u;
    neg_s = true;
  }
  if (u > R / 4) {  // cos(pi - a) = -cos a
    u = R / 2 - u;
    neg_c = true;
  }
  if (u > R / 8) {  // cos(pi/2 - a) = sin a
    u = R / 4 - u;
    swap = true;
  }
  const long double th = 2.0L * 3.141592653589793238462643383279502884L * (long double)u /
                         (long double)R;
  long double c = cosl(th), s = sinl(th);
  if (swap) std::swap(c, s);
  if (neg_c) c = -c;
  if (neg_s) s = -s;
  return make_double2(double(c), double(-s));
}

int ilog2(size_t n) {
  int l = 0;
  while ((size_t(1) << l) < n) ++l;
  return l;
}

}  // namespace
}  // namespace l2c

struct l2c_dct_plan_s {
  size_t N = 0, count = 0;
  int logM = 0;
  // Multi-pass form, M = 2^logM.  The first pass runs FFTs of length 2^l_first
  // on 2^c_first columns, the last pass FFTs of length 2^l_last on 2^c_last
  // columns; a three-pass plan (l_mid > 0) has M = 2^(l_last + l_mid + l_first)
  // and runs the middle pass of length 2^l_mid between them.  Two passes:
  // c_first = l_last and c_last = l_first.
  int l_first = 0, c_first = 0, l_mid = 0, l_last = 0, c_last = 0;
  int size = 1;      // elements per work-group in the passes: 0 -> 1024, 1 -> 2048, 2 -> 4096
  int size_mid = 1;  // the same for the middle pass
  int pt = 8;        // elements per thread in the passes: 8 (radix-8 stages) or 4 (radix 4)
  int dev = 0;
  l2c::Tabs tb{};
  double2 *d_tab = nullptr;   // every twiddle table, back to back
  double2 *d_work = nullptr;  // count * M per pass boundary (one or two)
  double *d_in = nullptr, *d_out = nullptr, *h_in = nullptr, *h_out = nullptr;  // host entry
  hipStream_t st = nullptr;
};

namespace l2c {
namespace {

template <int LOGM, int LAY>
void launch_1p_lay(const l2c_dct_plan *p, const double *in, double *out, int kind, size_t count,
                   const LineMap &lm, hipStream_t st) {
  constexpr int M = 1 << LOGM, NV = kElems / M;
  const size_t shmem =
      (size_t(NV) * (M + 1) + (LOGM > 3 ? M / 4 : 0)) * sizeof(double2) + NV * sizeof(size_t);
  const dim3 grid(unsigned((count + NV - 1) / NV)), blk(kElems / 8);
  if (kind == 2)
    hipLaunchKernelGGL((k_dct2_1p<LOGM, NV, LAY>), grid, blk, shmem, st, in, out, count, lm, p->tb);
  else
    hipLaunchKernelGGL((k_dct3_1p<LOGM, NV, LAY>), grid, blk, shmem, st, in, out, count, lm, p->tb);
}

template <int LOGM>
void launch_1p(const l2c_dct_plan *p, const double *in, double *out, int kind, size_t count,
               const LineMap &lm, int lay, hipStream_t st) {
  if (lay == kVecs)
    launch_1p_lay<LOGM, kVecs>(p, in, out, kind, count, lm, st);
  else if (lay == kLinesUnit)
    launch_1p_lay<LOGM, kLinesUnit>(p, in, out, kind, count, lm, st);
  else
    launch_1p_lay<LOGM, kLinesStrided>(p, in, out, kind, count, lm, st);
}

// Columns per work-group for a pass of FFT length 2^LOGL: E elements, and at
// least two columns (a chunk column and its pair).
template <int LOGL, int E>
constexpr int pass_nc() {
  return (1 << LOGL) >= E / 2 ? 2 : E >> LOGL;
}

template <int LOGL, int E, int PT>
void launch_pass_e(const l2c_dct_plan *p, int pass, const double *in, double *out, int kind,
                   hipStream_t st) {
  constexpr int L = 1 << LOGL, NC = pass_nc<LOGL, E>(), T = NC * L / PT;
  const size_t shmem = (size_t(NC) * (L + 1) + L / 4) * sizeof(double2);
  const int other = pass == 1 ? p->c_first : p->c_last;  // number of columns = 2^other
  const dim3 grid(unsigned((size_t(1) << other) / NC), unsigned(p->count));
  // the last pass reads the second workspace after a middle pass
  double2 *w1 = p->d_work;
  const double2 *wl = p->l_mid ? p->d_work + p->count * (size_t(1) << p->logM) : p->d_work;
  if (kind == 2 && pass == 1)
    hipLaunchKernelGGL((k_dct2_p1<LOGL, NC, PT>), grid, dim3(T), shmem, st, in, w1, other, p->tb);
  else if (kind == 2)
    hipLaunchKernelGGL((k_dct2_p2<LOGL, NC, PT>), grid, dim3(T), shmem, st, wl, out, other, p->tb);
  else if (pass == 1)
    hipLaunchKernelGGL((k_dct3_p1<LOGL, NC, PT>), grid, dim3(T), shmem, st, in, w1, other, p->tb);
  else
    hipLaunchKernelGGL((k_dct3_p2<LOGL, NC, PT>), grid, dim3(T), shmem, st, wl, out, other, p->tb);
}

template <int LOGL, int E, int PT>
void launch_mid_e(const l2c_dct_plan *p, int kind, hipStream_t st) {
  constexpr int L = 1 << LOGL, NC = E / L, T = E / PT;
  const size_t shmem = (size_t(NC) * (L + 1) + L / 4) * sizeof(double2);
  const dim3 grid(unsigned(size_t(1) << (p->l_last + p->l_first)) / NC, unsigned(p->count));
  const double2 *w1 = p->d_work;
  double2 *w2 = p->d_work + p->count * (size_t(1) << p->logM);
  if (kind == 2)
    hipLaunchKernelGGL((k_fft_mid<LOGL, NC, -1, PT>), grid, dim3(T), shmem, st, w1, w2, p->l_last,
                       p->l_first, p->tb);
  else
    hipLaunchKernelGGL((k_fft_mid<LOGL, NC, 1, PT>), grid, dim3(T), shmem, st, w1, w2, p->l_last,
                       p->l_first, p->tb);
}

// Work-group size by job size: wide work-groups (4096 elements, 512 threads,
// 64 KB of LDS) touch twice as many columns per row segment and are used when
// the job still gives at least two of them per CU; narrow ones (1024
// elements) when 2048-element work-groups would leave fewer than one per CU,
// so that one work-group's loads overlap another's FFT (at one per CU, N =
// 2^20, the 2048-element groups are faster).
template <int LOGL>
void launch_pass(const l2c_dct_plan *p, int pass, const double *in, double *out, int kind,
                 hipStream_t st) {
  if (p->size == 2)
    launch_pass_e<LOGL, 2 * kElems, 8>(p, pass, in, out, kind, st);
  else if (p->size == 1 && p->pt == 4)
    launch_pass_e<LOGL, kElems, 4>(p, pass, in, out, kind, st);
  else if (p->size == 1)
    launch_pass_e<LOGL, kElems, 8>(p, pass, in, out, kind, st);
  else if (p->pt == 4)
    launch_pass_e<LOGL, kElems / 2, 4>(p, pass, in, out, kind, st);
  else
    launch_pass_e<LOGL, kElems / 2, 8>(p, pass, in, out, kind, st);
}

template <int LOG>
int dispatch_mid(const l2c_dct_plan *p, int kind, hipStream_t st) {
  if constexpr (LOG > kMaxMidLog) {
    set_error("dct: middle pass length out of range");
    return -1;
  } else {
    if (p->l_mid == LOG) {
      if (p->size_mid == 2)
        launch_mid_e<LOG, 2 * kElems, 8>(p, kind, st);
      else if (p->size_mid == 1 && p->pt == 4)
        launch_mid_e<LOG, kElems, 4>(p, kind, st);
      else if (p->size_mid == 1)
        launch_mid_e<LOG, kElems, 8>(p, kind, st);
      else if (p->pt == 4)
        launch_mid_e<LOG, kElems / 2, 4>(p, kind, st);
      else
        launch_mid_e<LOG, kElems / 2, 8>(p, kind, st);
      return 0;
    }
    return dispatch_mid<LOG + 1>(p, kind, st);
  }
}

template <int LOG>
int dispatch_pass(int logl, const l2c_dct_plan *p, int pass, const double *in, double *out,
                  int kind, hipStream_t st) {
  if constexpr (LOG > kMaxPassLog) {
    set_error("dct: pass length out of range");
    return -1;
  } else {
    if (logl == LOG) {
      launch_pass<LOG>(p, pass, in, out, kind, st);
      return 0;
    }
    return dispatch_pass<LOG + 1>(logl, p, pass, in, out, kind, st);
  }
}

template <int LOG>
int dispatch_1p(const l2c_dct_plan *p, const double *in, double *out, int kind, size_t count,
                const LineMap &lm, int lay, hipStream_t st) {
  if constexpr (LOG > kOnePassLogM) {
    set_error("dct: one-pass length out of range");
    return -1;
  } else {
    if (p->logM == LOG) {
      launch_1p<LOG>(p, in, out, kind, count, lm, lay, st);
      return 0;
    }
    return dispatch_1p<LOG + 1>(p, in, out, kind, count, lm, lay, st);
  }
}

// Lines of a one-pass plan (N <= 4096): line l starts at line_off(lm, l).
int run_dct_lines(const l2c_dct_plan *p, const double *in, double *out, int kind, size_t nlines,
                  const LineMap &lm, int lay, hipStream_t st) {
  if (p->N <= 2) {
    const unsigned blocks = unsigned((nlines + 255) / 256);
    hipLaunchKernelGGL(k_dct_tiny, dim3(blocks), dim3(256), 0, st, in, out, nlines, lm,
                       int(p->N), kind);
  } else if (dispatch_1p<1>(p, in, out, kind, nlines, lm, lay, st)) {
    return -1;
  }
  L2C_HIP(hipGetLastError(), -1);
  return 0;
}

int run_dct(const l2c_dct_plan *p, const double *in, double *out, int kind, hipStream_t st) {
  if (p->logM <= kOnePassLogM) {
    return run_dct_lines(p, in, out, kind, p->count, LineMap{p->count, p->N, 0, 1}, kVecs, st);
  } else {
    // the passes put the vectors in grid.y: at most kMaxGridY per launch,
    // each chunk reusing the front of the workspace
    for (size_t v0 = 0; v0 < p->count; v0 += kMaxGridY) {
      l2c_dct_plan_s q = *p;
      q.count = std::min<size_t>(kMaxGridY, p->count - v0);
      const double *ci = in + v0 * p->N;
      double *co = out + v0 * p->N;
      if (dispatch_pass<kMinPassLog>(q.l_first, &q, 1, ci, co, kind, st)) return -1;
      if (q.l_mid && dispatch_mid<kMinPassLog>(&q, kind, st)) return -1;
      if (dispatch_pass<kMinPassLog>(q.l_last, &q, 2, ci, co, kind, st)) return -1;
    }
  }
  L2C_HIP(hipGetLastError(), -1);
  return 0;
}

void dct_release(l2c_dct_plan *p) {
  if (!p) return;
  int cur = 0;
  const bool swap = hipGetDevice(&cur) == hipSuccess && cur != p->dev;
  if (swap) (void)hipSetDevice(p->dev);
  (void)hipDeviceSynchronize();
  if (p->d_tab) (void)hipFree(p->d_tab);
  if (p->d_work) (void)hipFree(p->d_work);
  if (p->d_in) (void)hipFree(p->d_in);
  if (p->d_out) (void)hipFree(p->d_out);
  if (p->h_in) (void)hipHostFree(p->h_in);
  if (p->h_out) (void)hipHostFree(p->h_out);
  if (p->st) (void)hipStreamDestroy(p->st);
  if (swap) (void)hipSetDevice(cur);
  delete p;
}

l2c_dct_plan *create_dct_impl(size_t N, size_t count) {
  clear_error();
  if (N == 0 || count == 0 || (N & (N - 1))) {
    set_error("l2c_create_dct: N must be a power of two >= 1 and count >= 1");
    return nullptr;
  }
  const int logN = ilog2(N);
  if (logN - 1 > kMaxLogM) {
    set_error("l2c_create_dct: N above 2^25");
    return nullptr;
  }
  auto *p = new l2c_dct_plan_s();
  p->N = N;
  p->count = count;
  if (hipGetDevice(&p->dev) != hipSuccess) {
    set_error("l2c_create_dct: no HIP device");
    delete p;
    return nullptr;
  }
  if (N <= 2) return p;
  p->logM = logN - 1;
  const int lm = p->logM;
  const bool multi = lm > kOnePassLogM;
  if (!multi) {
    p->l_last = lm;  // one pass: the tables below see M as the last pass length
  } else if (lm < kThreePassLogM) {
    // the first pass takes the longer length when lm is odd (its column
    // runs are the long ones; the last pass's short FFTs leave it more
    // columns per work-group for its scattered outputs): 2^20 DCT-II 23.8
    // -> 22.8 us, 2^22 63.2 -> 56.2 us.  L2C_DCT_FIRST overrides (within
    // the pass length limits).
    p->l_first = lm - lm / 2;
    if (const char *e = std::getenv("L2C_DCT_FIRST"))
      p->l_first = std::max(std::max(lm - kMaxPassLog, kMinPassLog),
                            std::min(std::min(kMaxPassLog, lm - kMinPassLog), std::atoi(e)));
    p->l_last = lm - p->l_first;
  } else {
    // the last pass (whose real outputs need the widest column runs) gets the
    // shortest length
    p->l_last = lm / 3;
    p->l_mid = (lm - p->l_last) / 2;
    p->l_first = lm - p->l_last - p->l_mid;
  }
  p->c_first = lm - p->l_first;
  p->c_last = lm - p->l_last;
  const size_t M = size_t(1) << lm, R = 4 * N;
  const size_t Lq = size_t(1) << p->l_last, Cq = size_t(1) << p->c_last;    // last pass
  const size_t Lf = size_t(1) << p->l_first, Cf = size_t(1) << p->c_first;  // first pass
  const int b = (lm + 1) / 2;
  // table layout: g1 g4 (G entries), h1 h4 (Lq), e1 e4 (Lf), ph (M >> b), pl (2^b), tw (1024)
  const size_t G = multi ? std::max(Cf, Cq) : M;
  const size_t o_g = 0, o_h = 2 * G, o_e = o_h + 2 * Lq, o_p = o_e + 2 * Lf,
               o_l = o_p + (M >> b), o_t = o_l + (size_t(1) << b), total = o_t + 1024;
  std::vector<double2> h(total);
  for (size_t j = 0; j < G; ++j) {
    h[o_g + j] = host_root(j, R);
    h[o_g + G + j] = host_root(4 * j, R);
  }
  for (size_t j = 0; j < Lq; ++j) {  // rows of the last pass, k = Cq k1 + column
    h[o_h + j] = host_root(Cq * j, R);
    h[o_h + Lq + j] = host_root(4 * Cq * j, R);
  }
  for (size_t j = 0; j < Lf; ++j) {  // rows of the first pass, k = column + Cf n2
    h[o_e + j] = host_root(Cf * j, R);
    h[o_e + Lf + j] = host_root(4 * Cf * j, R);
  }
  for (size_t j = 0; j < (M >> b); ++j) h[o_p + j] = host_root(8 * (j << b), R);
  for (size_t j = 0; j < (size_t(1) << b); ++j) h[o_l + j] = host_root(8 * j, R);
  for (size_t j = 0; j < 1024; ++j) h[o_t + j] = host_root(j, size_t(1) << kTwLog);
  const size_t work = multi ? (p->l_mid ? 2 : 1) * count * M : 0;
  if (multi && (M / (2 * kElems)) * count >= 512)
    p->size = 2;
  else if ((M / kElems) * count < 256)  // 2^20: 1024-element groups 22.6, 2048 21.7 us
    p->size = 0;
  // the middle pass's runs of E / L_mid columns must fit in L_first
  p->size_mid = p->size == 2 && p->l_mid + p->l_first < 12 ? 1 : p->size;
  // passes of length <= 256 run faster with four elements per thread (twice
  // the waves per work-group; N = 2^13 in 7.7 against 9.0 us, 256 vectors of
  // 2^13 21.7 against 24.4 us); longer ones with eight (2^20: 23.8 against
  // 24.3 us).  L2C_DCT_PT = 4 or 8 overrides the rule.
  p->pt = std::max(p->l_first, std::max(p->l_mid, p->l_last)) <= 8 ? 4 : 8;
  if (const char *e = std::getenv("L2C_DCT_PT")) p->pt = std::atoi(e) == 4 ? 4 : 8;
  if (hipMalloc((void **)&p->d_tab, total * sizeof(double2)) != hipSuccess ||
      (work && hipMalloc((void **)&p->d_work, work * sizeof(double2)) != hipSuccess) ||
      hipMemcpy(p->d_tab, h.data(), total * sizeof(double2), hipMemcpyHostToDevice) !=
          hipSuccess) {
    set_error("l2c_create_dct: device allocation failed");
    dct_release(p);
    return nullptr;
  }
  const double2 *d = p->d_tab;
  p->tb = Tabs{d + o_g, d + o_g + G, d + o_h, d + o_h + Lq, d + o_e, d + o_e + Lf,
               d + o_p, d + o_l, d + o_t, b};
  return p;
}

bool aligned16(const void *q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; }

}  // namespace
}  // namespace l2c

using namespace l2c;

extern "C" {

l2c_dct_plan *l2c_create_dct(size_t N, size_t count) { return create_dct_impl(N, count); }

void l2c_free_dct(l2c_dct_plan *p) { dct_release(p); }

int l2c_dct_synchronize(l2c_dct_plan *p) {
  clear_error();
  if (!p) {
    set_error("l2c_dct_synchronize: NULL plan");
    return -1;
  }
  int cur = 0;
  L2C_HIP(hipGetDevice(&cur), -1);
  if (cur != p->dev) L2C_HIP(hipSetDevice(p->dev), -1);
  const hipError_t e = hipDeviceSynchronize();
  if (cur != p->dev) (void)hipSetDevice(cur);
  L2C_HIP(e, -1);
  return 0;
}

int l2c_execute_dct_device(l2c_dct_plan *p, const double *d_in, double *d_out, int kind,
                           void *stream) {
  clear_error();
  if (!p || !d_in || !d_out) {
    set_error("l2c_execute_dct_device: NULL argument");
    return -1;
  }
  if (kind != 2 && kind != 3) {
    set_error("l2c_execute_dct_device: kind must be 2 (REDFT10) or 3 (REDFT01)");
    return -1;
  }
  if (!aligned16(d_in) || !aligned16(d_out)) {
    set_error("l2c_execute_dct_device: device arrays must be 16-byte aligned");
    return -1;
  }
  return run_dct(p, d_in, d_out, kind, static_cast<hipStream_t>(stream));
}

int l2c_execute_dct_lines_device(l2c_dct_plan *p, const double *d_in, double *d_out, int kind,
                                 size_t nlines, size_t n0, size_t ls0, size_t ls1, size_t es,
                                 void *stream) {
  clear_error();
  if (!p || !d_in || !d_out || nlines == 0 || n0 == 0 || es == 0) {
    set_error("l2c_execute_dct_lines_device: invalid argument");
    return -1;
  }
  if (kind != 2 && kind != 3) {
    set_error("l2c_execute_dct_lines_device: kind must be 2 (REDFT10) or 3 (REDFT01)");
    return -1;
  }
  if (p->N > 2 && p->logM > kOnePassLogM) {
    set_error("l2c_execute_dct_lines_device: lines longer than 4096 are not supported "
              "(use l2c_execute_dct_device on contiguous vectors)");
    return -1;
  }
  // unit-stride lines at even offsets of 16-byte aligned arrays move element
  // pairs as 16-byte accesses; anything else takes the strided kernels
  const bool strided = es != 1 || (ls0 & 1) || (nlines > n0 && (ls1 & 1)) || !aligned16(d_in) ||
                       !aligned16(d_out);
  return run_dct_lines(p, d_in, d_out, kind, nlines, LineMap{n0, ls0, ls1, es},
                       strided ? kLinesStrided : kLinesUnit, static_cast<hipStream_t>(stream));
}

int l2c_execute_dct(l2c_dct_plan *p, const double *in, double *out, int kind) {
  clear_error();
  if (!p || !in || !out) {
    set_error("l2c_execute_dct: NULL argument");
    return -1;
  }
  if (kind != 2 && kind != 3) {
    set_error("l2c_execute_dct: kind must be 2 (REDFT10) or 3 (REDFT01)");
    return -1;
  }
  const size_t bytes = p->N * p->count * sizeof(double);
  if (!p->st) {
    // everything into locals first: the plan takes the stream and buffers
    // only when all of them exist, so a failed call leaves it unchanged
    hipStream_t st = nullptr;
    double *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr;
    if (hipStreamCreate(&st) != hipSuccess || hipMalloc((void **)&a, bytes) != hipSuccess ||
        hipMalloc((void **)&b, bytes) != hipSuccess ||
        hipHostMalloc((void **)&c, bytes) != hipSuccess ||
        hipHostMalloc((void **)&d, bytes) != hipSuccess) {
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
      if (c) (void)hipHostFree(c);
      if (d) (void)hipHostFree(d);
      if (st) (void)hipStreamDestroy(st);
      set_error("l2c_execute_dct: staging allocation failed");
      return -1;
    }
    p->d_in = a;
    p->d_out = b;
    p->h_in = c;
    p->h_out = d;
    p->st = st;
  }
  std::memcpy(p->h_in, in, bytes);
  L2C_HIP(hipMemcpyAsync(p->d_in, p->h_in, bytes, hipMemcpyHostToDevice, p->st), -1);
  if (run_dct(p, p->d_in, p->d_out, kind, p->st)) return -1;
  L2C_HIP(hipMemcpyAsync(p->h_out, p->d_out, bytes, hipMemcpyDeviceToHost, p->st), -1);
  L2C_HIP(hipStreamSynchronize(p->st), -1);
  std::memcpy(out, p->h_out, bytes);
  return 0;
}

}  // extern "C"
