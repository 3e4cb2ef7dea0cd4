// l2c_internal.h -- plan layout, geometry and the shared host/device Lambda
// of the gfx950 FMM Legendre<->Chebyshev library.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/leg2cheb_amd.h"

namespace l2c {

// ---------------------------------------------------------------------------
// errors: thread-local text, returned by l2c_last_error()
void set_error(const std::string &msg);
void clear_error();

#define L2C_HIP(call, ret)                                                   \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::l2c::set_error(std::string(#call) + ": " + hipGetErrorString(e_));   \
      return ret;                                                            \
    }                                                                        \
  } while (0)

// ---------------------------------------------------------------------------
// Lambda(z) = Gamma(z+1/2)/Gamma(z+1), same expressions on host and device
// (src/C/lambda.c:65-98).  Contraction is off so host and device round alike.
__host__ __device__ inline double lambda_series(double z) {
#pragma clang fp contract(off)
  const double zp = z + 0.25;
  const double sc = 1 / sqrt(zp);
  const double w = 1 / (zp * zp);
  // coefficients of the expansion in (z+1/4)^-2 (FMM_paper.pdf eq. 4.8)
  double y = 1 + (-1.0 / 64.0) * w;
  if (z > 2014) return y * sc;
  double p = w * w;
  y += (21.0 / 8192.0) * p;
  if (z > 141) return y * sc;
  p *= w;
  y += (-671.0 / 524288.0) * p;
  if (z > 40) return y * sc;
  p *= w;
  y += (180323.0 / 134217728.0) * p;
  if (z > 20) return y * sc;
  p *= w;
  y += (-20898423.0 / 8589934592.0) * p;
  if (z > 14) return y * sc;
  p *= w;
  y += (7426362705.0 / 1099511627776.0) * p;
  return y * sc;
}

// Below 10 the argument is raised by repeated +1 and the recursion
// Lambda(z) = Lambda(z+1)(z+1)/(z+1/2) unwound innermost first.
__host__ __device__ inline double lambda(double z) {
#pragma clang fp contract(off)
  int n = 0;
  double zz = z;
  while (zz < 10 && n < 64) {
    zz = zz + 1;
    n++;
  }
  double v = lambda_series(zz);
  // regenerate the same argument sequence from the bottom up
  for (int k = n - 1; k >= 0; k--) {
    double zk = z;
    for (int t = 0; t < k; t++) zk = zk + 1;
    v = v * (zk + 1) / (zk + 0.5);
  }
  return v;
}

// ---------------------------------------------------------------------------
// Plan geometry (src/C/leg2cheb.c:125-150, :860-871).
__host__ __device__ inline size_t nblocks(size_t level) { return ((size_t)2 << level) - 1; }
__host__ __device__ inline size_t total_blocks(size_t L) { return ((size_t)2 << L) - (L + 2); }
__host__ __device__ inline size_t total_sub(size_t L) { return 3 * total_blocks(L); }
// Chunk slots of the multipole store: level l holds 2^(l+2) chunks.
__host__ __device__ inline size_t slot_base(int level) { return ((size_t)4 << level) - 4; }

// Work-group barrier for LDS traffic only: it orders LDS reads and writes
// like __syncthreads() but does not wait for global loads and stores still
// in flight, so prefetches and result stores overlap the next phase.  Use it
// only where the threads exchange data through LDS alone.
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
struct LdsSync {
  __device__ __forceinline__ void operator()() const { lds_sync(); }
};

// Synchronisation among the waves of one work-group that run different code
// between two hardware barriers (wave roles), on LDS counters: a wave
// arrives by adding one (after its LDS writes are complete) and waits until
// a counter reaches a target.  The wait is bounded (2^22 polls, far beyond
// any real wait), so that a wave can never spin forever.  LDS-only fences,
// like lds_sync.
__device__ __forceinline__ void role_arrive(int *ctr) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  if ((threadIdx.x & 63) == 0)
    __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void role_wait(const int *ctr, int target) {
  for (int k = 0; k < (1 << 22); k++) {
    if (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= target) break;
    __builtin_amdgcn_s_sleep(1);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
// A barrier among the nw waves of one role: the k-th call returns once every
// wave of the role has made its k-th call.
struct RoleSync {
  int *ctr;
  int nw;
  int target;
  __device__ __forceinline__ void operator()() {
    target += nw;
    role_arrive(ctr);
    role_wait(ctr, target);
  }
};

// Line l of a batch starts at (l / n0) * ls1 + (l % n0) * ls0; its element i
// is at that offset + i * es.  A single transform is {1, 0, 0, stride}.
// Batches of uneven lengths set off (device array: start of line l) and n
// (device array: length of line l <= the plan's N); the kernels then read
// zeros past a line's end and store only its first n[l] outputs.
struct LineMap {
  size_t n0, ls0, ls1, es;
  const size_t *off = nullptr;
  const int *n = nullptr;
  // first line of this launch in the (n0, ls0, ls1) numbering, for batches
  // split over several launches (packed tables are offset by pointer instead)
  size_t l0 = 0;
};
__host__ __device__ inline size_t line_off(const LineMap &m, size_t l) {
  return m.off ? m.off[l] : ((m.l0 + l) / m.n0) * m.ls1 + ((m.l0 + l) % m.n0) * m.ls0;
}
__host__ __device__ inline int line_len(const LineMap &m, size_t l, int N) {
  return m.n ? m.n[l] : N;
}

struct Geo {
  int N, Nn, L, s, M;
  int nL;      // N / (2s): number of 2s-row near-field blocks
  int gu, lu;  // upward work-group: 2^gu leaf chunks; cut level lu = L-1-gu
  int gd, ld;  // downward work-group: 2^gd leaf chunks; cut level ld = L-1-gd
  int nodal;   // 1: Lagrange basis (lagrange=1), 0: Chebyshev basis
  int fast;    // 1: downward pass by k_down8 (s = 16, 32 or 64)
};

struct DevTables {
  double *a = nullptr;   // Lambda(k), k < N
  double *an = nullptr;  // C2L diagonal scale
  double *dn = nullptr;  // C2L band factor
  double *A[2] = {nullptr, nullptr};  // Nc x M x M per direction
  // the same blocks with each block's 3M rows across lanes (M <= 21): block
  // B's element (sub, row, col) at B 3 M^2 + col 3M + sub M + row, so that
  // lane sub M + row of a wave reads its whole row in M coalesced loads
  double *At[2] = {nullptr, nullptr};
  double *T = nullptr;   // 2 x s x M  [parity][row][k]
  double *U = nullptr;   // 2 x M x M  upward (child q -> parent), row-major
  double *D = nullptr;   // 2 x M x M  downward (parent -> child p), row-major
  double *P = nullptr;   // composite upward operators for the top levels
  double *Q = nullptr;   // composite downward operators for the ancestor chain
  // band head per direction: Lambda(k) (L2C) or dn(k) (C2L) for k < min(2s,
  // table length), zero up to 2s + 16
  double *head[2] = {nullptr, nullptr};
  // Split form of the M2L matrices (precision form 1), per direction: every
  // block (its three M x M submatrices, 3 M^2 coefficients) is one record of
  // rec_bytes: the coefficients of storage tier FP64 (n64), then FP32 (n32),
  // then BF16 (n16); the remaining coefficient positions are dropped (zero).
  // pos[i] is the padded LDS index (row (M+1) + column) of record slot i, the
  // dropped positions last.  Tiers are per coefficient position, chosen at
  // plan time from the largest magnitude relative to its block's largest.
  struct Split {
    unsigned char *rec = nullptr;
    unsigned short *pos = nullptr;
    int n64 = 0, n32 = 0, n16 = 0, rec_bytes = 0;
  } sp[2];
  // Sliced form of the M2L matrices (precision form 2), per direction, for
  // the blocks at levels >= lu: each block's stacked operator E (36 x 36:
  // rows 0..17 [A0 A1], rows 18..35 [0 A2]) with rows scaled by powers of
  // two sc (48 per block, rows past 35 zero) and cut into kSlices BF16
  // slices of 8 bits each, stored as MFMA A-operand fragments
  // [row tile 3][k-step 2][slice][lane 64][8 bf16].
  struct Slices {
    unsigned short *q = nullptr;
    double *sc = nullptr;
  } sl[2];
};
constexpr int kSlices = 6;

// Per-launch HIP-event timing of the execute kernels (bench support).
enum KernelId {
  K_UP_SUBTREE = 0, K_M2L = 1, K_DOWN_BAND = 2, K_DIRECT = 3, K_FAR = 4, K_BAND = 5, K_COUNT = 8
};
struct Prof {
  int on = 0;       // timing period in calls (0: off)
  size_t calls = 0;  // calls since enabled
  int cur = 0;      // the current call is timed
  int open = -1;     // kernel slot of the open bracket (-1: none)
  size_t brackets = 0;
  struct Rec {
    int id;
    size_t bracket;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  double ms[K_COUNT] = {0};
  size_t cnt[K_COUNT] = {0};
};
void prof_step(Prof *pf);  // at the start of each call
// A bracket around the launches of one kernel slot: every launch inside it
// adds its execution time to the slot, and the bracket counts one launch.
void prof_begin(Prof *pf, int id, hipStream_t st);
void prof_end(Prof *pf, hipStream_t st);
// A start/stop event pair for the next launch when the current call is
// timed and a bracket is open; false otherwise.
bool prof_events(Prof *pf, hipEvent_t *a, hipEvent_t *b);

// Launch kern on st.  Inside an open bracket of a timed call the launch
// carries the start/stop events that the runtime fills from the dispatch
// itself (hipExtLaunchKernel), so the recorded time is the kernel's
// execution alone, without the dispatch gaps that events recorded between
// launches would add.
template <typename... KArgs, typename... Args>
inline void launch_k(Prof *pf, void (*kern)(KArgs...), dim3 grid, dim3 block, size_t shm,
                     hipStream_t st, Args... args) {
  hipEvent_t a, b;
  if (prof_events(pf, &a, &b))
    hipExtLaunchKernelGGL(kern, grid, block, (std::uint32_t)shm, st, a, b, 0u, args...);
  else
    kern<<<grid, block, shm, st>>>(args...);
}
}  // namespace l2c

struct direct_plan_s {
  size_t N = 0, direction = 0, precompute = 0;
  int device = 0;
  double *d_a = nullptr, *d_an = nullptr, *d_dn = nullptr;
  // dense parity blocks for batched lines (N <= kGemmMaxN): per direction
  // 2 x NPP x NPP, NPP = ceil(N/2) rounded up to 16
  double *d_K[2] = {nullptr, nullptr};
  int NPP = 0;
  // dense kernel for line batches: 0 by the line count, 1 one tile of all
  // rows (k_lines_dense), 2 row tiles (k_lines_gemm); the batch tuner sets it
  int dense_tile = 0;
  // lines per grid row of k_direct (grid.y limit 65535; L2C_DIRECT_LINES_Y
  // sets a smaller value, which the tests use to reach grid.z > 1 cheaply)
  unsigned lines_y = 65535;
  // single transforms of N > 3072 on k_direct_band, else k_direct_rows
  // (L2C_DIRECT_BAND=0: always k_direct_rows)
  bool direct_band = true;
  double *d_in = nullptr, *d_out = nullptr;  // staging (device)
  double *h_in = nullptr, *h_out = nullptr;  // staging (pinned host)
  hipStream_t stream = nullptr;
  l2c::Prof *prof = nullptr;
};

struct fmm_plan_s {
  size_t N = 0, maxs = 0, M = 0, direction = 0, lagrange = 0, precompute = 0;
  size_t Nn = 0, L = 0, s = 0, Nc = 0, Nb = 0;
  int device = 0;
  // 0 plain FP64; 1 split storage of the M2L matrices; 2 the M2L products of
  // the blocks at levels >= lu as BF16 slices on MFMA (FP32 accumulation,
  // FP64 recombination)
  int precision = 0;
  // M2L of levels >= lu on FP64 MFMA for calls of at least this many lines
  // (0: never; from 32 lines on the MFMA form is faster: the uneven batch of
  // 4096 transforms 14.0 -> 14.9 M/s, while 8 lines of 2^20 lose 244 -> 259
  // us); L2C_M2L_MMA sets it at plan creation
  int m2l_mma_lines = 32;
  // downward pass of s = 16/32/64 plans: 0 (default) one fused kernel, 1
  // k_band + k_far, 2 k_band + k_far for calls of several lines with L >= 10;
  // L2C_DOWN_SPLIT
  int down_split = 0;
  // the fused downward pass with wave roles (k_down8r: tree and far field on
  // half the waves, the band on the other half; default) or without (0:
  // k_down8); L2C_DOWN_ROLES
  int down_roles = 1;
  // band columns moved from k_down8r's band role to its tree role (negative:
  // the other way); L2C_ROLES_SHIFT
  int roles_shift = -32;
  // single transforms (execute, l2c_execute_device) of N up to this run as
  // one direct O(N^2) launch (k_direct_rows) on the plan's Lambda tables
  // instead of the three FMM launches, whose dependent chains set a floor of
  // about 22 us whatever the size; 0: always the FMM; L2C_DIRECT_MAXN
  int direct_max_n = 8192;
  // line calls of such plans of at most this many elements (lines x N) take
  // the same direct sums, all lines in one launch; L2C_DIRECT_LINES
  int direct_line_elems = 16384;
  // M2L of the blocks below the top levels from At (rows in registers, no
  // LDS staging: k_m2l_t) for the VALU M2L; L2C_M2L_T=0: k_m2l from A
  bool m2l_t = true;
  // single direct transforms of N > 3072 on k_direct_band (the band's
  // register rings over LDS windows) rather than k_direct_rows (three loads
  // per term); L2C_DIRECT_BAND=0 selects k_direct_rows
  bool direct_band = true;
  // work-group geometry of line calls (geo) and of single-line calls (geo1):
  // the same tree, cut into subtrees of different size; the composite
  // operators P and Q hold levels 1..lu_tab and 1..ld_tab, enough for both
  l2c::Geo geo{};
  l2c::Geo geo1{};
  int lu_tab = 0, ld_tab = 0;
  l2c::DevTables t;
  direct_plan_s *dplan = nullptr;  // L < 1: direct path only
  double *d_w = nullptr;           // multipole store: slots x 2 x M
  double *d_wl = nullptr;          // multipole store for line batches
  size_t wl_cap = 0;               // doubles held by d_wl
  // stores replaced by a larger one; a launch on another stream may still
  // use them, so they are released only when the plan is freed
  std::vector<double *> wl_retired;
  // host staging, allocated by the first call with host arrays
  double *d_in = nullptr, *d_out = nullptr;  // staging (device, contiguous)
  double *h_in = nullptr, *h_out = nullptr;  // staging (pinned host)
  size_t hbm_bytes = 0;
  hipStream_t stream = nullptr;
  l2c::Prof *prof = nullptr;
  // HIP graphs of the execute launch sequence (l2c_plan_graph): one per
  // (input, output, direction, stride, workspace, precision form), at most
  // kMaxGraphs; further keys run the launches directly
  struct Graph {
    const double *in;
    double *out;
    size_t dir, stride;
    const double *ws;
    int precision;
    hipGraphExec_t exec;
  };
  static constexpr int kMaxGraphs = 8;
  bool graph_on = false;
  std::vector<Graph> graphs;
};

struct fmm_plan_2d_s {
  fmm_plan_s *p0 = nullptr, *p1 = nullptr;
  size_t N0 = 0, N1 = 0;
  int axis = 0;
  double *d_in = nullptr, *d_tmp = nullptr, *d_out = nullptr;
};

namespace l2c {
// batched lines of the direct path up to this length run as dense MFMA products
constexpr int kDenseMaxN = 256;
// direct plans up to this length build the dense parity blocks; batched lines
// with kDenseMaxN < N <= kGemmMaxN run them as a tiled product
constexpr int kGemmMaxN = 4096;
// implemented in l2c_exec.hip
int run_fmm(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir, size_t stride,
            hipStream_t st, int band_only);
// the same without advancing the timing period (the caller did)
int run_fmm_launch(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir, size_t stride,
                   hipStream_t st, int band_only);
int run_direct(const direct_plan_s *dp, const double *d_in, double *d_out, size_t dir,
               size_t stride, hipStream_t st, size_t precompute);
int run_lines(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir, size_t nlines,
              size_t n0, size_t ls0, size_t ls1, size_t es, hipStream_t st);
int run_packed(fmm_plan_s *p, const direct_plan_s *dp, const double *d_in, double *d_out,
               size_t dir, size_t nlines, const size_t *d_off, const int *d_n, hipStream_t st);
int reserve_lines(fmm_plan_s *p, size_t nlines);
// packed lines of one direct plan for a fused dense launch
struct DenseItem {
  const direct_plan_s *dp;
  const size_t *off;
  const int *n;
  size_t nlines;
};
bool dense_tiled(const direct_plan_s *dp, size_t dir, size_t nlines);
int run_dense_classes(const DenseItem *items, int count, const double *d_in, double *d_out,
                      size_t dir, hipStream_t st);
int build_slices(fmm_plan_s *p, int dir);  // precision form 2 tables of one direction
void drop_graphs(fmm_plan_s *p);  // synchronises the plan's stream first
size_t flops_fmm(const fmm_plan_s *p, size_t dir);
size_t down_shared_bytes(const Geo &g);  // of the downward kernel g selects
size_t band_shared_bytes(const Geo &g);
size_t far_shared_bytes(const Geo &g);
bool down8_eligible(const Geo &g);
size_t up_shared_bytes(const Geo &g);
size_t ws_doubles(const Geo &g);  // execute workspace per line
size_t flops_direct(size_t N, size_t dir, size_t precompute);
// implemented in l2c_plan.hip
direct_plan_s *make_direct(size_t N, size_t direction, size_t precompute);
fmm_plan_s *create_fmm_impl(size_t N, size_t maxs, size_t M, size_t direction, size_t lagrange,
                            size_t precompute, size_t v);
void destroy_direct(direct_plan_s *dp);
}  // namespace l2c
