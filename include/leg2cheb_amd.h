/*
 * leg2cheb_amd.h -- C ABI of the MI355X (gfx950) fast multipole
 * Legendre<->Chebyshev transform.
 *
 * Drop-in for the plan API of spectralDNS/Legendre-to-Chebyshev
 * (src/C/leg2cheb.h).  Every reference entry point keeps its name, argument
 * list and argument meaning; the plan structs are opaque (the reference's
 * own bindings already treat them as opaque pointers: src/cython/l2c.pyx:62,
 * src/ctypes/l2ctypes.py:11-39 never read a field).
 *
 * Semantics that differ from the reference, by design:
 *   - execute()/direct()/execute2D() OVERWRITE the output.  The reference's
 *     FMM path accumulates into it (src/C/leg2cheb.c:1475,1480) and every
 *     reference caller zeroes it first (src/cython/l2c.pyx:90,
 *     src/ctypes/l2ctypes.py:108, src/bin/l2c.c:189), so zeroed callers see
 *     identical results.
 *   - errors never abort: a failing call returns NULL or (size_t)-1 and
 *     l2c_last_error() describes the failure (the reference only asserts,
 *     src/C/leg2cheb.c:1312).
 *   - input and output may be host or device pointers (detected per call);
 *     host pointers are staged through the plan's pinned buffers.
 *   - a plan owns one HIP stream and its scratch; like the reference, one
 *     plan must not be executed from two threads at once.
 *
 * Extensions (l2c_*) expose device-resident execution, batched and strided
 * line execution, per-kernel timing and plan introspection.
 */
#ifndef LEG2CHEB_AMD_H
#define LEG2CHEB_AMD_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
/* The library is built with hidden visibility; exactly these are exported. */
#pragma GCC visibility push(default)

/* src/C/leg2cheb.h:27 */
enum { L2C = 0, C2L = 1, BOTH = 2 };

typedef struct fmm_plan_s fmm_plan;       /* src/C/leg2cheb.h:58-78 (opaque) */
typedef struct direct_plan_s direct_plan; /* src/C/leg2cheb.h:48-56 (opaque) */
typedef struct fmm_plan_2d_s fmm_plan_2d; /* src/C/leg2cheb.h:80-86 (opaque) */

/* ---- plan API: src/C/leg2cheb.h:88-105 ---------------------------------- */

/* Replaces create_fmm, src/C/leg2cheb.c:820-1234 (decl. leg2cheb.h:94-96).
 * Builds every table in HBM once: Lambda tables, the per-level Chebyshev
 * (lagrange=0) or Lagrange (lagrange=1) M2L matrices, T, the M2M/L2L
 * operators.  direction: L2C, C2L or BOTH.  precompute is accepted; the
 * device band forms the same coefficient products on the fly.  v > 1 prints
 * the plan summary like the reference. */
fmm_plan *create_fmm(const size_t N, const size_t maxs, const size_t M,
                     const size_t direction, const size_t lagrange,
                     const size_t precompute, const size_t v);

/* Replaces execute, src/C/leg2cheb.c:1299-1500 (decl. leg2cheb.h:104-105).
 * Elements are input_array[i*stride], output_array[i*stride], i < N.
 * Returns the reference's algorithmic flop count, (size_t)-1 on error.
 * Plans of N <= 8192 compute the transform as the direct O(N^2) sums in one
 * launch (faster there than the FMM's three; environment L2C_DIRECT_MAXN at
 * plan creation moves the bound, 0 keeps the FMM); l2c_execute_device and
 * the plan's graph replays do the same, and so do line calls of at most
 * 16,384 elements (environment L2C_DIRECT_LINES). */
size_t execute(const double *input_array, double *output_array,
               fmm_plan *fmmplan, size_t direction, const size_t stride);

/* Replaces free_fmm, src/C/leg2cheb.c:688-756 (decl. leg2cheb.h:88). */
void free_fmm(fmm_plan *plan);

/* Replaces create_direct / direct / free_direct, src/C/leg2cheb.c:758-818,
 * :152-234, :650-673 (decl. leg2cheb.h:90, :98-99, :93).  The O(N^2) sums. */
direct_plan *create_direct(size_t N, size_t direction, const size_t precompute);
size_t direct(const double *input_array, double *output_array,
              direct_plan *dplan, size_t direction, size_t stride);
void free_direct(direct_plan *dplan);

/* Replaces directM, src/C/leg2cheb.c:236-376 (decl. leg2cheb.h:100-101):
 * the L2C near-field band alone, with the final 2/pi and b_0/2 scaling. */
size_t directM(const double *input_array, double *output_array,
               fmm_plan *fmmplan, const size_t strides);

/* Replaces create_fmm_2d / execute2D / free_fmm_2d,
 * src/C/leg2cheb.c:1236-1297, :675-686 (decl. leg2cheb.h:89, :91-93, :102-103).
 * axis 0: along rows of an N0 x N1 row-major array (stride N1); axis 1: along
 * each row (stride 1); axis -1: axis 0 then axis 1.  Runs every line of an
 * axis in one batched launch. */
fmm_plan_2d *create_fmm_2d(size_t N0, size_t N1, int axis, size_t maxs,
                           size_t M, size_t direction, const size_t lagrange,
                           const size_t precompute, size_t v);
size_t execute2D(const double *input_array, double *output_array,
                 fmm_plan_2d *fmmplan2d, size_t direction);
void free_fmm_2d(fmm_plan_2d *plan);

/* Replaces matvectri, src/C/leg2cheb.c:537-619 (host helper; compact
 * triangular B(0) times the even/odd folded vector, as in the reference). */
void matvectri(const double *A, const double *x, double *b, const size_t m,
               const bool upper);

/* Plan geometry helpers, src/C/leg2cheb.c:125-150 (decl. leg2cheb.h:108-114). */
size_t get_number_of_blocks(const size_t level);
size_t get_total_number_of_blocks(const size_t L);
size_t get_h(const size_t level, const size_t L);
size_t get_number_of_submatrices(const size_t level);
size_t get_total_number_of_submatrices(const size_t L);
void get_ij(size_t *ij, const size_t level, const size_t block, const size_t s,
            const size_t L);

/* Lambda(z) = Gamma(z+1/2)/Gamma(z+1): src/C/lambda.c:65-98, :101-192,
 * :15-18, :20 (decl. leg2cheb.h:116-120).  Host functions; the device plan
 * kernels evaluate the same expressions. */
double _Lambda0(const double z);
double Lambda(const double z);
double LambdaI(const size_t z);
void Lambdavec(const double *z, double *w, const size_t N);

/* Replaces dct2, src/C/dct.c:582-598 (decl. leg2cheb.h:122): 18 x 18 2-D
 * DCT-II scaled by 4/18^2 with row 0 and column 0 halved (host helper). */
void dct2(double *input, double *output);

/* ---- extensions (no counterpart in the reference) ----------------------- */

/* Text of the last error on the calling thread ("" if none). */
const char *l2c_last_error(void);

/* Plan summary: N, Nn, L, s, M, number of submatrices, bytes of HBM held. */
typedef struct {
  size_t N, Nn, L, s, M, maxs, direction, lagrange, Nc, Nb;
  size_t hbm_bytes;    /* device memory held by the plan */
  size_t a_bytes;      /* bytes of one direction's M2L matrices */
  int subtree_up;      /* leaf chunks per upward work-group */
  int subtree_down;    /* leaf chunks per downward work-group */
} l2c_plan_info;
int l2c_get_plan_info(const fmm_plan *plan, l2c_plan_info *info);

/* Plan introspection: copy one of the plan's device tables into host memory
 * dst (room for `count` doubles) and return its length in doubles; dst = NULL
 * returns the length alone.  (size_t)-1 with an error text for an unknown
 * table, one the plan does not hold (an unplanned direction, the FMM tables
 * of a direct-path plan, P or Q of a tree without composite levels) or too
 * small a count.  No counterpart in the reference, whose plan struct is
 * public and whose CLI reads the same tables as fields (src/C/leg2cheb.h:58-78,
 * src/bin/l2c.c:644-660); tests compare them with the oracle's tables.
 *   L2C_TAB_LAMBDA  a[k] = Lambda(k), k < N            (src/C/leg2cheb.c:758-795)
 *   L2C_TAB_AN      the C2L diagonal scale, N           (:796-818)
 *   L2C_TAB_DN      the C2L band factor, ceil(N/2)      (:796-818)
 *   L2C_TAB_HEAD    the band head of `direction`: a (L2C) or dn (C2L) for
 *                   k < 2s, zero-padded to 2s + 16
 *   L2C_TAB_M2L     the M2L matrices of `direction`, Nc x M x M, level ->
 *                   block -> submatrix, row-major       (:924-995)
 *   L2C_TAB_T       2 x s x M [parity][row][k]          (:621-648, :1009-1031)
 *   L2C_TAB_U       U(0), U(1): child -> parent (modal U(0) = B(0)),
 *                   row-major M x M                     (:6-123, :1036-1105)
 *   L2C_TAB_D       D(p) = U(p)^T: parent -> child p
 *   L2C_TAB_P       P(k, d), k = 1..lu, d < 2^k, at (2^k - 2 + d) M^2:
 *                   P(1, q) = U(q), P(k, d) = U(d >> (k-1)) P(k-1, d mod 2^(k-1))
 *   L2C_TAB_Q       Q(k, d), k = 1..ld, d < 2^k, at (2^k - 2 + d) M^2,
 *                   stored column-major: Q(1, q) = D(q),
 *                   Q(k, d) = D(d & 1) Q(k-1, d >> 1) */
enum {
  L2C_TAB_LAMBDA = 0, L2C_TAB_AN = 1, L2C_TAB_DN = 2, L2C_TAB_HEAD = 3, L2C_TAB_M2L = 4,
  L2C_TAB_T = 5, L2C_TAB_U = 6, L2C_TAB_D = 7, L2C_TAB_P = 8, L2C_TAB_Q = 9
};
size_t l2c_get_plan_table(const fmm_plan *plan, int table, size_t direction, double *dst,
                          size_t count);

/* Device-resident execute: d_in, d_out are device pointers; runs on `stream`
 * (a hipStream_t, NULL = the plan's own stream) and returns without
 * synchronising.  Same overwrite semantics.  Returns 0 or -1. */
int l2c_execute_device(fmm_plan *plan, const double *d_in, double *d_out,
                       size_t direction, size_t stride, void *stream);

/* HIP-graph replay of l2c_execute_device (no counterpart in the reference).
 * With enable = 1 the plan captures its execute launch sequence into a HIP
 * graph on the first l2c_execute_device call for a given (d_in, d_out,
 * direction, stride) and replays it on later calls with the same arguments
 * (the launches are identical; replay saves host launch work).  Up to eight
 * argument sets are kept; further ones launch directly, as do the calls that
 * l2c_profile times (their launches carry the timing events).
 * enable = 0 releases the graphs (after synchronising the plan's stream).
 * Returns 0 or -1. */
int l2c_plan_graph(fmm_plan *plan, int enable);

/* Many lines of length N through one plan in one call (one launch sequence
 * per 65,535 lines; any nlines).  Line l (l < nlines) has first element at
 * (l / n0) * line_stride1 + (l % n0) * line_stride0 and its elements at that
 * offset + i * elem_stride.  Device pointers.  With n0 = nlines,
 * line_stride0 = distance between consecutive lines. */
int l2c_execute_lines_device(fmm_plan *plan, const double *d_in, double *d_out,
                             size_t direction, size_t nlines, size_t n0,
                             size_t line_stride0, size_t line_stride1,
                             size_t elem_stride, void *stream);

/* l2c_execute_lines_device for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, semantics and results;
 * the FMM workspace (multipoles and M2L vectors) is a fixed 32 MiB device
 * array instead of the plan's line workspace, so the call allocates nothing
 * (and times nothing).  Lines run as many per launch as the array holds,
 * launches in order on `stream`.  Returns -1 with an error text when one
 * line of the plan needs more workspace than the array holds (at M = 18 and
 * maxs = 64: N above 2^21).  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (another stream, another thread, another
 * plan) race and must be ordered by the caller. */
int l2c_execute_lines_device_noalloc(fmm_plan *plan, const double *d_in, double *d_out,
                                     size_t direction, size_t nlines, size_t n0,
                                     size_t line_stride0, size_t line_stride1,
                                     size_t elem_stride, void *stream);

/* Batches of independent transforms of uneven length (no counterpart in the
 * reference, whose callers loop over execute(); src/C/leg2cheb.c:1272-1293
 * is the same loop for the lines of a 2-D array).
 *
 * l2c_create_batch plans `count` transforms of lengths n[0..count).  Their
 * inputs and outputs are packed back to back in that order: transform t
 * occupies elements [off_t, off_t + n[t]) with off_t = n[0] + ... + n[t-1].
 * Transforms are grouped into size classes (n in (2^(k-1), 2^k]); each class
 * runs as one batched launch sequence of the direct O(N^2) sum or of the FMM
 * with leaf size s in {16, 32, 64} on the zero-padded length 2^k, so that a
 * class shares one set of tables.  tune = 0 picks the method per class by a
 * fixed rule (direct for 2^k <= 4 maxs with maxs = 64, else the FMM with
 * s = 64); tune = 1 times every candidate on the class's own lines at plan
 * time and keeps the fastest.  M is the rank (18 as in the reference).
 * direction: L2C, C2L or BOTH.  Returns NULL on error. */
typedef struct l2c_batch_plan_s l2c_batch_plan;
l2c_batch_plan *l2c_create_batch(const size_t *n, size_t count, size_t M, size_t direction,
                                 int tune);
void l2c_free_batch(l2c_batch_plan *bp);

/* Execute every transform of the batch: d_in, d_out are device pointers to the
 * packed arrays (overwrite semantics).  Runs on the batch's own streams and
 * returns without synchronising the host; a later call on `stream` (a
 * hipStream_t, NULL = the batch's main stream) or l2c_batch_synchronize()
 * sees the result.  With l2c_batch_graph(bp, 1) the launch sequence is
 * captured into a HIP graph on the first call for a given (d_in, d_out,
 * direction) and replayed by later calls.  Returns 0 or -1. */
int l2c_execute_batch_device(l2c_batch_plan *bp, const double *d_in, double *d_out,
                             size_t direction, void *stream);
/* Same with host arrays (staged through pinned buffers; synchronous). */
int l2c_execute_batch(l2c_batch_plan *bp, const double *in, double *out, size_t direction);
int l2c_batch_graph(l2c_batch_plan *bp, int enable);
int l2c_batch_synchronize(l2c_batch_plan *bp);

/* One size class of a batch plan: its padded length, the number of transforms
 * in it, the method chosen (0 direct, 1 FMM), the FMM leaf size s and depth
 * L, the tuner's time for the class in milliseconds (-1 if untuned), the
 * kernel variant (FMM: 0 fused downward pass, 1 split into band and far
 * field; direct: 0 dense kernel by line count, 1 one tile of all rows, 2 row
 * tiles) and whether the choice came from the tuning cache. */
typedef struct {
  size_t n_max, count;
  int method;
  size_t s, L;
  double ms;
  int variant;
  int cached;
} l2c_batch_class;
/* Tuning cache for l2c_create_batch(..., tune = 1): a text file holding the
 * tuner's choice per (device, size class, population, direction, M); classes
 * found there skip the timing, new choices are appended.  path = NULL turns
 * it off.  The environment variable L2C_TUNE_CACHE gives the default path.
 * No counterpart in the reference (its CLI keeps FFTW wisdom, src/bin/l2c.c:
 * 417-454).  Returns 0. */
int l2c_tune_cache(const char *path);
size_t l2c_batch_num_classes(const l2c_batch_plan *bp);
int l2c_batch_get_class(const l2c_batch_plan *bp, size_t i, l2c_batch_class *out);

/* Precision form of a plan.  0 = plain FP64 (default).  1 = split storage:
 * the M2L matrices are stored per coefficient as FP64, FP32, BF16 or dropped,
 * chosen against 2^-29, 2^-44 and 2^-53 of each block's largest coefficient;
 * all arithmetic stays FP64.  2 = split products (M = 18): the M2L products
 * of the blocks below the top levels run as BF16 slices on the matrix cores
 * (each operand scaled by powers of two and cut into 6 slices of 8 bits,
 * every anti-diagonal of slice pairs accumulated exactly in FP32,
 * recombined in FP64); the Lambda tables, the near-field band, the leaf
 * products, the L2L passes and the final accumulation stay FP64.  Form 2
 * builds its slice tables on first use (about 4.7x the bytes of the FP64
 * M2L matrices).  Any other value returns -1 with an error text. */
int l2c_set_precision(fmm_plan *plan, int form);

/* Length-N DCT-II and DCT-III of real FP64 vectors on the GPU: the steps
 * between function values at the N Chebyshev-Gauss points and Chebyshev
 * coefficients around L2C / C2L.  No counterpart in the reference library;
 * it replaces the FFTW REDFT10 the reference times as its yardstick
 * (src/bin/l2c.c:414-463, fftw_plan_r2r_1d(N, ..., FFTW_REDFT10, ...)).
 * Unnormalised, as FFTW:
 *   kind 2 (REDFT10): y_k = 2 sum_{n<N} x_n cos(pi k (2n+1) / (2N))
 *   kind 3 (REDFT01): y_k = x_0 + 2 sum_{1<=n<N} x_n cos(pi n (2k+1) / (2N))
 * so REDFT01(REDFT10(x)) = 2N x.  l2c_create_dct plans `count` vectors of
 * length N stored back to back (N a power of two, 1 <= N <= 2^25) and returns
 * NULL on error.  In-place calls (in == out) are allowed. */
typedef struct l2c_dct_plan_s l2c_dct_plan;
l2c_dct_plan *l2c_create_dct(size_t N, size_t count);
void l2c_free_dct(l2c_dct_plan *p);
/* d_in, d_out: device arrays of count * N doubles, 16-byte aligned.  Runs on
 * `stream` (a hipStream_t, NULL = the default stream) without synchronising
 * the host; calls that share a plan must be ordered on one stream, since they
 * share its workspace.  Returns 0 or -1. */
int l2c_execute_dct_device(l2c_dct_plan *p, const double *d_in, double *d_out, int kind,
                           void *stream);
/* Same with host arrays (staged through pinned buffers; synchronous). */
int l2c_execute_dct(l2c_dct_plan *p, const double *in, double *out, int kind);
/* Wait for all work on the plan's device. */
int l2c_dct_synchronize(l2c_dct_plan *p);
/* The DCT of `nlines` lines of length N (N <= 4096) in one launch, with the
 * line map of l2c_execute_lines_device: line l starts at
 * (l / n0) * line_stride1 + (l % n0) * line_stride0 and its elements are
 * elem_stride apart (the values <-> coefficients step along one axis of a
 * field).  The plan's count is not used.  Device pointers, in place allowed,
 * no synchronisation.  Returns 0 or -1. */
int l2c_execute_dct_lines_device(l2c_dct_plan *p, const double *d_in, double *d_out, int kind,
                                 size_t nlines, size_t n0, size_t line_stride0,
                                 size_t line_stride1, size_t elem_stride, void *stream);

/* Per-kernel timing of the execute kernels with HIP events.  Each timed
 * launch carries a start/stop event pair that the runtime fills from its own
 * dispatch (hipExtLaunchKernel), so a launch's time is its kernel's
 * execution, without launch gaps.  l2c_profile(plan, n) resets the sums and,
 * for n >= 1, times the kernels of every n-th call (execute, directM, lines);
 * n = 0 turns it off.  l2c_profile_read fills ms[8] (summed milliseconds)
 * and counts[8] (timed calls per slot; a slot that runs two kernels in one
 * call, as the M2L of precision form 2 does, sums both and counts one)
 * indexed 0 k_up, 1 k_m2l, 2 k_down, 3 k_direct (and the dense line
 * products), 4 k_far, 5 k_band, 6-7 unused. */
int l2c_profile(fmm_plan *plan, int enable);
int l2c_profile_read(fmm_plan *plan, double *ms, size_t *counts);

/* Synchronise the plan's stream. */
int l2c_synchronize(fmm_plan *plan);

/* Device memory helpers for hosts without their own allocator. */
void *l2c_malloc_device(size_t bytes);
int l2c_free_device(void *p);
int l2c_memcpy_h2d(void *dst, const void *src, size_t bytes);
int l2c_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* Device-to-device copy, complete when the call returns (a plain copy kernel
 * of known traffic: the bench scales its HBM counters by it, and a copy
 * larger than the 256 MB last-level cache evicts a plan's tables). */
int l2c_memcpy_d2d(void *dst, const void *src, size_t bytes);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif
#endif
