// l2c_exec.hip -- C ABI entry points of the gfx950 FMM Legendre<->Chebyshev
// library: execute / direct / directM / execute2D and the l2c_* extensions
// (src/C/leg2cheb.c:152-234, 236-376, 1267-1500), the reference's flop
// formulas, and per-launch HIP-event timing.  The kernels live in
// l2c_fmm.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "l2c_internal.h"

namespace l2c {

// ---------------------------------------------------------------------------
// per-launch event timing
void prof_step(Prof *pf) {
  if (!pf || !pf->on) return;
  pf->cur = pf->calls++ % (size_t)pf->on == 0;
}

void prof_begin(Prof *pf, int id, hipStream_t st) {
  (void)st;
  if (!pf || !pf->on || !pf->cur) return;
  pf->open = id;
  pf->brackets++;
}

void prof_end(Prof *pf, hipStream_t st) {
  (void)st;
  if (pf) pf->open = -1;
}

bool prof_events(Prof *pf, hipEvent_t *a, hipEvent_t *b) {
  if (!pf || !pf->on || !pf->cur || pf->open < 0) return false;
  if (pf->used + 2 > pf->pool.size()) {
    for (int i = 0; i < 256; i++) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return false;
      pf->pool.push_back(e);
    }
  }
  *a = pf->pool[pf->used];
  *b = pf->pool[pf->used + 1];
  pf->used += 2;
  pf->recs.push_back(Prof::Rec{pf->open, pf->brackets, *a, *b});
  return true;
}

static void prof_harvest(Prof *pf) {
  size_t last[K_COUNT];
  for (int i = 0; i < K_COUNT; i++) last[i] = 0;
  for (auto &r : pf->recs) {
    (void)hipEventSynchronize(r.b);
    float ms = 0;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      pf->ms[r.id] += ms;
      if (last[r.id] != r.bracket) pf->cnt[r.id]++;
      last[r.id] = r.bracket;
    }
  }
  pf->recs.clear();
  pf->used = 0;
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

// One execute through a cached HIP graph of its launch sequence: captured on
// the plan's stream by the first call with this key, replayed on st after.
// A call that l2c_profile times launches directly (its launches carry the
// timing events, which stay out of the graphs), as does a call whose key
// finds the cache full.  Returns 0 or -1.
static int run_fmm_graph(fmm_plan_s *p, const double *d_in, double *d_out, size_t dir,
                         size_t stride, hipStream_t st) {
  prof_step(p->prof);
  if ((p->prof && p->prof->on && p->prof->cur) ||
      (int)p->graphs.size() >= fmm_plan_s::kMaxGraphs) {
    bool cached = false;
    for (const auto &g : p->graphs)
      cached = cached || (g.in == d_in && g.out == d_out && g.dir == dir && g.stride == stride &&
                          g.ws == p->d_wl && g.precision == p->precision);
    if (!cached || (p->prof && p->prof->on && p->prof->cur))
      return run_fmm_launch(p, d_in, d_out, dir, stride, st, 0);
  }
  for (const auto &g : p->graphs)
    if (g.in == d_in && g.out == d_out && g.dir == dir && g.stride == stride && g.ws == p->d_wl &&
        g.precision == p->precision) {
      L2C_HIP(hipGraphLaunch(g.exec, st), -1);
      return 0;
    }
  hipStream_t cap = p->stream;
  L2C_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), -1);
  const int rc = run_fmm_launch(p, d_in, d_out, dir, stride, cap, 0);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(cap, &graph);
  if (rc || e != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    if (!rc) set_error(std::string("execute graph capture failed: ") + hipGetErrorString(e));
    return -1;
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    return -1;
  }
  p->graphs.push_back({d_in, d_out, dir, stride, p->d_wl, p->precision, exec});
  L2C_HIP(hipGraphLaunch(exec, st), -1);
  return 0;
}

void drop_graphs(fmm_plan_s *p) {
  if (p->graphs.empty()) return;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (auto &g : p->graphs) (void)hipGraphExecDestroy(g.exec);
  p->graphs.clear();
}

// Where the two arrays of a call live: 0 both host, 1 both device, -1 mixed
// (one host, one device), which no entry point accepts: the host-staged
// path would read or write device memory from the CPU.
static int pointer_kind(const void *in, const void *out, const char *fn) {
  const bool di = is_device_ptr(in), dout = is_device_ptr(out);
  if (di != dout) {
    set_error(std::string(fn) + ": input and output must both be device pointers or both host "
                                "pointers");
    return -1;
  }
  return di ? 1 : 0;
}

// Whether direction `dir` (0 or 1) was planned by create_fmm's direction
// argument (L2C, C2L or BOTH); the tables of an unplanned direction do not
// exist, so every entry point checks before any launch.
static bool dir_planned(const fmm_plan_s *p, size_t dir) {
  if (dir > 1) return false;
  if (p->L == 0) return p->dplan && (p->dplan->direction == BOTH || p->dplan->direction == dir);
  return p->t.A[dir] != nullptr;
}

// Host staging buffers (device copy and pinned host copy of N doubles each
// way), allocated on the first call with host arrays and kept by the plan.
static int ensure_staging(size_t N, double **din, double **dout, double **hin, double **hout) {
  if (*din && *dout && *hin && *hout) return 0;
  double *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr;
  if (hipMalloc((void **)&a, N * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&b, N * sizeof(double)) != hipSuccess ||
      hipHostMalloc((void **)&c, N * sizeof(double)) != hipSuccess ||
      hipHostMalloc((void **)&d, N * sizeof(double)) != hipSuccess) {
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (c) (void)hipHostFree(c);
    if (d) (void)hipHostFree(d);
    set_error("host staging allocation failed");
    return -1;
  }
  *din = a;
  *dout = b;
  *hin = c;
  *hout = d;
  return 0;
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
  if (!dir_planned(p, direction)) {
    set_error("execute: direction was not planned");
    return (size_t)-1;
  }
  const size_t N = p->N;
  hipStream_t st = p->stream;
  const int kind = pointer_kind(input_array, output_array, "execute");
  if (kind < 0) return (size_t)-1;
  int rc;
  if (kind == 1) {
    rc = p->L == 0 ? run_direct(p->dplan, input_array, output_array, direction, stride, st,
                                p->precompute)
                   : run_fmm(p, input_array, output_array, direction, stride, st, 0);
    if (rc) return (size_t)-1;
    L2C_HIP(hipStreamSynchronize(st), (size_t)-1);
  } else {
    direct_plan_s *dq = p->dplan;
    if (p->L == 0 ? ensure_staging(N, &dq->d_in, &dq->d_out, &dq->h_in, &dq->h_out)
                  : ensure_staging(N, &p->d_in, &p->d_out, &p->h_in, &p->h_out))
      return (size_t)-1;
    double *hin = p->L == 0 ? dq->h_in : p->h_in;
    double *hout = p->L == 0 ? dq->h_out : p->h_out;
    double *din = p->L == 0 ? dq->d_in : p->d_in;
    double *dout = p->L == 0 ? dq->d_out : p->d_out;
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
  const int kind = pointer_kind(input_array, output_array, "directM");
  if (kind < 0) return (size_t)-1;
  if (kind == 1) {
    if (run_fmm(p, input_array, output_array, L2C, strides, st, 1)) return (size_t)-1;
    L2C_HIP(hipStreamSynchronize(st), (size_t)-1);
  } else {
    if (ensure_staging(N, &p->d_in, &p->d_out, &p->h_in, &p->h_out)) return (size_t)-1;
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
  if (dp->direction != BOTH && dp->direction != direction) {
    set_error("direct: direction was not planned");
    return (size_t)-1;
  }
  const size_t N = dp->N;
  hipStream_t st = dp->stream;
  const int kind = pointer_kind(input_array, output_array, "direct");
  if (kind < 0) return (size_t)-1;
  if (kind == 1) {
    if (run_direct(dp, input_array, output_array, direction, stride, st, dp->precompute))
      return (size_t)-1;
    L2C_HIP(hipStreamSynchronize(st), (size_t)-1);
  } else {
    if (ensure_staging(N, &dp->d_in, &dp->d_out, &dp->h_in, &dp->h_out)) return (size_t)-1;
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
  if (!dir_planned(p, direction)) {
    set_error("l2c_execute_device: direction was not planned");
    return -1;
  }
  hipStream_t st = stream ? (hipStream_t)stream : p->stream;
  if (p->L == 0) return run_direct(p->dplan, d_in, d_out, direction, stride, st, p->precompute);
  if (p->graph_on) return run_fmm_graph(p, d_in, d_out, direction, stride, st);
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
  if (!dir_planned(p, direction)) {
    set_error("l2c_execute_lines_device: direction was not planned");
    return -1;
  }
  hipStream_t st = stream ? (hipStream_t)stream : p->stream;
  return run_lines(p, d_in, d_out, direction, nlines, n0, ls0, ls1, es, st);
}

int l2c_plan_graph(fmm_plan *p, int enable) {
  clear_error();
  if (!p) {
    set_error("l2c_plan_graph: NULL plan");
    return -1;
  }
  if (!enable) drop_graphs(p);
  p->graph_on = enable != 0;
  return 0;
}

int l2c_profile(fmm_plan *p, int enable) {
  if (!p) {
    set_error("l2c_profile: NULL plan");
    return -1;
  }
  if (!p->prof) p->prof = new Prof();
  if (p->prof->on) prof_harvest(p->prof);
  p->prof->on = enable > 0 ? enable : 0;
  p->prof->calls = 0;
  for (int i = 0; i < K_COUNT; i++) {
    p->prof->ms[i] = 0;
    p->prof->cnt[i] = 0;
  }
  if (p->dplan) p->dplan->prof = p->prof;
  return 0;
}

int l2c_profile_read(fmm_plan *p, double *ms, size_t *counts) {
  if (!p || !p->prof) {
    set_error("l2c_profile_read: profiling not enabled");
    return -1;
  }
  prof_harvest(p->prof);
  for (int i = 0; i < K_COUNT; i++) {
    if (ms) ms[i] = p->prof->ms[i];
    if (counts) counts[i] = p->prof->cnt[i];
  }
  return 0;
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
  if ((q->p0 && !dir_planned(q->p0, direction)) || (q->p1 && !dir_planned(q->p1, direction))) {
    set_error("execute2D: direction was not planned");
    return (size_t)-1;
  }
  hipStream_t st = pp->stream;
  const int kind = pointer_kind(input_array, output_array, "execute2D");
  if (kind < 0) return (size_t)-1;
  const bool dev = kind == 1;
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
