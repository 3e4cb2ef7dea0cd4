// l2c_batch.hip -- batches of independent transforms of uneven length.
//
// The reference transforms one vector per execute() call; a caller with many
// vectors loops (src/C/leg2cheb.c:1272-1293 does so for the lines of a 2-D
// array).  Here a batch plan sorts the transforms into size classes
// n in (2^(k-1), 2^k], gives each class one plan on the padded length 2^k
// (direct O(N^2) sum, or FMM with leaf size s in {16, 32, 64}), and runs every
// class as one batched launch sequence over its lines, each line reading
// zeros past its own end and storing only its own outputs.  Transforms of a
// class share the class plan's tables (Lambda, M2L matrices, T, operators).
//
// Padding to 2^k changes the FMM partition, not the transform: the L2C and
// C2L matrices are upper triangular, so rows i < n of the padded transform
// of the zero-extended vector are the length-n transform.
//
// Classes are spread over a few streams (default priority) that fork from and
// join back to the caller's stream; the whole sequence can be captured into
// one HIP graph.  The tuner times every candidate method on the class's own
// lines at plan time and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "l2c_internal.h"

namespace l2c {

struct BatchClass {
  int k = 0;             // n in (2^(k-1), 2^k]
  size_t n_max = 0;      // 2^k: the class plan's length
  std::vector<size_t> idx;  // transforms of the class, in batch order
  int method = 0;        // 0 direct, 1 FMM
  size_t s = 0, L = 0;   // FMM leaf size and depth
  // FMM: 0 fused downward kernel, 1 k_band + k_far; direct: 0 the default
  // dense kernel for the line count, 1 the one-tile kernel, 2 the tiled one
  int variant = 0;
  double ms = -1;        // tuner time of the chosen method
  int alt_method = -1;   // the tuner's runner-up (-1: none) and its time
  size_t alt_s = 0;
  int alt_variant = 0;
  double alt_ms = -1;
  bool cached = false;   // the choice came from the tuning cache
  double work = 0;       // estimated cost (orders classes over streams)
  int stream = 0;        // stream index the class runs on
  fmm_plan_s *fp = nullptr;
  direct_plan_s *dp = nullptr;
  size_t first = 0;      // index of the class's first line in d_off / d_n
};

constexpr int kBatchStreams = 4;

}  // namespace l2c

struct l2c_batch_plan_s {
  size_t count = 0, M = 0, direction = 0, total = 0;
  int device = 0;
  std::vector<size_t> n, off;
  std::vector<l2c::BatchClass> cls;
  size_t *d_off = nullptr;  // per line, class by class
  int *d_n = nullptr;
  hipStream_t st[l2c::kBatchStreams] = {};
  hipEvent_t fork = nullptr, join[l2c::kBatchStreams] = {};
  bool graph_on = false;
  hipGraphExec_t exec = nullptr;
  const double *g_in = nullptr;
  double *g_out = nullptr;
  size_t g_dir = 99;
  // classes on the tiled dense product run as one launch (L2C_BATCH_FUSE=0
  // launches them one by one)
  bool fuse_dense = true;
  // device buffers of the tuner (plan time only), then the staging of the
  // host-pointer form (allocated by its first call)
  double *d_in = nullptr, *d_out = nullptr;
  double *h_in = nullptr, *h_out = nullptr;
};

namespace l2c {

static void destroy_class_plans(BatchClass &c) {
  if (c.fp) {
    fmm_plan_s *p = c.fp;
    c.fp = nullptr;
    free_fmm(p);
  }
  if (c.dp) {
    direct_plan_s *d = c.dp;
    c.dp = nullptr;
    destroy_direct(d);
  }
}

static void destroy_batch(l2c_batch_plan_s *bp) {
  if (!bp) return;
  for (auto &s : bp->st)
    if (s) (void)hipStreamSynchronize(s);
  for (auto &c : bp->cls) destroy_class_plans(c);
  if (bp->exec) (void)hipGraphExecDestroy(bp->exec);
  if (bp->d_off) (void)hipFree(bp->d_off);
  if (bp->d_n) (void)hipFree(bp->d_n);
  if (bp->d_in) (void)hipFree(bp->d_in);
  if (bp->d_out) (void)hipFree(bp->d_out);
  if (bp->h_in) (void)hipHostFree(bp->h_in);
  if (bp->h_out) (void)hipHostFree(bp->h_out);
  if (bp->fork) (void)hipEventDestroy(bp->fork);
  for (auto &e : bp->join)
    if (e) (void)hipEventDestroy(e);
  for (auto &s : bp->st)
    if (s) (void)hipStreamDestroy(s);
  delete bp;
}

// Launch one class on stream st.
static int run_class(l2c_batch_plan_s *bp, const BatchClass &c, const double *d_in,
                     double *d_out, size_t dir, hipStream_t st) {
  return run_packed(c.fp, c.dp, d_in, d_out, dir, c.idx.size(), bp->d_off + c.first,
                    bp->d_n + c.first, st);
}

// Make the plan for one candidate of a class (method 0: direct with dense
// kernel variant `variant`, 1: FMM with leaf size s and downward variant
// `variant`).  Returns 0 or -1 (candidate not available).
static int make_candidate(l2c_batch_plan_s *bp, BatchClass &c, int method, size_t s,
                          int variant = 0) {
  c.method = method;
  c.s = s;
  c.variant = variant;
  c.L = 0;
  if (method == 0) {
    c.dp = make_direct(c.n_max, bp->direction, 0);
    if (c.dp) c.dp->dense_tile = variant;
    return c.dp ? 0 : -1;
  }
  c.fp = create_fmm_impl(c.n_max, s, bp->M, bp->direction, 0, 0, 0);
  if (!c.fp) return -1;
  if (c.fp->L < 1 || c.fp->s != s || !c.fp->geo.fast || reserve_lines(c.fp, c.idx.size())) {
    destroy_class_plans(c);
    return -1;
  }
  c.fp->down_split = variant;
  c.L = c.fp->L;
  return 0;
}

// Device buffers of the batch's size for the tuner's timings (zeros in).
static int ensure_tune_bufs(l2c_batch_plan_s *bp) {
  if (bp->d_in) return 0;
  const size_t bytes = std::max<size_t>(bp->total, 1) * sizeof(double);
  L2C_HIP(hipMalloc(&bp->d_in, bytes), -1);
  L2C_HIP(hipMalloc(&bp->d_out, bytes), -1);
  L2C_HIP(hipMemsetAsync(bp->d_in, 0, bytes, bp->st[0]), -1);
  return 0;
}

// Time the class's current plan on the tuning buffers (min of 3 after one
// warm-up), milliseconds; < 0 on failure.
static double time_class(l2c_batch_plan_s *bp, const BatchClass &c, size_t dir) {
  if (ensure_tune_bufs(bp)) return -1;
  hipStream_t st = bp->st[0];
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return -1;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return -1;
  }
  double best = -1;
  for (int r = 0; r < 4; r++) {
    (void)hipEventRecord(a, st);
    if (run_class(bp, c, bp->d_in, bp->d_out, dir, st)) {
      best = -1;
      break;
    }
    (void)hipEventRecord(b, st);
    if (hipEventSynchronize(b) != hipSuccess) {
      best = -1;
      break;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    if (r > 0 && (best < 0 || ms < best)) best = ms;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best;
}

// Candidates of size class k: the direct sum up to 2^13 (for 32 < 2^k <= 256
// with each dense kernel: one tile of all rows, or row tiles), the FMM with
// s = 16, 32, 64 where the depth L = k - log2(s) - 2 is at least 1, each with
// the fused and the split downward pass.
struct Cand {
  int method;
  size_t s;
  int variant;
};
static std::vector<Cand> candidates(int k) {
  std::vector<Cand> v;
  if (k <= 13) {
    if (k > 5 && k <= 8) {
      v.push_back({0, 0, 1});
      v.push_back({0, 0, 2});
    } else {
      v.push_back({0, 0, 0});
    }
  }
  for (int ls = 4; ls <= 6; ls++)
    if (k - ls - 2 >= 1) {
      v.push_back({1, (size_t)1 << ls, 0});
      v.push_back({1, (size_t)1 << ls, 1});
    }
  return v;
}

static int launch_batch(l2c_batch_plan_s *bp, const double *d_in, double *d_out, size_t dir,
                        hipStream_t origin);

// The whole batch, eager, on the staging buffers (min of 3 after one warm-up),
// milliseconds; < 0 on failure.
static double time_batch(l2c_batch_plan_s *bp, size_t dir) {
  if (ensure_tune_bufs(bp)) return -1;
  hipStream_t st = bp->st[0];
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return -1;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return -1;
  }
  double best = -1;
  for (int r = 0; r < 4; r++) {
    (void)hipEventRecord(a, st);
    if (launch_batch(bp, bp->d_in, bp->d_out, dir, st)) {
      best = -1;
      break;
    }
    (void)hipEventRecord(b, st);
    if (hipEventSynchronize(b) != hipSuccess) {
      best = -1;
      break;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    if (r > 0 && (best < 0 || ms < best)) best = ms;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best;
}

// Classes share the GPU on four streams, so a method that wins alone (a
// dense product keeping the matrix cores busy) can lose beside the others
// (latency-bound FMM work overlaps better).  For every class whose runner-up
// came within 25 % alone, time the whole batch with each and keep the faster.
static int refine_close_calls(l2c_batch_plan_s *bp, size_t dir) {
  for (auto &c : bp->cls) {
    if (c.alt_method < 0 || c.alt_ms < 0 || c.alt_ms > 1.25 * c.ms) continue;
    const double t_cur = time_batch(bp, dir);
    const int m0 = c.method, v0 = c.variant;
    const size_t s0 = c.s;
    destroy_class_plans(c);
    if (make_candidate(bp, c, c.alt_method, c.alt_s, c.alt_variant)) {
      clear_error();
      if (make_candidate(bp, c, m0, s0, v0)) return -1;
      continue;
    }
    const double t_alt = time_batch(bp, dir);
    if (t_cur >= 0 && t_alt >= 0 && t_alt < t_cur) {
      std::swap(c.ms, c.alt_ms);
      c.alt_method = m0;
      c.alt_s = s0;
      c.alt_variant = v0;
      continue;
    }
    destroy_class_plans(c);
    if (make_candidate(bp, c, m0, s0, v0)) return -1;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Tuning cache: the tuner's choice per (device, size class, class population,
// direction, rank), kept in a text file (l2c_tune_cache or L2C_TUNE_CACHE),
// one "key method s variant ms" line per choice, so that later batch plans
// of the same shapes skip the timing (the reference keeps FFTW wisdom for the
// same purpose, src/bin/l2c.c:417-454).
static std::mutex g_tc_mu;
static std::string g_tc_path;
static bool g_tc_path_set = false, g_tc_loaded = false;
static std::map<std::string, std::pair<Cand, double>> g_tc;

static std::string tune_cache_path() {
  if (!g_tc_path_set) {
    g_tc_path_set = true;
    if (const char *e = getenv("L2C_TUNE_CACHE")) g_tc_path = e;
  }
  return g_tc_path;
}

static std::string tune_key(const l2c_batch_plan_s *bp, const BatchClass &c, size_t dir) {
  static std::string dev;
  if (dev.empty()) {
    hipDeviceProp_t prop;
    int d = 0;
    (void)hipGetDevice(&d);
    if (hipGetDeviceProperties(&prop, d) == hipSuccess)
      dev = std::string(prop.gcnArchName) + ":" + std::to_string(prop.multiProcessorCount);
    else
      dev = "unknown";
    for (auto &ch : dev)
      if (ch == ' ') ch = '_';
  }
  // the population enters as its power of two (the choice moves with it)
  size_t pop = 1;
  while (pop < c.idx.size()) pop <<= 1;
  return dev + "/n" + std::to_string(c.n_max) + "/count" + std::to_string(pop) + "/dir" +
         std::to_string(dir) + "/M" + std::to_string(bp->M);
}

static void tune_cache_load() {  // under g_tc_mu
  if (g_tc_loaded) return;
  g_tc_loaded = true;
  const std::string path = tune_cache_path();
  if (path.empty()) return;
  std::ifstream f(path);
  std::string key;
  Cand cd{};
  double ms;
  while (f >> key >> cd.method >> cd.s >> cd.variant >> ms) g_tc[key] = {cd, ms};
}

static bool tune_cache_get(const std::string &key, Cand &cd, double &ms) {
  std::lock_guard<std::mutex> lock(g_tc_mu);
  if (tune_cache_path().empty()) return false;
  tune_cache_load();
  auto it = g_tc.find(key);
  if (it == g_tc.end()) return false;
  cd = it->second.first;
  ms = it->second.second;
  return true;
}

static void tune_cache_put(const std::string &key, const Cand &cd, double ms) {
  std::lock_guard<std::mutex> lock(g_tc_mu);
  const std::string path = tune_cache_path();
  if (path.empty()) return;
  tune_cache_load();
  g_tc[key] = {cd, ms};
  std::ofstream f(path, std::ios::app);
  if (f) f << key << ' ' << cd.method << ' ' << cd.s << ' ' << cd.variant << ' ' << ms << '\n';
}

// The untuned rule: direct while 2^k <= 4 maxs (maxs = 64: the reference's
// own switch to the direct path, src/C/leg2cheb.c:860-866), else FMM s = 64.
static std::pair<int, size_t> default_method(int k) {
  if (k <= 8) return {0, 0};
  return {1, 64};
}

static l2c_batch_plan_s *create_batch_impl(const size_t *n, size_t count, size_t M,
                                           size_t direction, int tune) {
  clear_error();
  if (!n || count == 0 || direction > 2) {
    set_error("l2c_create_batch: need lengths, count > 0 and direction in {0,1,2}");
    return nullptr;
  }
  if (M < 2 || M > 48) {
    set_error("l2c_create_batch: M must be in [2, 48]");
    return nullptr;
  }
  for (size_t t = 0; t < count; t++)
    if (n[t] == 0 || n[t] > ((size_t)1 << 24)) {
      set_error("l2c_create_batch: every length must be in [1, 2^24]");
      return nullptr;
    }
  l2c_batch_plan_s *bp = new l2c_batch_plan_s();
  bp->count = count;
  bp->M = M;
  bp->direction = direction;
  if (const char *e = getenv("L2C_BATCH_FUSE")) bp->fuse_dense = atoi(e) != 0;
  (void)hipGetDevice(&bp->device);
  bp->n.assign(n, n + count);
  bp->off.resize(count);
  size_t tot = 0;
  for (size_t t = 0; t < count; t++) {
    bp->off[t] = tot;
    tot += n[t];
  }
  bp->total = tot;
  for (int i = 0; i < kBatchStreams; i++) {
    if (hipStreamCreateWithFlags(&bp->st[i], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&bp->join[i], hipEventDisableTiming) != hipSuccess) {
      set_error("l2c_create_batch: stream or event creation failed");
      destroy_batch(bp);
      return nullptr;
    }
  }
  if (hipEventCreateWithFlags(&bp->fork, hipEventDisableTiming) != hipSuccess) {
    set_error("l2c_create_batch: event creation failed");
    destroy_batch(bp);
    return nullptr;
  }

  // size classes
  std::vector<int> kof(count);
  int kmax = 0;
  for (size_t t = 0; t < count; t++) {
    int k = 0;
    while (((size_t)1 << k) < n[t]) k++;
    kof[t] = k;
    kmax = std::max(kmax, k);
  }
  for (int k = 0; k <= kmax; k++) {
    BatchClass c;
    c.k = k;
    c.n_max = (size_t)1 << k;
    for (size_t t = 0; t < count; t++)
      if (kof[t] == k) c.idx.push_back(t);
    if (!c.idx.empty()) bp->cls.push_back(std::move(c));
  }
  // per-line offsets and lengths, class by class
  std::vector<size_t> hoff;
  std::vector<int> hn;
  for (auto &c : bp->cls) {
    c.first = hoff.size();
    for (size_t t : c.idx) {
      hoff.push_back(bp->off[t]);
      hn.push_back((int)n[t]);
    }
  }
  // the tuner times the candidates on device buffers of the batch's size
  // (zeros in, scratch out), made on its first timing and released once the
  // methods are chosen; the host staging of l2c_execute_batch is allocated by
  // its first call
  if (hipMalloc(&bp->d_off, count * sizeof(size_t)) != hipSuccess ||
      hipMalloc(&bp->d_n, count * sizeof(int)) != hipSuccess) {
    set_error("l2c_create_batch: allocation failed");
    destroy_batch(bp);
    return nullptr;
  }
  hipStream_t s0 = bp->st[0];
  L2C_HIP(hipMemcpyAsync(bp->d_off, hoff.data(), count * sizeof(size_t), hipMemcpyHostToDevice, s0),
          (destroy_batch(bp), nullptr));
  L2C_HIP(hipMemcpyAsync(bp->d_n, hn.data(), count * sizeof(int), hipMemcpyHostToDevice, s0),
          (destroy_batch(bp), nullptr));
  L2C_HIP(hipStreamSynchronize(s0), (destroy_batch(bp), nullptr));

  // method per class
  const size_t tdir = direction == C2L ? C2L : L2C;
  for (auto &c : bp->cls) {
    if (!tune) {
      const auto m = default_method(c.k);
      if (make_candidate(bp, c, m.first, m.second)) {
        if (m.first == 0 || make_candidate(bp, c, 0, 0)) {
          destroy_batch(bp);
          return nullptr;
        }
      }
      continue;
    }
    Cand hit{};
    double hit_ms = -1;
    if (tune_cache_get(tune_key(bp, c, tdir), hit, hit_ms)) {
      if (!make_candidate(bp, c, hit.method, hit.s, hit.variant)) {
        c.ms = hit_ms;
        c.cached = true;
        continue;
      }
      clear_error();
    }
    double best = -1, second = -1;
    Cand pick{-1, 0, 0}, alt{-1, 0, 0};
    for (const auto &m : candidates(c.k)) {
      if (make_candidate(bp, c, m.method, m.s, m.variant)) {
        clear_error();
        continue;
      }
      const double ms = time_class(bp, c, tdir);
      destroy_class_plans(c);
      if (ms < 0) continue;
      if (best < 0 || ms < best) {
        second = best;
        alt = pick;
        best = ms;
        pick = m;
      } else if (second < 0 || ms < second) {
        second = ms;
        alt = m;
      }
    }
    c.alt_method = alt.method;
    c.alt_s = alt.s;
    c.alt_variant = alt.variant;
    c.alt_ms = second;
    if (pick.method < 0 || make_candidate(bp, c, pick.method, pick.s, pick.variant)) {
      if (pick.method < 0) set_error("l2c_create_batch: no method ran for a size class");
      destroy_batch(bp);
      return nullptr;
    }
    c.ms = best;
  }

  // spread the classes over the streams, largest first, each to the least
  // loaded stream
  for (auto &c : bp->cls) {
    double w = 0;
    for (size_t t : c.idx) {
      const double nn = (double)c.n_max;
      (void)t;
      w += c.method == 0 ? nn * nn / 4 : 120.0 * nn;
    }
    c.work = c.ms > 0 ? c.ms : w;
  }
  std::vector<size_t> order(bp->cls.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  std::sort(order.begin(), order.end(),
            [&](size_t a, size_t b) { return bp->cls[a].work > bp->cls[b].work; });
  double load[kBatchStreams] = {0};
  for (size_t i : order) {
    int j = 0;
    for (int q = 1; q < kBatchStreams; q++)
      if (load[q] < load[j]) j = q;
    bp->cls[i].stream = j;
    load[j] += bp->cls[i].work;
  }
  std::vector<BatchClass> sorted;
  for (size_t i : order) sorted.push_back(std::move(bp->cls[i]));
  bp->cls = std::move(sorted);
  if (tune && refine_close_calls(bp, tdir)) {
    destroy_batch(bp);
    return nullptr;
  }
  if (tune)
    for (const auto &c : bp->cls)
      if (!c.cached) tune_cache_put(tune_key(bp, c, tdir), {c.method, c.s, c.variant}, c.ms);
  if (bp->d_in) {
    L2C_HIP(hipStreamSynchronize(s0), (destroy_batch(bp), nullptr));
    (void)hipFree(bp->d_in);
    (void)hipFree(bp->d_out);
    bp->d_in = bp->d_out = nullptr;
  }
  return bp;
}

// Fork from `origin`, run every class on its stream, join back to `origin`.
// The classes that take the tiled dense product run as one launch (on the
// stream of the first of them) when fuse_dense is set.
static int launch_batch(l2c_batch_plan_s *bp, const double *d_in, double *d_out, size_t dir,
                        hipStream_t origin) {
  L2C_HIP(hipEventRecord(bp->fork, origin), -1);
  for (int i = 0; i < kBatchStreams; i++)
    if (bp->st[i] != origin) L2C_HIP(hipStreamWaitEvent(bp->st[i], bp->fork, 0), -1);
  std::vector<DenseItem> fused;
  int fused_stream = -1;
  for (const auto &c : bp->cls) {
    if (bp->fuse_dense && c.dp && (int)fused.size() < 8 && dense_tiled(c.dp, dir, c.idx.size())) {
      fused.push_back({c.dp, bp->d_off + c.first, bp->d_n + c.first, c.idx.size()});
      if (fused_stream < 0) fused_stream = c.stream;
      continue;
    }
    if (run_class(bp, c, d_in, d_out, dir, bp->st[c.stream])) return -1;
  }
  if (!fused.empty() &&
      run_dense_classes(fused.data(), (int)fused.size(), d_in, d_out, dir, bp->st[fused_stream]))
    return -1;
  for (int i = 0; i < kBatchStreams; i++) {
    if (bp->st[i] == origin) continue;
    L2C_HIP(hipEventRecord(bp->join[i], bp->st[i]), -1);
    L2C_HIP(hipStreamWaitEvent(origin, bp->join[i], 0), -1);
  }
  return 0;
}

static int execute_batch_impl(l2c_batch_plan_s *bp, const double *d_in, double *d_out,
                              size_t dir, hipStream_t user) {
  if (!bp || !d_in || !d_out) {
    set_error("l2c_execute_batch_device: NULL argument");
    return -1;
  }
  if (dir > 1 || (bp->direction != BOTH && bp->direction != dir)) {
    set_error("l2c_execute_batch_device: direction not planned");
    return -1;
  }
  const hipStream_t origin = user ? user : bp->st[0];
  if (!bp->graph_on) return launch_batch(bp, d_in, d_out, dir, origin);
  if (!bp->exec || bp->g_in != d_in || bp->g_out != d_out || bp->g_dir != dir) {
    if (bp->exec) {
      L2C_HIP(hipStreamSynchronize(origin), -1);
      (void)hipGraphExecDestroy(bp->exec);
      bp->exec = nullptr;
    }
    hipStream_t cap = bp->st[0];
    L2C_HIP(hipStreamSynchronize(cap), -1);
    L2C_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), -1);
    const int rc = launch_batch(bp, d_in, d_out, dir, cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cap, &g);
    if (rc || e != hipSuccess || !g) {
      if (g) (void)hipGraphDestroy(g);
      if (!rc) set_error(std::string("graph capture failed: ") + hipGetErrorString(e));
      return -1;
    }
    const hipError_t ei = hipGraphInstantiate(&bp->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      bp->exec = nullptr;
      set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
      return -1;
    }
    bp->g_in = d_in;
    bp->g_out = d_out;
    bp->g_dir = dir;
  }
  L2C_HIP(hipGraphLaunch(bp->exec, origin), -1);
  return 0;
}

}  // namespace l2c

using namespace l2c;

extern "C" {

l2c_batch_plan *l2c_create_batch(const size_t *n, size_t count, size_t M, size_t direction,
                                 int tune) {
  return create_batch_impl(n, count, M, direction, tune);
}

void l2c_free_batch(l2c_batch_plan *bp) { destroy_batch(bp); }

int l2c_execute_batch_device(l2c_batch_plan *bp, const double *d_in, double *d_out,
                             size_t direction, void *stream) {
  clear_error();
  return execute_batch_impl(bp, d_in, d_out, direction, (hipStream_t)stream);
}

int l2c_execute_batch(l2c_batch_plan *bp, const double *in, double *out, size_t direction) {
  clear_error();
  if (!bp || !in || !out) {
    set_error("l2c_execute_batch: NULL argument");
    return -1;
  }
  const size_t bytes = bp->total * sizeof(double);
  hipStream_t st = bp->st[0];
  if (!bp->h_in) {  // host staging, on the first call with host arrays
    double *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr;
    if (hipMalloc((void **)&a, bytes) != hipSuccess || hipMalloc((void **)&b, bytes) != hipSuccess ||
        hipHostMalloc((void **)&c, bytes) != hipSuccess ||
        hipHostMalloc((void **)&d, bytes) != hipSuccess) {
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
      if (c) (void)hipHostFree(c);
      if (d) (void)hipHostFree(d);
      set_error("l2c_execute_batch: host staging allocation failed");
      return -1;
    }
    bp->d_in = a;
    bp->d_out = b;
    bp->h_out = d;
    bp->h_in = c;
  }
  std::memcpy(bp->h_in, in, bytes);
  L2C_HIP(hipMemcpyAsync(bp->d_in, bp->h_in, bytes, hipMemcpyHostToDevice, st), -1);
  if (execute_batch_impl(bp, bp->d_in, bp->d_out, direction, st)) return -1;
  L2C_HIP(hipMemcpyAsync(bp->h_out, bp->d_out, bytes, hipMemcpyDeviceToHost, st), -1);
  L2C_HIP(hipStreamSynchronize(st), -1);
  std::memcpy(out, bp->h_out, bytes);
  return 0;
}

int l2c_tune_cache(const char *path) {
  clear_error();
  std::lock_guard<std::mutex> lock(g_tc_mu);
  g_tc_path = path ? path : "";
  g_tc_path_set = true;
  g_tc_loaded = false;
  g_tc.clear();
  return 0;
}

int l2c_batch_graph(l2c_batch_plan *bp, int enable) {
  clear_error();
  if (!bp) {
    set_error("l2c_batch_graph: NULL plan");
    return -1;
  }
  bp->graph_on = enable != 0;
  if (!bp->graph_on && bp->exec) {
    (void)hipStreamSynchronize(bp->st[0]);
    (void)hipGraphExecDestroy(bp->exec);
    bp->exec = nullptr;
  }
  return 0;
}

int l2c_batch_synchronize(l2c_batch_plan *bp) {
  clear_error();
  if (!bp) {
    set_error("l2c_batch_synchronize: NULL plan");
    return -1;
  }
  for (auto &s : bp->st) L2C_HIP(hipStreamSynchronize(s), -1);
  return 0;
}

size_t l2c_batch_num_classes(const l2c_batch_plan *bp) { return bp ? bp->cls.size() : 0; }

int l2c_batch_get_class(const l2c_batch_plan *bp, size_t i, l2c_batch_class *out) {
  clear_error();
  if (!bp || !out || i >= bp->cls.size()) {
    set_error("l2c_batch_get_class: bad argument");
    return -1;
  }
  const BatchClass &c = bp->cls[i];
  out->n_max = c.n_max;
  out->count = c.idx.size();
  out->method = c.method;
  out->s = c.s;
  out->L = c.L;
  out->ms = c.ms;
  out->variant = c.variant;
  out->cached = c.cached ? 1 : 0;
  return 0;
}

}  // extern "C"
