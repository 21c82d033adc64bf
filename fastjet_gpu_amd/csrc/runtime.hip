/*
 * runtime.hip — host side of libfjmi: the C ABI declared in include/fjmi.h.
 *
 * Holds the per-device context (preallocated device buffers, the multiplicity
 * bucket table, kernel attributes), the one-event-per-launch path, the batched
 * path (bucketed persistent queues captured into one hipGraph) and the
 * host-memory conveniences used by tests and tools.  Replaces the reference's
 * per-call allocation and launch sequences (tri_matrix.cu:286-336,
 * n_array.cu:229-243, grid.cu:580-640) and its launch-shape helpers
 * (launch.h:11-60).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fjmi.h"
#include "fjmi_internal.h"
#include "fjmi_kernels.h"
#include "fjmi_runtime.h"

using fjmi::KernelDesc;
using fjmi::LaunchArgs;

thread_local std::string g_err;

int fjmi::rt::fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

namespace {

using fjmi::rt::DevBuf;
using fjmi::rt::fail;

const fjmi_bucket kDefaultBuckets[] = {
#include "tune_table.inc"
};

const KernelDesc *find_kernel(int T, int K, int p4g) {
  for (int i = 0; i < fjmi::kNumDenseKernels; ++i) {
    const KernelDesc &d = fjmi::kDenseKernels[i];
    if (d.T == T && d.K == K && (int)d.p4g == p4g) return &d;
  }
  return nullptr;
}

size_t dyn_lds_bytes(const KernelDesc &d) {
  const size_t cap = (size_t)d.T * d.K;
  const size_t ch = d.p4g ? (size_t)fjmi::kStageChunk : cap;
  return 32 * ch + 4 * cap;
}

}  // namespace

/* Scratch of fjmi_cluster_event_device_noalloc: the tiled engine's per-event
 * scratch for up to kNoAllocMaxN particles, then one status word.  Each device
 * holds its own copy for the life of the process; host code never uses this
 * name as an address (it is a host-side stand-in): fjmi_create fetches each
 * device's address with hipGetSymbolAddress and the entry point passes that. */
namespace {
constexpr int kNoAllocMaxN = 20000;
constexpr size_t kNoAllocTiledBytes =
    (fjmi::tiled_scratch_bytes(kNoAllocMaxN, 1, fjmi::kTiledPoolPerEvent) + 255) / 256 * 256;
constexpr size_t kNoAllocBytes = kNoAllocTiledBytes + 256;
}  // namespace
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];


struct BucketRun {
  bool tiled;
  fjmi::TiledScratch ts;
  int threads;
  const void *fn;
  const KernelDesc *kd;
  int grid;
  size_t lds;
  LaunchArgs args;
  int n_events;
  int max_n;
  int tile_q;
  int tile_table;
  int64_t particles;
  double sum_n2;
};

/* The device side of one batch of events: multiplicity buckets, their
 * longest-first queues and scratch, and one persistent-kernel launch per
 * bucket (BucketRun).  Built by make_plan; launched inside a captured graph
 * (fjmi_batch_*) or directly on a stream (the host-memory pipeline). */
struct Plan {
  fjmi::rt::DevBuf queue, heads, p4s, coords, tscratch, times;
  std::vector<BucketRun> runs;
  std::vector<int32_t> flat; /* the queues, host copy */
  uint32_t *d_status = nullptr;
  int nb = 0;
  void release() {
    for (fjmi::rt::DevBuf *b : {&queue, &heads, &p4s, &coords, &tscratch, &times}) b->release();
    runs.clear();
  }
};

/* Host-memory batches (fjmi_cluster_batch): chunks of events in flight, each
 * in a slot with its own stream, device buffers and pinned staging, kept by
 * the context for the next call. */
constexpr int kHostSlots = 4;
constexpr int kHostMaxChunks = 4;                 /* chunks per call, at most */
constexpr int64_t kHostChunkParticles = 1 << 18;  /* and at least this many particles each */

struct HostSlot {
  Plan plan;
  fjmi::rt::DevBuf d_in, d_off, d_p, d_R, d_h1, d_h2, d_hd, d_jets, d_nj, d_joff, d_jc, d_defs, d_hj;
  char *pin = nullptr; /* pinned staging */
  size_t pin_cap = 0;
  int32_t *h_h1 = nullptr, *h_h2 = nullptr, *h_nj = nullptr, *h_joff = nullptr;
  double *h_hd = nullptr, *h_jc = nullptr, *h_hj = nullptr; /* h_hj: history momenta, or null */
  hipStream_t s = nullptr;
  std::vector<hipStream_t> fork; /* further buckets' kernels */
  std::vector<hipEvent_t> ev_join;
  hipEvent_t ev_fork = nullptr, ev_out = nullptr;
  std::vector<int32_t> events; /* the chunk's events (caller's indices) */
  std::vector<int64_t> off;    /* their particle offsets in the chunk */
  int n_events = 0;
  int64_t particles = 0;
  uint32_t *h_status = nullptr;
  bool busy = false;
  void release() {
    if (s) (void)hipStreamSynchronize(s);
    plan.release();
    for (fjmi::rt::DevBuf *b :
         {&d_in, &d_off, &d_p, &d_R, &d_h1, &d_h2, &d_hd, &d_jets, &d_nj, &d_joff, &d_jc, &d_defs, &d_hj})
      b->release();
    if (pin) (void)hipHostFree(pin);
    pin = nullptr;
    pin_cap = 0;
    for (hipStream_t f : fork) (void)hipStreamDestroy(f);
    for (hipEvent_t j : ev_join) (void)hipEventDestroy(j);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_out) (void)hipEventDestroy(ev_out);
    if (s) (void)hipStreamDestroy(s);
    fork.clear();
    ev_join.clear();
    s = nullptr;
    ev_fork = ev_out = nullptr;
  }
};

struct HostPath {
  HostSlot slot[kHostSlots];
};

struct fjmi_batch {
  fjmi_context *ctx = nullptr;
  int n_events = 0;
  Plan plan;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  /* timing events recorded inside the graph: [0] after the queue reset, [i+1] after bucket i */
  std::vector<hipEvent_t> tev;
};

namespace {

int set_device(fjmi_context *ctx) {
  HIPCHK(hipSetDevice(ctx->device));
  return FJMI_OK;
}

int check_table(const fjmi_bucket *rows, int n) {
  if (n <= 0) return fail(FJMI_EINVAL, "empty bucket table");
  for (int i = 0; i < n; ++i) {
    const fjmi_bucket &b = rows[i];
    if (i > 0 && b.max_n <= rows[i - 1].max_n) return fail(FJMI_EINVAL, "bucket table not sorted by max_n");
    if (b.engine == FJMI_ENGINE_TILED) {
      if (b.threads != 64 && b.threads != fjmi::kTiledCoopThreads)
        return fail(FJMI_EINVAL, "tiled rows run 64 threads (one wave per event) or %d (%d waves per event)",
                    fjmi::kTiledCoopThreads, fjmi::kTiledCoopThreads / 64);
      if (b.slots != 0 && (b.slots < 4 || b.slots > 16))
        return fail(FJMI_EINVAL, "tiled rows: slots is the tile side in quarters of R, 4..16 (0 = 4)");
      if (b.p4_global < 0 || b.p4_global >= fjmi::kNumTiledKernels)
        return fail(FJMI_EINVAL, "tiled rows: p4_global selects the tile table size, 0..%d", fjmi::kNumTiledKernels - 1);
      continue;
    }
    if (b.engine != FJMI_ENGINE_DENSE) return fail(FJMI_EINVAL, "bucket engine must be dense or tiled");
    const KernelDesc *kd = find_kernel(b.threads, b.slots, b.p4_global);
    if (!kd)
      return fail(FJMI_EINVAL, "no compiled dense kernel with threads=%d slots=%d p4_global=%d", b.threads, b.slots,
                  b.p4_global);
    if ((int64_t)b.threads * b.slots < b.max_n)
      return fail(FJMI_EINVAL, "bucket max_n=%d exceeds capacity %d", b.max_n, b.threads * b.slots);
  }
  return FJMI_OK;
}

/* The row serving an event of n particles: the first row with max_n >= n whose
 * engine matches `engine` (any engine for FJMI_ENGINE_AUTO). */
const fjmi_bucket *bucket_for(const fjmi_context *ctx, int n, int engine) {
  for (const fjmi_bucket &b : ctx->buckets)
    if (n <= b.max_n && (engine == FJMI_ENGINE_AUTO || b.engine == engine)) return &b;
  return nullptr;
}

const fjmi_bucket kTiledRow = {0x7fffffff, FJMI_ENGINE_TILED, 64, 0, 0};

/* A dense row for an event no dense table row covers (engine forced to dense):
 * the compiled kernel of capacity >= n with the fewest slots per thread (the
 * shortest per-step scan), then the smallest capacity, then LDS four-momenta. */
bool dense_row_for(int n, fjmi_bucket *out) {
  const KernelDesc *best = nullptr;
  for (int i = 0; i < fjmi::kNumDenseKernels; ++i) {
    const KernelDesc &d = fjmi::kDenseKernels[i];
    if ((int64_t)d.T * d.K < n) continue;
    if (!best || d.K < best->K || (d.K == best->K && d.T < best->T) ||
        (d.K == best->K && d.T == best->T && !d.p4g && best->p4g))
      best = &d;
  }
  if (!best) return false;
  *out = fjmi_bucket{best->T * best->K, FJMI_ENGINE_DENSE, best->T, best->K, best->p4g ? 1 : 0};
  return true;
}

/* The engine that finishes one event alone soonest.  The dense engine's step
 * costs grow with N (a scan of every slot) and not with R; the tiled engine's
 * with the jets in a 3x3 tile neighbourhood (about 0.14 N R^2 for tiles of side
 * R over |y| < 5), on top of a larger fixed cost per step.  Measured on MI355X,
 * one event alone, both engines at their best shapes, N from 100 to 8000:
 *   R = 0.2: dense up to N = 354, a tie at 500, tiled from 808;
 *   R = 0.3, 0.4, 0.45: dense up to N = 1000, tiled at 1500 and 2000;
 *   R = 0.5: tiled 2 % ahead at N = 1500, dense ahead at 2000 and 5194;
 *   R >= 0.6: dense at every N it holds (up to 8192).
 * So: dense up to N_c(R), with N_c = 500 at R <= 0.2 rising linearly to 1200 at
 * R = 0.3 and staying there, and dense at every N from R = 0.5 (no case measured
 * more than 2 % from the faster engine).  Both engines give the same results
 * bit for bit. */
bool latency_prefers_dense(int n, double R) {
  if (R >= 0.5) return true;
  const double nc = R <= 0.2 ? 500.0 : (R >= 0.3 ? 1200.0 : 500.0 + 7000.0 * (R - 0.2));
  return (double)n <= nc;
}

/* Whether one event alone runs faster on the cooperative kernel (8 waves share
 * gathers of several chunks) than on one wave: its 3x3 neighbourhoods hold
 * about 0.14 N R^2 jets.  Measured on MI355X, anti-kt, 384 tiles: at R = 0.4,
 * 0.96x at N = 2000, 1.07x at 9278, 1.17x at 20000; at R = 0.8, 0.98x at
 * N = 2000, 1.33x at 9278, 1.67x at 20000.  So: N R^2 > 1200. */
bool latency_prefers_coop(int n, double R) { return (double)n * R * R > 1200.0; }

/* The tiled row for one event alone (tile side R): the 384-tile variant, the
 * shortest step at 3 waves' registers, except for small R and large N, where
 * the 512-tile variant's finer grid and full register budget win (measured at
 * R = 0.2: 9.5 against 9.8 ms at N = 2000, 37.5 against 41.6 ms at N = 5194; at
 * R = 0.4 the two are within 1 % or the 384-tile variant is ahead).  The
 * default table's variants hold more events per CU, which a batch needs and a
 * single event does not. */
fjmi_bucket latency_tiled_row(int n, double R) {
  return fjmi_bucket{0x7fffffff, FJMI_ENGINE_TILED, latency_prefers_coop(n, R) ? fjmi::kTiledCoopThreads : 64, 0,
                     (R < 0.3 && n > 2000) ? 0 : 1};
}

LaunchArgs base_args() {
  LaunchArgs a;
  std::memset(&a, 0, sizeof a);
  a.in_stride = 4;
  a.in_off = 0;
  a.stage_chunk = fjmi::kStageChunk;
  return a;
}

int launch_dense(const KernelDesc &kd, int grid, size_t lds, const LaunchArgs &a, hipStream_t stream) {
  LaunchArgs copy = a;
  void *params[] = {&copy};
  HIPCHK(hipLaunchKernel(kd.fn, dim3(grid), dim3(kd.T), params, lds, stream));
  return FJMI_OK;
}

/* The state of the one-event entry points on `stream`, created on its first use
 * (its status word zeroed synchronously there).  Called with ctx->mtx held. */
int stream_state(fjmi_context *ctx, hipStream_t stream, fjmi::rt::StreamState **out) {
  auto it = ctx->streams.find(static_cast<void *>(stream));
  if (it == ctx->streams.end()) {
    fjmi::rt::StreamState st;
    int rc = st.status.ensure(sizeof(uint32_t));
    if (rc) return rc;
    hipError_t e = hipMemset(st.status.ptr, 0, sizeof(uint32_t));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      st.release_all();
      return fail(FJMI_EHIP, "status word init: %s", hipGetErrorString(e));
    }
    it = ctx->streams.emplace(static_cast<void *>(stream), st).first;
  }
  *out = &it->second;
  return FJMI_OK;
}

/* Zeroes n 32-bit words on `stream` with a fill kernel (no memset node, so a
 * caller may capture the call into a graph). */
int zero_words(void *p, int n, hipStream_t stream) {
  unsigned *w = static_cast<unsigned *>(p);
  unsigned long long *t = nullptr;
  int nt = 0;
  void *params[] = {&w, &n, &t, &nt};
  HIPCHK(hipLaunchKernel(fjmi::kZeroKernel, dim3(1), dim3(64), params, 0, stream));
  return FJMI_OK;
}

/* One event on the device: one kernel launch with one workgroup and no queue
 * head (the kernel takes the single queue entry and leaves), using the scratch
 * of `st`, the state of `stream`. */
int cluster_one_device(fjmi_context *ctx, fjmi::rt::StreamState &st, int engine, const double *d_in, int in_stride,
                       int in_off, int n, int p, double R, int32_t *h1, int32_t *h2, double *hd, double *jets,
                       int32_t *nj, unsigned char *pj, hipStream_t stream, const fjmi::JetDef *def = nullptr,
                       double *hist_jets = nullptr) {
  /* under the default table: FJMI_ENGINE_AUTO follows the latency rule (the
   * dense shape with the fewest slots per thread, or the latency tiled row), and
   * the tiled engine takes the latency tiled row; under a caller's table, its
   * row for n, else (engine forced) tile side R with the 512-tile variant or the
   * dense shape with the fewest slots per thread.  A general jet definition
   * (def) runs on the tiled engine. */
  if (def) engine = FJMI_ENGINE_TILED;
  fjmi_bucket forced;
  const fjmi_bucket *b = nullptr;
  if (engine == FJMI_ENGINE_AUTO && !ctx->caller_table) {
    if (!(latency_prefers_dense(n, R) && dense_row_for(n, &forced))) forced = latency_tiled_row(n, R);
    b = &forced;
  } else if (engine == FJMI_ENGINE_TILED && !ctx->caller_table) {
    forced = latency_tiled_row(n, R);
    b = &forced;
  }
  if (!b) b = bucket_for(ctx, n, engine);
  if (!b && engine == FJMI_ENGINE_TILED) b = &kTiledRow;
  if (!b && engine == FJMI_ENGINE_DENSE && dense_row_for(n, &forced)) b = &forced;
  if (!b) return fail(FJMI_ECAPACITY, "event with %d particles exceeds the largest bucket of engine %d", n, engine);
  if (b->engine == FJMI_ENGINE_TILED && n > fjmi::kTiledMaxN)
    return fail(FJMI_ECAPACITY, "event with %d particles: the tiled engine takes at most %d", n, fjmi::kTiledMaxN);
  int rc;
  const bool tiled = b->engine == FJMI_ENGINE_TILED;
  const KernelDesc *kd = tiled ? nullptr : find_kernel(b->threads, b->slots, b->p4_global);
  fjmi::TiledScratch ts{};
  if (tiled) {
    if ((rc = st.tscratch.ensure(fjmi::tiled_scratch_bytes(n, 1, fjmi::kTiledPoolPerEvent)))) return rc;
    ts = fjmi::tiled_scratch_carve(st.tscratch.ptr, n, fjmi::kTiledPoolPerEvent);
  } else if (kd->p4g) {
    if ((rc = st.p4s.ensure(sizeof(double) * 4 * (size_t)n))) return rc;
    if ((rc = st.coords.ensure(sizeof(double) * 4 * (size_t)n))) return rc;
  }
  LaunchArgs a = base_args();
  a.in = d_in;
  a.in_stride = in_stride;
  a.in_off = in_off;
  a.offsets = nullptr; /* single event: particles [0, n) */
  a.n_single = n;
  a.queue = nullptr;
  a.n_queue = 1;
  a.head = nullptr; /* one pass: queue entry 0 */
  a.head_base = 0;
  a.p_uni = p;
  a.R_uni = R;
  a.hist_p1 = h1;
  a.hist_p2 = h2;
  a.hist_dij = hd;
  a.jets = jets;
  a.njets = nj;
  a.pj_out = pj;
  a.p4_scratch = st.p4s.as<double>();
  a.coord_scratch = st.coords.as<double>();
  a.status = st.status.as<unsigned int>();
  a.tile_q = tiled ? b->slots : 0;
  if (def) {
    a.def_uni = *def;
    a.hist_jets = hist_jets;
  }
  if (tiled) {
    void *params[] = {&a, &ts};
    /* general definitions run one wave per event; a row of several waves
     * per event runs the cooperative kernel */
    const bool coop = !def && b->threads == fjmi::kTiledCoopThreads;
    const void *fn = def ? fjmi::kTiledGenKernels[b->p4_global]
                         : (coop ? fjmi::kTiledCoopKernels[b->p4_global] : fjmi::kTiledKernels[b->p4_global]);
    HIPCHK(hipLaunchKernel(fn, dim3(1), dim3(coop ? fjmi::kTiledCoopThreads : 64), params, 0, stream));
  } else if ((rc = launch_dense(*kd, 1, dyn_lds_bytes(*kd), a, stream))) {
    return rc;
  }
  return FJMI_OK;
}

/* The fjmi_status of a kernel status word. */
int status_error(uint32_t status) {
  if (status & fjmi::STATUS_INTERNAL)
    return fail(FJMI_EINTERNAL, "a clustering kernel found an inconsistent state (status 0x%x)", status);
  if (status & fjmi::STATUS_DEFINITION)
    return fail(FJMI_EINVAL, "an event's jet definition has a non-finite exponent or an unknown scheme (status 0x%x)",
                status);
  return fail(FJMI_ECAPACITY, "a clustering kernel reported an event beyond its capacity (status 0x%x)", status);
}

int check_engine(int engine) {
  if (engine != FJMI_ENGINE_AUTO && engine != FJMI_ENGINE_DENSE && engine != FJMI_ENGINE_TILED)
    return fail(FJMI_EUNSUPPORTED, "engine %d not available for this entry point", engine);
  return FJMI_OK;
}

/* Buckets, longest-first queues, scratch and per-bucket launch arguments of a
 * batch whose event e holds particles [h_offsets[e], h_offsets[e+1]); the
 * device buffers of `pl` grow as needed and are kept for the next plan.  The
 * queues are left in pl.flat for the caller to upload to pl.queue. */
int make_plan(fjmi_context *ctx, int n_events, const int64_t *h_offsets, const double *d_p4, const int64_t *d_offsets,
              const int32_t *d_p, const double *d_R, int p_uni, double R_uni, int32_t *d_h1, int32_t *d_h2,
              double *d_hd, double *d_jets, int32_t *d_nj, Plan &pl, const fjmi::JetDef *d_defs = nullptr,
              const fjmi::JetDef *def_uni = nullptr, double *d_hist_jets = nullptr) {
  const int nb = (int)ctx->buckets.size();
  /* general jet definitions run on the tiled engine: a dense row of the table
   * becomes the one-event tiled row (384 tiles, side R) for its sizes */
  const bool gen = d_defs || def_uni;
  std::vector<fjmi_bucket> rows(ctx->buckets.begin(), ctx->buckets.end());
  if (gen)
    for (fjmi_bucket &r : rows) {
      if (r.engine != FJMI_ENGINE_TILED) r = fjmi_bucket{r.max_n, FJMI_ENGINE_TILED, 64, 0, 1};
      r.threads = 64; /* the general kernels run one wave per event */
    }
  pl.nb = nb;
  pl.runs.clear();
  pl.flat.clear();
  std::vector<std::vector<int32_t>> lists(nb);
  int64_t max_p4g_total = 0;
  bool any_tiled = false;
  for (int e = 0; e < n_events; ++e) {
    const int64_t n = h_offsets[e + 1] - h_offsets[e];
    if (n < 0 || h_offsets[e] < 0) return fail(FJMI_EINVAL, "offsets not monotonic at event %d", e);
    int b = 0;
    while (b < nb && n > rows[b].max_n) ++b;
    if (b == nb) return fail(FJMI_ECAPACITY, "event %d has %lld particles, above the largest bucket", e, (long long)n);
    if (rows[b].engine == FJMI_ENGINE_TILED && n > fjmi::kTiledMaxN)
      return fail(FJMI_ECAPACITY, "event %d has %lld particles: the tiled engine takes at most %d", e, (long long)n,
                  fjmi::kTiledMaxN);
    lists[b].push_back(e);
    if (rows[b].engine == FJMI_ENGINE_TILED)
      any_tiled = true;
    else if (rows[b].p4_global)
      max_p4g_total = h_offsets[n_events];
  }
  /* cost-ordered queues (longest first): the last events to start are the cheapest */
  std::vector<size_t> start(nb);
  for (int b = 0; b < nb; ++b) {
    std::stable_sort(lists[b].begin(), lists[b].end(), [&](int32_t x, int32_t y) {
      return h_offsets[x + 1] - h_offsets[x] > h_offsets[y + 1] - h_offsets[y];
    });
    start[b] = pl.flat.size();
    pl.flat.insert(pl.flat.end(), lists[b].begin(), lists[b].end());
  }
  int rc;
  if ((rc = pl.queue.ensure(sizeof(int32_t) * std::max<size_t>(pl.flat.size(), 1)))) return rc;
  if ((rc = pl.heads.ensure(sizeof(uint32_t) * (nb + 1)))) return rc;
  if ((rc = pl.times.ensure(sizeof(unsigned long long) * (nb + 1)))) return rc;
  if (max_p4g_total > 0) {
    if ((rc = pl.p4s.ensure(32 * (size_t)max_p4g_total))) return rc;
    if ((rc = pl.coords.ensure(32 * (size_t)max_p4g_total))) return rc;
  }
  if (any_tiled) {
    const int64_t P = h_offsets[n_events];
    if ((rc = pl.tscratch.ensure(fjmi::tiled_scratch_bytes(P, n_events, fjmi::kTiledPoolPerEvent)))) return rc;
  }
  uint32_t *heads = pl.heads.as<uint32_t>();
  pl.d_status = heads + nb;
  for (int b = nb - 1; b >= 0; --b) { /* largest multiplicities first */
    if (lists[b].empty()) continue;
    const fjmi_bucket &bk = rows[b];
    BucketRun r;
    r.tiled = bk.engine == FJMI_ENGINE_TILED;
    r.kd = r.tiled ? nullptr : find_kernel(bk.threads, bk.slots, bk.p4_global);
    /* a row of several waves per event runs the cooperative kernel while its
     * bucket holds at most one event per CU; with more events, one wave per
     * event keeps more of them in flight (10 per CU), which a batch needs */
    const bool coop = r.tiled && !gen && bk.threads == fjmi::kTiledCoopThreads && (int)lists[b].size() <= ctx->n_cu;
    r.fn = r.tiled ? (gen ? fjmi::kTiledGenKernels[bk.p4_global]
                          : (coop ? fjmi::kTiledCoopKernels[bk.p4_global] : fjmi::kTiledKernels[bk.p4_global]))
                   : r.kd->fn;
    r.threads = r.tiled ? (coop ? fjmi::kTiledCoopThreads : 64) : r.kd->T;
    r.lds = r.tiled ? 0 : dyn_lds_bytes(*r.kd);
    r.ts = r.tiled ? fjmi::tiled_scratch_carve(pl.tscratch.ptr, h_offsets[n_events], fjmi::kTiledPoolPerEvent)
                   : fjmi::TiledScratch{};
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, r.fn, r.threads, r.lds);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    /* FJMI_MAX_WG_PER_CU caps the persistent workgroups per CU (fewer events in
     * flight: a smaller working set for the caches) */
    if (const char *cap = std::getenv("FJMI_MAX_WG_PER_CU")) {
      const int c = std::atoi(cap);
      if (c >= 1 && c < per_cu) per_cu = c;
    }
    r.n_events = (int)lists[b].size();
    r.max_n = bk.max_n;
    r.tile_q = r.tiled ? bk.slots : 0;
    r.tile_table = r.tiled ? bk.p4_global : 0;
    r.particles = 0;
    r.sum_n2 = 0.0;
    for (int32_t ev : lists[b]) {
      const double nev = (double)(h_offsets[ev + 1] - h_offsets[ev]);
      r.particles += (int64_t)nev;
      r.sum_n2 += nev * nev;
    }
    r.grid = (int)std::min<int64_t>(r.n_events, (int64_t)per_cu * ctx->n_cu);
    LaunchArgs a = base_args();
    a.in = d_p4;
    a.offsets = d_offsets;
    a.queue = pl.queue.as<int32_t>() + start[b];
    a.n_queue = r.n_events;
    a.head = heads + b;
    a.head_base = 0;
    a.p_arr = d_p;
    a.R_arr = d_R;
    a.p_uni = p_uni;
    a.R_uni = R_uni;
    a.hist_p1 = d_h1;
    a.hist_p2 = d_h2;
    a.hist_dij = d_hd;
    a.jets = d_jets;
    a.njets = d_nj;
    a.p4_scratch = pl.p4s.as<double>();
    a.coord_scratch = pl.coords.as<double>();
    a.status = pl.d_status;
    a.t_end = pl.times.as<unsigned long long>() + 1 + pl.runs.size();
    a.tile_q = r.tile_q;
    a.def_arr = d_defs;
    if (def_uni) a.def_uni = *def_uni;
    if (gen) a.hist_jets = d_hist_jets;
    r.args = a;
    pl.runs.push_back(r);
  }
  return FJMI_OK;
}

/* ---- host-memory batches: a pipeline of chunks through pinned staging ---- */

/* Runs f(i) for i in [0, n) on up to 8 host threads (the staging copies). */
template <class F>
void parallel_for(int n, const F &f) {
  const int T = std::max(1, std::min(8, (int)std::thread::hardware_concurrency()));
  if (n < 256 || T == 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int> next{0};
  auto work = [&]() {
    for (;;) {
      const int i0 = next.fetch_add(64);
      if (i0 >= n) return;
      for (int i = i0; i < std::min(n, i0 + 64); ++i) f(i);
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < T; ++t) th.emplace_back(work);
  work();
  for (std::thread &x : th) x.join();
}

int host_alloc_pinned(HostSlot &sl, size_t bytes) {
  if (bytes <= sl.pin_cap) return FJMI_OK;
  if (sl.pin) HIPCHK(hipHostFree(sl.pin));
  sl.pin = nullptr;
  sl.pin_cap = 0;
  void *q = nullptr;
  HIPCHK(hipHostMalloc(&q, bytes, hipHostMallocDefault));
  sl.pin = static_cast<char *>(q);
  sl.pin_cap = bytes;
  return FJMI_OK;
}

/* Stages the chunk's events (original indices ev[0..n)) and enqueues its
 * copies in, its kernels and its copies out on the slot's stream.  A general
 * jet definition comes per event (defs, host memory) or for all (def_uni);
 * want_hj brings the history momenta back too. */
int host_start(fjmi_context *ctx, HostSlot &sl, const int32_t *ev, int n, const int64_t *offsets, const double *p4,
               const int32_t *p, const double *R, int p_uni, double R_uni, const fjmi::JetDef *defs = nullptr,
               const fjmi::JetDef *def_uni = nullptr, bool want_hj = false) {
  int rc;
  if (!sl.s) {
    HIPCHK(hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&sl.ev_out, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&sl.ev_fork, hipEventDisableTiming));
  }
  sl.events.assign(ev, ev + n);
  sl.n_events = n;
  sl.off.resize(n + 1);
  sl.off[0] = 0;
  for (int i = 0; i < n; ++i) sl.off[i + 1] = sl.off[i] + (offsets[ev[i] + 1] - offsets[ev[i]]);
  const int64_t P = sl.off[n];
  sl.particles = P;
  const size_t Pn = (size_t)std::max<int64_t>(P, 1), En = (size_t)n;
  /* pinned layout: in (32 P) | h1 (4 P) | h2 (4 P) | hd (8 P) | jc (32 P) | offsets (8 (E+1)) | p (4 E) |
   * R (8 E) | queue (4 E) | nj (4 E) | joff (4 (E+1)) | status (4, padded to 8) | defs (16 E) | hj (32 P) */
  const size_t o_in = 0, o_h1 = o_in + 32 * Pn, o_h2 = o_h1 + 4 * Pn, o_hd = o_h2 + 4 * Pn, o_jc = o_hd + 8 * Pn,
               o_off = o_jc + 32 * Pn, o_p = o_off + 8 * (En + 1), o_R = o_p + 4 * En, o_q = o_R + 8 * En,
               o_nj = o_q + 4 * En, o_jo = o_nj + 4 * En, o_st = o_jo + 4 * (En + 1),
               o_df = (o_st + 4 + 7) & ~(size_t)7, o_hj = o_df + (defs ? 16 * En : 0),
               o_end = o_hj + (want_hj ? 32 * Pn : 0);
  if ((rc = host_alloc_pinned(sl, o_end))) return rc;
  double *in = reinterpret_cast<double *>(sl.pin + o_in);
  int64_t *off = reinterpret_cast<int64_t *>(sl.pin + o_off);
  int32_t *pp = reinterpret_cast<int32_t *>(sl.pin + o_p);
  double *RR = reinterpret_cast<double *>(sl.pin + o_R);
  std::memcpy(off, sl.off.data(), 8 * (En + 1));
  fjmi::JetDef *dd = reinterpret_cast<fjmi::JetDef *>(sl.pin + o_df);
  parallel_for(n, [&](int i) {
    const int e = ev[i];
    std::memcpy(in + 4 * off[i], p4 + 4 * offsets[e], 32 * (size_t)(offsets[e + 1] - offsets[e]));
    if (p) pp[i] = p[e];
    if (R) RR[i] = R[e];
    if (defs) dd[i] = defs[e];
  });
  sl.h_hj = want_hj ? reinterpret_cast<double *>(sl.pin + o_hj) : nullptr;
  sl.h_h1 = reinterpret_cast<int32_t *>(sl.pin + o_h1);
  sl.h_h2 = reinterpret_cast<int32_t *>(sl.pin + o_h2);
  sl.h_hd = reinterpret_cast<double *>(sl.pin + o_hd);
  sl.h_jc = reinterpret_cast<double *>(sl.pin + o_jc);
  sl.h_nj = reinterpret_cast<int32_t *>(sl.pin + o_nj);
  sl.h_joff = reinterpret_cast<int32_t *>(sl.pin + o_jo);
  sl.h_status = reinterpret_cast<uint32_t *>(sl.pin + o_st);
  /* device buffers */
  if ((rc = sl.d_in.ensure(32 * Pn)) || (rc = sl.d_off.ensure(8 * (En + 1))) || (rc = sl.d_p.ensure(4 * En)) ||
      (rc = sl.d_R.ensure(8 * En)) || (rc = sl.d_h1.ensure(4 * Pn)) || (rc = sl.d_h2.ensure(4 * Pn)) ||
      (rc = sl.d_hd.ensure(8 * Pn)) || (rc = sl.d_jets.ensure(32 * Pn)) || (rc = sl.d_nj.ensure(4 * En)) ||
      (rc = sl.d_joff.ensure(4 * (En + 1))) || (rc = sl.d_jc.ensure(32 * Pn)) ||
      (defs && (rc = sl.d_defs.ensure(16 * En))) || (want_hj && (rc = sl.d_hj.ensure(32 * Pn))))
    return rc;
  if ((rc = make_plan(ctx, n, sl.off.data(), sl.d_in.as<double>(), sl.d_off.as<int64_t>(),
                      p ? sl.d_p.as<int32_t>() : nullptr, R ? sl.d_R.as<double>() : nullptr, p_uni, R_uni,
                      sl.d_h1.as<int32_t>(), sl.d_h2.as<int32_t>(), sl.d_hd.as<double>(), sl.d_jets.as<double>(),
                      sl.d_nj.as<int32_t>(), sl.plan, defs ? sl.d_defs.as<fjmi::JetDef>() : nullptr, def_uni,
                      want_hj ? sl.d_hj.as<double>() : nullptr)))
    return rc;
  int32_t *q = reinterpret_cast<int32_t *>(sl.pin + o_q);
  std::memcpy(q, sl.plan.flat.data(), 4 * sl.plan.flat.size());
  hipStream_t s = sl.s;
  HIPCHK(hipMemcpyAsync(sl.d_in.ptr, in, 32 * (size_t)P, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(sl.d_off.ptr, off, 8 * (En + 1), hipMemcpyHostToDevice, s));
  if (p) HIPCHK(hipMemcpyAsync(sl.d_p.ptr, pp, 4 * En, hipMemcpyHostToDevice, s));
  if (R) HIPCHK(hipMemcpyAsync(sl.d_R.ptr, RR, 8 * En, hipMemcpyHostToDevice, s));
  if (defs) HIPCHK(hipMemcpyAsync(sl.d_defs.ptr, dd, 16 * En, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(sl.plan.queue.ptr, q, 4 * sl.plan.flat.size(), hipMemcpyHostToDevice, s));
  /* the kernels: queue heads and status cleared, then each bucket's persistent
   * kernel (buckets after the first on streams of their own, joined again) */
  {
    unsigned *hq = sl.plan.heads.as<unsigned>();
    int cnt = sl.plan.nb + 1;
    unsigned long long *tp = sl.plan.times.as<unsigned long long>();
    int tcnt = (int)sl.plan.runs.size() + 1;
    void *zp[] = {&hq, &cnt, &tp, &tcnt};
    HIPCHK(hipLaunchKernel(fjmi::kZeroKernel, dim3(1), dim3(64), zp, 0, s));
  }
  const size_t nr = sl.plan.runs.size();
  if (nr > 1) {
    HIPCHK(hipEventRecord(sl.ev_fork, s));
    while (sl.fork.size() < nr - 1) {
      hipStream_t f = nullptr;
      hipEvent_t j = nullptr;
      HIPCHK(hipStreamCreateWithFlags(&f, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&j, hipEventDisableTiming));
      sl.fork.push_back(f);
      sl.ev_join.push_back(j);
    }
  }
  for (size_t i = 0; i < nr; ++i) {
    BucketRun &r = sl.plan.runs[i];
    void *params[] = {&r.args, &r.ts};
    hipStream_t rs = i == 0 ? s : sl.fork[i - 1];
    if (i > 0) HIPCHK(hipStreamWaitEvent(rs, sl.ev_fork, 0));
    HIPCHK(hipLaunchKernel(r.fn, dim3(r.grid), dim3(r.threads), params, r.lds, rs));
    if (i > 0) HIPCHK(hipEventRecord(sl.ev_join[i - 1], rs));
  }
  for (size_t i = 1; i < nr; ++i) HIPCHK(hipStreamWaitEvent(s, sl.ev_join[i - 1], 0));
  /* jets packed back to back, then the copies out of everything but the jets */
  {
    const int32_t *dn = sl.d_nj.as<int32_t>();
    int32_t *jo = sl.d_joff.as<int32_t>();
    int nn = n;
    void *sp[] = {&dn, &nn, &jo};
    HIPCHK(hipLaunchKernel(fjmi::kJetsScanKernel, dim3(1), dim3(256), sp, 0, s));
    if (n > 0) {
      const double *jets = sl.d_jets.as<double>();
      const int64_t *doff = sl.d_off.as<int64_t>();
      double *out = sl.d_jc.as<double>();
      void *pk[] = {&jets, &doff, &dn, &jo, &out};
      HIPCHK(hipLaunchKernel(fjmi::kJetsPackKernel, dim3(n), dim3(64), pk, 0, s));
    }
  }
  HIPCHK(hipMemcpyAsync(sl.h_h1, sl.d_h1.ptr, 4 * (size_t)P, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(sl.h_h2, sl.d_h2.ptr, 4 * (size_t)P, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(sl.h_hd, sl.d_hd.ptr, 8 * (size_t)P, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(sl.h_nj, sl.d_nj.ptr, 4 * En, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(sl.h_joff, sl.d_joff.ptr, 4 * (En + 1), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(sl.h_status, sl.plan.d_status, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (want_hj) HIPCHK(hipMemcpyAsync(sl.h_hj, sl.d_hj.ptr, 32 * (size_t)P, hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(sl.ev_out, s));
  sl.busy = true;
  return FJMI_OK;
}

/* Waits for a slot's chunk, copies its packed jets out, and writes its results
 * into the caller's arrays at the events' own positions. */
int host_finish(HostSlot &sl, const int64_t *offsets, int32_t *h1, int32_t *h2, double *hd, double *jets,
                int32_t *nj, double *hist_jets = nullptr) {
  sl.busy = false;
  HIPCHK(hipEventSynchronize(sl.ev_out));
  const int n = sl.n_events;
  const int64_t njt = sl.h_joff[n];
  if (njt > 0)
    HIPCHK(hipMemcpyAsync(sl.h_jc, sl.d_jc.ptr, 32 * (size_t)njt, hipMemcpyDeviceToHost, sl.s));
  HIPCHK(hipStreamSynchronize(sl.s));
  if (*sl.h_status) return status_error(*sl.h_status);
  parallel_for(n, [&](int i) {
    const int e = sl.events[i];
    const int64_t lo = sl.off[i], m = sl.off[i + 1] - lo, dst = offsets[e];
    std::memcpy(h1 + dst, sl.h_h1 + lo, 4 * (size_t)m);
    std::memcpy(h2 + dst, sl.h_h2 + lo, 4 * (size_t)m);
    std::memcpy(hd + dst, sl.h_hd + lo, 8 * (size_t)m);
    const int k = sl.h_nj[i];
    nj[e] = k;
    if (k > 0) std::memcpy(jets + 4 * dst, sl.h_jc + 4 * (size_t)sl.h_joff[i], 32 * (size_t)k);
    if (hist_jets && sl.h_hj) std::memcpy(hist_jets + 4 * dst, sl.h_hj + 4 * lo, 32 * (size_t)m);
  });
  return FJMI_OK;
}

int check_args(int n, int p, double R) {
  if (n < 0) return fail(FJMI_EINVAL, "negative particle count");
  if (p < -1 || p > 1) return fail(FJMI_EINVAL, "exponent p must be -1, 0 or 1");
  if (!(R > 0.0) || !(R < 1e150)) return fail(FJMI_EINVAL, "radius must be positive and finite");
  return FJMI_OK;
}

}  // namespace

extern "C" {

const char *fjmi_last_error(void) { return g_err.c_str(); }

int fjmi_version(void) { return 100; }

int fjmi_kernel_configs(fjmi_bucket *out, int cap) {
  for (int i = 0; i < fjmi::kNumDenseKernels && i < cap; ++i) {
    const KernelDesc &d = fjmi::kDenseKernels[i];
    out[i].max_n = d.T * d.K;
    out[i].engine = FJMI_ENGINE_DENSE;
    out[i].threads = d.T;
    out[i].slots = d.K;
    out[i].p4_global = d.p4g ? 1 : 0;
  }
  for (int k = 0; k < fjmi::kNumTiledKernels; ++k) {
    if (fjmi::kNumDenseKernels + k >= cap) break;
    fjmi_bucket r = kTiledRow;
    r.p4_global = k;
    out[fjmi::kNumDenseKernels + k] = r;
  }
  for (int k = 0; k < 2; ++k) { /* the cooperative kernels: 512 and 384 tiles, 8 waves per event */
    if (fjmi::kNumDenseKernels + fjmi::kNumTiledKernels + k >= cap) break;
    fjmi_bucket r = kTiledRow;
    r.threads = fjmi::kTiledCoopThreads;
    r.p4_global = k;
    out[fjmi::kNumDenseKernels + fjmi::kNumTiledKernels + k] = r;
  }
  return fjmi::kNumDenseKernels + fjmi::kNumTiledKernels + 2;
}

int fjmi_create(int device, fjmi_context **out) {
  if (!out) return fail(FJMI_EINVAL, "null output pointer");
  *out = nullptr;
  int count = 0;
  HIPCHK(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(FJMI_EINVAL, "device %d not present (%d devices)", device, count);
  auto *ctx = new fjmi_context();
  ctx->device = device;
  int rc = set_device(ctx);
  if (rc) {
    delete ctx;
    return rc;
  }
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    delete ctx;
    return fail(FJMI_EHIP, "hipGetDeviceProperties: %s", hipGetErrorString(e));
  }
  ctx->n_cu = prop.multiProcessorCount;
  ctx->max_lds = (int)prop.sharedMemPerBlock;
  e = hipGetSymbolAddress(reinterpret_cast<void **>(&ctx->static_scratch), HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ctx->static_scratch) {
    delete ctx;
    return fail(FJMI_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  }
  for (int i = 0; i < fjmi::kNumDenseKernels; ++i) {
    const KernelDesc &d = fjmi::kDenseKernels[i];
    size_t lds = dyn_lds_bytes(d);
    e = hipFuncSetAttribute(d.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
      delete ctx;
      return fail(FJMI_EHIP, "hipFuncSetAttribute(%d,%d): %s", d.T, d.K, hipGetErrorString(e));
    }
  }
  const int nb = (int)(sizeof kDefaultBuckets / sizeof kDefaultBuckets[0]);
  if ((rc = check_table(kDefaultBuckets, nb))) {
    delete ctx;
    return rc;
  }
  ctx->buckets.assign(kDefaultBuckets, kDefaultBuckets + nb);
  *out = ctx;
  return FJMI_OK;
}

int fjmi_destroy(fjmi_context *ctx) {
  if (!ctx) return FJMI_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  for (auto &kv : ctx->streams) kv.second.release_all();
  if (ctx->host) {
    for (HostSlot &sl : ctx->host->slot) sl.release();
    delete ctx->host;
  }
  delete ctx;
  return FJMI_OK;
}

int fjmi_set_buckets(fjmi_context *ctx, const fjmi_bucket *rows, int n_rows) {
  if (!ctx || !rows) return fail(FJMI_EINVAL, "null argument");
  int rc = check_table(rows, n_rows);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  ctx->buckets.assign(rows, rows + n_rows);
  ctx->caller_table = true;
  return FJMI_OK;
}

int fjmi_get_buckets(fjmi_context *ctx, fjmi_bucket *rows, int cap, int *n_rows) {
  if (!ctx || !n_rows) return fail(FJMI_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(ctx->mtx);
  *n_rows = (int)ctx->buckets.size();
  for (int i = 0; i < *n_rows && i < cap; ++i) rows[i] = ctx->buckets[i];
  return FJMI_OK;
}

int fjmi_cluster_event_device(fjmi_context *ctx, const double *d_p4, int n, int p, double R, int engine,
                              int32_t *d_h1, int32_t *d_h2, double *d_hd, double *d_jets, int32_t *d_nj,
                              void *stream) {
  if (!ctx) return fail(FJMI_EINVAL, "null context");
  int rc = check_args(n, p, R);
  if (rc) return rc;
  if ((rc = check_engine(engine))) return rc;
  if (n > 0 && (!d_p4 || !d_h1 || !d_h2 || !d_hd || !d_jets || !d_nj)) return fail(FJMI_EINVAL, "null buffer");
  if (n == 0 && !d_nj) return fail(FJMI_EINVAL, "null njets");
  std::lock_guard<std::mutex> lk(ctx->mtx);
  if ((rc = set_device(ctx))) return rc;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (n == 0) return zero_words(d_nj, 1, s);
  fjmi::rt::StreamState *st = nullptr;
  if ((rc = stream_state(ctx, s, &st))) return rc;
  return cluster_one_device(ctx, *st, engine, d_p4, 4, 0, n, p, R, d_h1, d_h2, d_hd, d_jets, d_nj, nullptr, s);
}

int fjmi_cluster_pseudojets(fjmi_context *ctx, void *d_particles, int n, int p, double R, int engine, void *stream) {
  if (!ctx) return fail(FJMI_EINVAL, "null context");
  int rc = check_args(n, p, R);
  if (rc) return rc;
  if (n == 0) return FJMI_OK;
  if (!d_particles) return fail(FJMI_EINVAL, "null particle buffer");
  std::lock_guard<std::mutex> lk(ctx->mtx);
  if ((rc = set_device(ctx))) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (engine == FJMI_ENGINE_EMPTY) {
    /* the per-step launch schedule of the reference's tri_matrix_empty.cu:257-292:
     * init + set_jets + 2 launches per step + output, all empty */
    const int launches = 2 * n + 3;
    for (int i = 0; i < launches; ++i) {
      int arg = i;
      void *params[] = {&arg};
      HIPCHK(hipLaunchKernel(fjmi::kEmptyKernel, dim3(1), dim3(64), params, 0, s));
    }
    return FJMI_OK;
  }
  if ((rc = check_engine(engine))) return rc;
  fjmi::rt::StreamState *st = nullptr;
  if ((rc = stream_state(ctx, s, &st))) return rc;
  const size_t nn = (size_t)n;
  if ((rc = st->hist1.ensure(4 * nn)) || (rc = st->hist2.ensure(4 * nn)) || (rc = st->hdij.ensure(8 * nn)) ||
      (rc = st->jets.ensure(32 * nn)) || (rc = st->njets.ensure(4)))
    return rc;
  return cluster_one_device(ctx, *st, engine, static_cast<const double *>(d_particles), 5, 1, n, p, R,
                            st->hist1.as<int32_t>(), st->hist2.as<int32_t>(), st->hdij.as<double>(),
                            st->jets.as<double>(), st->njets.as<int32_t>(), static_cast<unsigned char *>(d_particles),
                            s);
}

int fjmi_cluster_event_device_noalloc(fjmi_context *ctx, const double *d_p4, int n, int p, double R,
                                      int32_t *d_h1, int32_t *d_h2, double *d_hd, double *d_jets, int32_t *d_nj,
                                      void *stream) {
  if (!ctx) return fail(FJMI_EINVAL, "null context");
  int rc = check_args(n, p, R);
  if (rc) return rc;
  if (!d_nj || (n > 0 && (!d_p4 || !d_h1 || !d_h2 || !d_hd || !d_jets))) return fail(FJMI_EINVAL, "null buffer");
  if (n > kNoAllocMaxN)
    return fail(FJMI_ECAPACITY, "event with %d particles: the no-allocation entry point takes at most %d", n,
                kNoAllocMaxN);
  std::lock_guard<std::mutex> lk(ctx->mtx);
  if ((rc = set_device(ctx))) return rc;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (n == 0) return zero_words(d_nj, 1, s);
  /* under the default table the latency row for one event alone; else the
   * table's tiled row for n, else tile side R with the 512-tile variant */
  const fjmi_bucket latency_row = latency_tiled_row(n, R);
  const fjmi_bucket *b = ctx->caller_table ? bucket_for(ctx, n, FJMI_ENGINE_TILED) : &latency_row;
  if (!b) b = &kTiledRow;
  /* every byte of the tiled scratch that the kernel reads it has written for this
   * event first, so nothing is cleared; the status word is written on failure only
   * (the caller sees njets = -1) and is never read */
  fjmi::TiledScratch ts = fjmi::tiled_scratch_carve(ctx->static_scratch, n, fjmi::kTiledPoolPerEvent);
  LaunchArgs a = base_args();
  a.in = d_p4;
  a.offsets = nullptr;
  a.n_single = n;
  a.queue = nullptr;
  a.n_queue = 1;
  a.head = nullptr;
  a.p_uni = p;
  a.R_uni = R;
  a.hist_p1 = d_h1;
  a.hist_p2 = d_h2;
  a.hist_dij = d_hd;
  a.jets = d_jets;
  a.njets = d_nj;
  a.status = reinterpret_cast<unsigned int *>(ctx->static_scratch + kNoAllocTiledBytes);
  a.tile_q = b->slots;
  void *params[] = {&a, &ts};
  HIPCHK(hipLaunchKernel(fjmi::kTiledKernels[b->p4_global], dim3(1), dim3(64), params, 0, s));
  return FJMI_OK;
}

int fjmi_inclusive_jets(fjmi_context *ctx, int n_events, const int64_t *d_offsets, const double *d_jets,
                        const int32_t *d_njets, double ptmin, double *d_out_jets, int32_t *d_out_njets,
                        void *stream) {
  if (!ctx) return fail(FJMI_EINVAL, "null context");
  if (n_events < 0 || (n_events > 1 && !d_offsets)) return fail(FJMI_EINVAL, "offsets needed for several events");
  if (!(ptmin >= 0.0)) return fail(FJMI_EINVAL, "ptmin must be >= 0");
  if (n_events == 0) return FJMI_OK;
  if (!d_jets || !d_njets || !d_out_jets || !d_out_njets) return fail(FJMI_EINVAL, "null buffer");
  if (d_out_jets == d_jets) return fail(FJMI_EINVAL, "the output must not overlap the input");
  std::lock_guard<std::mutex> lk(ctx->mtx);
  int rc;
  if ((rc = set_device(ctx))) return rc;
  void *params[] = {&d_jets, &d_njets, &d_offsets, &ptmin, &d_out_jets, &d_out_njets};
  HIPCHK(hipLaunchKernel(fjmi::kInclusiveKernel, dim3(n_events), dim3(256), params, 0, (hipStream_t)stream));
  return FJMI_OK;
}

int fjmi_inclusive_pseudojets(fjmi_context *ctx, const void *d_particles, int n, double ptmin, void *d_out,
                              int32_t *d_count, void *stream) {
  if (!ctx) return fail(FJMI_EINVAL, "null context");
  if (n < 0 || !(ptmin >= 0.0)) return fail(FJMI_EINVAL, "bad argument");
  if (!d_count || (n > 0 && (!d_particles || !d_out))) return fail(FJMI_EINVAL, "null buffer");
  if (d_out == d_particles && n > 0) return fail(FJMI_EINVAL, "the output must not overlap the input");
  std::lock_guard<std::mutex> lk(ctx->mtx);
  int rc;
  if ((rc = set_device(ctx))) return rc;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (n == 0) return zero_words(d_count, 1, s);
  const unsigned char *in = static_cast<const unsigned char *>(d_particles);
  unsigned char *out = static_cast<unsigned char *>(d_out);
  void *params[] = {&in, &n, &ptmin, &out, &d_count};
  HIPCHK(hipLaunchKernel(fjmi::kInclusivePjKernel, dim3(1), dim3(256), params, 0, s));
  return FJMI_OK;
}

int fjmi_history_view(fjmi_context *ctx, int n_events, const int64_t *d_offsets, int n_single,
                      const int32_t *d_hist_p1, const int32_t *d_hist_p2, const double *d_hist_dij, int32_t *d_child,
                      double *d_max_dij, void *stream) {
  if (!ctx) return fail(FJMI_EINVAL, "null context");
  if (n_events < 0 || (!d_offsets && (n_events > 1 || n_single < 0))) return fail(FJMI_EINVAL, "bad event layout");
  if (n_events == 0) return FJMI_OK;
  if (!d_hist_p1 || !d_hist_p2 || !d_hist_dij || !d_child || !d_max_dij) return fail(FJMI_EINVAL, "null buffer");
  std::lock_guard<std::mutex> lk(ctx->mtx);
  int rc;
  if ((rc = set_device(ctx))) return rc;
  void *params[] = {&d_offsets, &n_single, &d_hist_p1, &d_hist_p2, &d_hist_dij, &d_child, &d_max_dij};
  HIPCHK(hipLaunchKernel(fjmi::kHistoryKernel, dim3(n_events), dim3(256), params, 0, (hipStream_t)stream));
  return FJMI_OK;
}

int fjmi_exclusive_jets(fjmi_context *ctx, int n_events, const int64_t *d_offsets, int n_single, const double *d_p4,
                        const int32_t *d_hist_p1, const int32_t *d_hist_p2, const double *d_hist_dij, int mode,
                        double param, double *d_work, double *d_out_jets, int32_t *d_out_njets, void *stream) {
  if (!ctx) return fail(FJMI_EINVAL, "null context");
  if (n_events < 0 || (!d_offsets && (n_events > 1 || n_single < 0))) return fail(FJMI_EINVAL, "bad event layout");
  if (mode == FJMI_EXCLUSIVE_NJETS) {
    if (!(param >= 0.0) || param != (double)(int)param) return fail(FJMI_EINVAL, "njets must be an integer >= 0");
  } else if (mode == FJMI_EXCLUSIVE_DCUT) {
    if (!(param >= 0.0)) return fail(FJMI_EINVAL, "dcut must be >= 0");
  } else {
    return fail(FJMI_EINVAL, "mode must be FJMI_EXCLUSIVE_NJETS or FJMI_EXCLUSIVE_DCUT");
  }
  if (n_events == 0) return FJMI_OK;
  if (!d_p4 || !d_hist_p1 || !d_hist_p2 || !d_hist_dij || !d_work || !d_out_jets || !d_out_njets)
    return fail(FJMI_EINVAL, "null buffer");
  std::lock_guard<std::mutex> lk(ctx->mtx);
  int rc;
  if ((rc = set_device(ctx))) return rc;
  const double *no_hist = nullptr;
  int scheme = FJMI_SCHEME_E;
  void *params[] = {&d_offsets, &n_single, &d_p4, &d_hist_p1, &d_hist_p2, &d_hist_dij, &mode, &param, &d_work,
                    &d_out_jets, &d_out_njets, &no_hist, &scheme};
  HIPCHK(hipLaunchKernel(fjmi::kExclusiveKernel, dim3(n_events), dim3(256), params, 0, (hipStream_t)stream));
  return FJMI_OK;
}

int fjmi_exclusive_jets_def(fjmi_context *ctx, int n_events, const int64_t *d_offsets, int n_single,
                            const double *d_p4, const int32_t *d_hist_p1, const int32_t *d_hist_p2,
                            const double *d_hist_dij, const double *d_hist_jets, int scheme, int mode, double param,
                            double *d_out_jets, int32_t *d_out_njets, void *stream) {
  if (!ctx) return fail(FJMI_EINVAL, "null context");
  if (n_events < 0 || (!d_offsets && (n_events > 1 || n_single < 0))) return fail(FJMI_EINVAL, "bad event layout");
  if (scheme < 0 || scheme >= FJMI_NUM_SCHEMES) return fail(FJMI_EINVAL, "unknown recombination scheme %d", scheme);
  if (mode == FJMI_EXCLUSIVE_NJETS) {
    if (!(param >= 0.0) || param != (double)(int)param) return fail(FJMI_EINVAL, "njets must be an integer >= 0");
  } else if (mode == FJMI_EXCLUSIVE_DCUT) {
    if (!(param >= 0.0)) return fail(FJMI_EINVAL, "dcut must be >= 0");
  } else {
    return fail(FJMI_EINVAL, "mode must be FJMI_EXCLUSIVE_NJETS or FJMI_EXCLUSIVE_DCUT");
  }
  if (n_events == 0) return FJMI_OK;
  if (!d_p4 || !d_hist_p1 || !d_hist_p2 || !d_hist_dij || !d_hist_jets || !d_out_jets || !d_out_njets)
    return fail(FJMI_EINVAL, "null buffer");
  std::lock_guard<std::mutex> lk(ctx->mtx);
  int rc;
  if ((rc = set_device(ctx))) return rc;
  double *no_work = nullptr;
  void *params[] = {&d_offsets, &n_single, &d_p4, &d_hist_p1, &d_hist_p2, &d_hist_dij, &mode, &param, &no_work,
                    &d_out_jets, &d_out_njets, &d_hist_jets, &scheme};
  HIPCHK(hipLaunchKernel(fjmi::kExclusiveKernel, dim3(n_events), dim3(256), params, 0, (hipStream_t)stream));
  return FJMI_OK;
}

int fjmi_stream_check(fjmi_context *ctx, void *stream) {
  if (!ctx) return fail(FJMI_EINVAL, "null context");
  std::lock_guard<std::mutex> lk(ctx->mtx);
  int rc;
  if ((rc = set_device(ctx))) return rc;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  HIPCHK(hipStreamSynchronize(s));
  auto it = ctx->streams.find(static_cast<void *>(s));
  if (it == ctx->streams.end()) return FJMI_OK; /* nothing launched on it */
  uint32_t status = 0;
  HIPCHK(hipMemcpy(&status, it->second.status.ptr, sizeof status, hipMemcpyDeviceToHost));
  if (!status) return FJMI_OK;
  const uint32_t zero = 0;
  HIPCHK(hipMemcpy(it->second.status.ptr, &zero, sizeof zero, hipMemcpyHostToDevice));
  return status_error(status);
}

}  // extern "C"

namespace {

/* One event from host memory, synchronously: the standard algorithms (def
 * null) or a general jet definition. */
int cluster_event_host(fjmi_context *ctx, const double *p4, int n, int p, double R, int engine,
                       const fjmi::JetDef *def, int32_t *h1, int32_t *h2, double *hd, double *jets, int32_t *nj,
                       double *hist_jets = nullptr) {
  if (!ctx) return fail(FJMI_EINVAL, "null context");
  int rc = check_args(n, p, R);
  if (rc) return rc;
  if (!nj) return fail(FJMI_EINVAL, "null njets");
  if (n == 0) {
    *nj = 0;
    return FJMI_OK;
  }
  if (!p4 || !h1 || !h2 || !hd || !jets) return fail(FJMI_EINVAL, "null buffer");
  if ((rc = check_engine(engine))) return rc;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  if ((rc = set_device(ctx))) return rc;
  fjmi::rt::StreamState *st = nullptr;
  if ((rc = stream_state(ctx, nullptr, &st))) return rc;
  const size_t nn = (size_t)n;
  if ((rc = st->p4in.ensure(32 * nn)) || (rc = st->hist1.ensure(4 * nn)) || (rc = st->hist2.ensure(4 * nn)) ||
      (rc = st->hdij.ensure(8 * nn)) || (rc = st->jets.ensure(32 * nn)) || (rc = st->njets.ensure(4)) ||
      (hist_jets && (rc = st->hjets.ensure(32 * nn))))
    return rc;
  HIPCHK(hipMemcpy(st->p4in.ptr, p4, 32 * nn, hipMemcpyHostToDevice));
  if ((rc = cluster_one_device(ctx, *st, engine, st->p4in.as<double>(), 4, 0, n, p, R, st->hist1.as<int32_t>(),
                               st->hist2.as<int32_t>(), st->hdij.as<double>(), st->jets.as<double>(),
                               st->njets.as<int32_t>(), nullptr, nullptr, def,
                               hist_jets ? st->hjets.as<double>() : nullptr)))
    return rc;
  HIPCHK(hipStreamSynchronize(nullptr));
  uint32_t status = 0;
  HIPCHK(hipMemcpy(&status, st->status.ptr, sizeof status, hipMemcpyDeviceToHost));
  if (status) {
    const uint32_t zero = 0;
    HIPCHK(hipMemcpy(st->status.ptr, &zero, sizeof zero, hipMemcpyHostToDevice));
    return status_error(status);
  }
  HIPCHK(hipMemcpy(h1, st->hist1.ptr, 4 * nn, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(h2, st->hist2.ptr, 4 * nn, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(hd, st->hdij.ptr, 8 * nn, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(nj, st->njets.ptr, 4, hipMemcpyDeviceToHost));
  if (*nj < 0 || *nj > n) return fail(FJMI_EHIP, "corrupt jet count %d", *nj);
  HIPCHK(hipMemcpy(jets, st->jets.ptr, 32 * (size_t)(*nj), hipMemcpyDeviceToHost));
  if (hist_jets) HIPCHK(hipMemcpy(hist_jets, st->hjets.ptr, 32 * nn, hipMemcpyDeviceToHost));
  return FJMI_OK;
}

/* A general jet definition the kernels can run, as the kernels' JetDef. */
int check_def(const fjmi_jetdef *def, fjmi::JetDef *out) {
  if (!def) return fail(FJMI_EINVAL, "null jet definition");
  if (!(def->p == def->p && def->p > -1e300 && def->p < 1e300))
    return fail(FJMI_EINVAL, "exponent p must be finite (|p| < 1e300)");
  if (def->scheme < 0 || def->scheme >= FJMI_NUM_SCHEMES)
    return fail(FJMI_EINVAL, "unknown recombination scheme %d", def->scheme);
  *out = fjmi::JetDef{def->p, def->scheme, 0};
  return FJMI_OK;
}

}  // namespace

extern "C" {

int fjmi_cluster_event(fjmi_context *ctx, const double *p4, int n, int p, double R, int engine, int32_t *h1,
                       int32_t *h2, double *hd, double *jets, int32_t *nj) {
  return cluster_event_host(ctx, p4, n, p, R, engine, nullptr, h1, h2, hd, jets, nj);
}

int fjmi_cluster_event_def(fjmi_context *ctx, const double *p4, int n, const fjmi_jetdef *def, double R,
                           int32_t *h1, int32_t *h2, double *hd, double *jets, int32_t *nj, double *hist_jets) {
  fjmi::JetDef d;
  const int rc = check_def(def, &d);
  if (rc) return rc;
  return cluster_event_host(ctx, p4, n, 0, R, FJMI_ENGINE_TILED, &d, h1, h2, hd, jets, nj, hist_jets);
}

int fjmi_cluster_event_device_def(fjmi_context *ctx, const double *d_p4, int n, const fjmi_jetdef *def, double R,
                                  int32_t *d_h1, int32_t *d_h2, double *d_hd, double *d_jets, int32_t *d_nj,
                                  double *d_hist_jets, void *stream) {
  if (!ctx) return fail(FJMI_EINVAL, "null context");
  fjmi::JetDef d;
  int rc = check_def(def, &d);
  if (rc) return rc;
  if ((rc = check_args(n, 0, R))) return rc;
  if (n > 0 && (!d_p4 || !d_h1 || !d_h2 || !d_hd || !d_jets || !d_nj)) return fail(FJMI_EINVAL, "null buffer");
  if (n == 0 && !d_nj) return fail(FJMI_EINVAL, "null njets");
  std::lock_guard<std::mutex> lk(ctx->mtx);
  if ((rc = set_device(ctx))) return rc;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (n == 0) return zero_words(d_nj, 1, s);
  fjmi::rt::StreamState *st = nullptr;
  if ((rc = stream_state(ctx, s, &st))) return rc;
  return cluster_one_device(ctx, *st, FJMI_ENGINE_TILED, d_p4, 4, 0, n, 0, R, d_h1, d_h2, d_hd, d_jets, d_nj, nullptr,
                            s, &d, d_hist_jets);
}

/* ---- batched path ---- */

}  // extern "C"

namespace {

/* A batch plan and its captured graph: the standard algorithms (d_defs and
 * def_uni null) or general jet definitions. */
int batch_create(fjmi_context *ctx, int n_events, const int64_t *h_offsets, const double *d_p4,
                 const int64_t *d_offsets, const int32_t *d_p, const double *d_R, int p_uni, double R_uni,
                 const fjmi::JetDef *d_defs, const fjmi::JetDef *def_uni, int32_t *d_h1, int32_t *d_h2, double *d_hd,
                 double *d_jets, int32_t *d_nj, double *d_hist_jets, fjmi_batch **out) {
  if (!ctx || !out || !h_offsets || n_events < 0) return fail(FJMI_EINVAL, "bad argument");
  *out = nullptr;
  if (!d_p && (p_uni < -1 || p_uni > 1)) return fail(FJMI_EINVAL, "exponent p must be -1, 0 or 1");
  if (!d_R && !(R_uni > 0.0 && R_uni < 1e150)) return fail(FJMI_EINVAL, "radius must be positive and finite");
  std::lock_guard<std::mutex> lk(ctx->mtx);
  int rc;
  if ((rc = set_device(ctx))) return rc;
  auto *bt = new fjmi_batch();
  bt->ctx = ctx;
  bt->n_events = n_events;
  auto cleanup = [&](int code) {
    fjmi_batch_destroy(bt);
    return code;
  };
  Plan &pl = bt->plan;
  if ((rc = make_plan(ctx, n_events, h_offsets, d_p4, d_offsets, d_p, d_R, p_uni, R_uni, d_h1, d_h2, d_hd, d_jets,
                      d_nj, pl, d_defs, def_uni, d_hist_jets)))
    return cleanup(rc);
  if (!pl.flat.empty()) {
    hipError_t e = hipMemcpy(pl.queue.ptr, pl.flat.data(), sizeof(int32_t) * pl.flat.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return cleanup(fail(FJMI_EHIP, "queue upload: %s", hipGetErrorString(e)));
  }
  /* capture: zero the heads and status, then every bucket's persistent kernel on
   * its own forked stream so the buckets run concurrently */
  hipStream_t s0 = nullptr;
  std::vector<hipStream_t> ss(pl.runs.size(), nullptr);
  bt->tev.assign(pl.runs.size() + 1, nullptr);
  hipError_t e = hipStreamCreateWithFlags(&s0, hipStreamNonBlocking);
  for (size_t i = 0; e == hipSuccess && i < ss.size(); ++i) e = hipStreamCreateWithFlags(&ss[i], hipStreamNonBlocking);
  for (size_t i = 0; e == hipSuccess && i < bt->tev.size(); ++i) e = hipEventCreate(&bt->tev[i]);
  if (e == hipSuccess) e = hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    unsigned *hp = pl.heads.as<unsigned>();
    int cnt = pl.nb + 1;
    unsigned long long *tp = pl.times.as<unsigned long long>();
    int tcnt = (int)pl.runs.size() + 1;
    void *zp[] = {&hp, &cnt, &tp, &tcnt};
    e = hipLaunchKernel(fjmi::kZeroKernel, dim3(1), dim3(64), zp, 0, s0);
    if (e == hipSuccess) e = hipEventRecord(bt->tev[0], s0);
    for (size_t i = 0; e == hipSuccess && i < pl.runs.size(); ++i) {
      e = hipStreamWaitEvent(ss[i], bt->tev[0], 0);
      BucketRun &r = pl.runs[i];
      void *params[] = {&r.args, &r.ts};
      if (e == hipSuccess) e = hipLaunchKernel(r.fn, dim3(r.grid), dim3(r.threads), params, r.lds, ss[i]);
      if (e == hipSuccess) e = hipEventRecord(bt->tev[i + 1], ss[i]);
      if (e == hipSuccess) e = hipStreamWaitEvent(s0, bt->tev[i + 1], 0);
    }
    hipError_t e2 = hipStreamEndCapture(s0, &bt->graph);
    if (e == hipSuccess) e = e2;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0);
  for (hipStream_t st : ss)
    if (st) (void)hipStreamDestroy(st);
  if (s0) (void)hipStreamDestroy(s0);
  if (e != hipSuccess) return cleanup(fail(FJMI_EHIP, "graph capture: %s", hipGetErrorString(e)));
  *out = bt;
  return FJMI_OK;
}

}  // namespace

extern "C" {

int fjmi_batch_create(fjmi_context *ctx, int n_events, const int64_t *h_offsets, const double *d_p4,
                      const int64_t *d_offsets, const int32_t *d_p, const double *d_R, int p_uni, double R_uni,
                      int32_t *d_h1, int32_t *d_h2, double *d_hd, double *d_jets, int32_t *d_nj, fjmi_batch **out) {
  return batch_create(ctx, n_events, h_offsets, d_p4, d_offsets, d_p, d_R, p_uni, R_uni, nullptr, nullptr, d_h1, d_h2,
                      d_hd, d_jets, d_nj, nullptr, out);
}

int fjmi_batch_create_def(fjmi_context *ctx, int n_events, const int64_t *h_offsets, const double *d_p4,
                          const int64_t *d_offsets, const fjmi_jetdef *d_defs, const fjmi_jetdef *def,
                          const double *d_R, double R_uni, int32_t *d_h1, int32_t *d_h2, double *d_hd, double *d_jets,
                          int32_t *d_nj, double *d_hist_jets, fjmi_batch **out) {
  fjmi::JetDef d{};
  if (!d_defs) {
    const int rc = check_def(def, &d);
    if (rc) return rc;
  }
  static_assert(sizeof(fjmi_jetdef) == sizeof(fjmi::JetDef) && offsetof(fjmi_jetdef, scheme) == offsetof(fjmi::JetDef, scheme),
                "fjmi_jetdef is the kernels' JetDef");
  return batch_create(ctx, n_events, h_offsets, d_p4, d_offsets, nullptr, d_R, 0, R_uni,
                      reinterpret_cast<const fjmi::JetDef *>(d_defs), d_defs ? nullptr : &d, d_h1, d_h2, d_hd, d_jets,
                      d_nj, d_hist_jets, out);
}

int fjmi_batch_launch(fjmi_batch *b, void *stream) {
  if (!b || !b->exec) return fail(FJMI_EINVAL, "null batch");
  HIPCHK(hipSetDevice(b->ctx->device));
  HIPCHK(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return FJMI_OK;
}

int fjmi_batch_sync(fjmi_batch *b, void *stream) {
  if (!b) return fail(FJMI_EINVAL, "null batch");
  HIPCHK(hipSetDevice(b->ctx->device));
  HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  uint32_t st = 0;
  HIPCHK(hipMemcpy(&st, b->plan.d_status, sizeof st, hipMemcpyDeviceToHost));
  if (st) return status_error(st);
  return FJMI_OK;
}

int fjmi_batch_stats(fjmi_batch *b, fjmi_bucket_stats *out, int cap, int *n) {
  if (!b || !n) return fail(FJMI_EINVAL, "null argument");
  HIPCHK(hipSetDevice(b->ctx->device));
  *n = (int)b->plan.runs.size();
  for (int i = 0; i < *n && i < cap; ++i) {
    const BucketRun &r = b->plan.runs[i];
    fjmi_bucket_stats &o = out[i];
    o.bucket.max_n = r.max_n;
    o.bucket.engine = r.tiled ? FJMI_ENGINE_TILED : FJMI_ENGINE_DENSE;
    o.bucket.threads = r.threads;
    o.bucket.slots = r.tiled ? r.tile_q : r.kd->K;
    o.bucket.p4_global = r.tiled ? r.tile_table : (r.kd->p4g ? 1 : 0);
    o.n_events = r.n_events;
    o.grid = r.grid;
    o.particles = r.particles;
    o.sum_n2 = r.sum_n2;
    o.ms = -1.0f;
  }
  /* completion times measured by the kernels themselves (100 MHz wall clock) */
  std::vector<unsigned long long> t(b->plan.runs.size() + 1, 0);
  HIPCHK(hipMemcpy(t.data(), b->plan.times.ptr, sizeof(unsigned long long) * t.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < *n && i < cap; ++i)
    if (t[i + 1] >= t[0]) out[i].ms = (float)((double)(t[i + 1] - t[0]) / 100e3);
  return FJMI_OK;
}

int fjmi_batch_destroy(fjmi_batch *b) {
  if (!b) return FJMI_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipDeviceSynchronize();
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (hipEvent_t ev : b->tev)
    if (ev) (void)hipEventDestroy(ev);
  b->plan.release();
  delete b;
  return FJMI_OK;
}

}  // extern "C"

namespace {

/* A host-memory batch: the standard algorithms (defs and def_uni null) or
 * general jet definitions (per event in host memory, or one for all), with the
 * history momenta back into hist_jets when it is given. */
int cluster_batch_host(fjmi_context *ctx, int n_events, const int64_t *offsets, const double *p4, const int32_t *p,
                       const double *R, int p_uni, double R_uni, const fjmi::JetDef *defs,
                       const fjmi::JetDef *def_uni, int32_t *h1, int32_t *h2, double *hd, double *jets, int32_t *nj,
                       double *hist_jets) {
  const bool gen = defs || def_uni;
  if (!ctx || !offsets || n_events < 0) return fail(FJMI_EINVAL, "bad argument");
  if (n_events == 0) return FJMI_OK;
  const int64_t total = offsets[n_events];
  if (total > 0 && (!p4 || !h1 || !h2 || !hd || !jets)) return fail(FJMI_EINVAL, "null buffer");
  if (!nj) return fail(FJMI_EINVAL, "null njets");
  if (!p && (p_uni < -1 || p_uni > 1)) return fail(FJMI_EINVAL, "exponent p must be -1, 0 or 1");
  if (!R && !(R_uni > 0.0)) return fail(FJMI_EINVAL, "radius must be positive");
  std::lock_guard<std::mutex> lk(ctx->mtx);
  /* every event must fit a bucket before anything is started */
  for (int e = 0; e < n_events; ++e) {
    const int64_t n = offsets[e + 1] - offsets[e];
    if (n < 0 || offsets[e] < 0) return fail(FJMI_EINVAL, "offsets not monotonic at event %d", e);
    const fjmi_bucket *b = bucket_for(ctx, (int)std::min<int64_t>(n, INT32_MAX), FJMI_ENGINE_AUTO);
    if (!b) return fail(FJMI_ECAPACITY, "event %d has %lld particles, above the largest bucket", e, (long long)n);
    if ((gen || b->engine == FJMI_ENGINE_TILED) && n > fjmi::kTiledMaxN)
      return fail(FJMI_ECAPACITY, "event %d has %lld particles: the tiled engine takes at most %d", e, (long long)n,
                  fjmi::kTiledMaxN);
  }
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if (!ctx->host) ctx->host = new HostPath();
  HostPath &hp = *ctx->host;

  /* events longest first, cut into chunks of about equal particle counts */
  std::vector<int32_t> order(n_events);
  for (int e = 0; e < n_events; ++e) order[e] = e;
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
    return offsets[x + 1] - offsets[x] > offsets[y + 1] - offsets[y];
  });
  const int n_chunks = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)kHostMaxChunks, (int64_t)n_events,
                                                                     (total + kHostChunkParticles - 1) /
                                                                         kHostChunkParticles}));
  std::vector<int> cut(n_chunks + 1, n_events);
  cut[0] = 0;
  {
    int64_t acc = 0;
    int c = 1;
    for (int i = 0; i < n_events && c < n_chunks; ++i) {
      acc += offsets[order[i] + 1] - offsets[order[i]];
      if (acc * n_chunks >= total * c) cut[c++] = i + 1;
    }
  }
  /* on an error, the chunks still in flight are waited for and dropped, so
   * that no later call writes their results */
  auto abort_all = [&](int code) {
    for (HostSlot &sl : hp.slot)
      if (sl.busy) {
        (void)hipStreamSynchronize(sl.s);
        sl.busy = false;
      }
    return code;
  };
  for (int c = 0; c < n_chunks; ++c) { /* a slot's earlier chunk must be finished before it is reused */
    if (cut[c + 1] == cut[c]) continue;
    HostSlot &sl = hp.slot[c % kHostSlots];
    if (sl.busy && (rc = host_finish(sl, offsets, h1, h2, hd, jets, nj, hist_jets))) return abort_all(rc);
    if ((rc = host_start(ctx, sl, order.data() + cut[c], cut[c + 1] - cut[c], offsets, p4, p, R, p_uni, R_uni, defs,
                         def_uni, hist_jets != nullptr)))
      return abort_all(rc);
  }
  /* finished in completion order: the first chunk holds the largest events and
   * ends last, while the others' copies out overlap its kernels */
  for (;;) {
    HostSlot *ready = nullptr;
    bool any = false;
    for (HostSlot &sl : hp.slot) {
      if (!sl.busy) continue;
      any = true;
      const hipError_t q = hipEventQuery(sl.ev_out);
      if (q == hipSuccess) {
        ready = &sl;
        break;
      }
      if (q != hipErrorNotReady) return abort_all(fail(FJMI_EHIP, "batch chunk: %s", hipGetErrorString(q)));
    }
    if (!any) break;
    if (!ready) {
      std::this_thread::sleep_for(std::chrono::microseconds(50));
      continue;
    }
    if ((rc = host_finish(*ready, offsets, h1, h2, hd, jets, nj, hist_jets))) return abort_all(rc);
  }
  return FJMI_OK;
}

}  // namespace

extern "C" {

int fjmi_cluster_batch(fjmi_context *ctx, int n_events, const int64_t *offsets, const double *p4, const int32_t *p,
                       const double *R, int p_uni, double R_uni, int32_t *h1, int32_t *h2, double *hd, double *jets,
                       int32_t *nj) {
  return cluster_batch_host(ctx, n_events, offsets, p4, p, R, p_uni, R_uni, nullptr, nullptr, h1, h2, hd, jets, nj,
                            nullptr);
}

int fjmi_cluster_batch_def(fjmi_context *ctx, int n_events, const int64_t *offsets, const double *p4,
                           const fjmi_jetdef *defs, const fjmi_jetdef *def, const double *R, double R_uni,
                           int32_t *h1, int32_t *h2, double *hd, double *jets, int32_t *nj, double *hist_jets) {
  fjmi::JetDef d{};
  int rc;
  if (defs) {
    for (int e = 0; e < n_events; ++e) {
      fjmi::JetDef x;
      if ((rc = check_def(defs + e, &x))) return fail(rc, "event %d: %s", e, fjmi_last_error());
    }
  } else if ((rc = check_def(def, &d))) {
    return rc;
  }
  return cluster_batch_host(ctx, n_events, offsets, p4, nullptr, R, 0, R_uni,
                            reinterpret_cast<const fjmi::JetDef *>(defs), defs ? nullptr : &d, h1, h2, hd, jets, nj,
                            hist_jets);
}

}  // extern "C"
