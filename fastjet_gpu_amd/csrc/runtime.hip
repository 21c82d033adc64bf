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
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
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

struct fjmi_batch {
  fjmi_context *ctx = nullptr;
  int n_events = 0;
  DevBuf queue, heads, p4s, coords, tscratch, times;
  std::vector<BucketRun> runs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t *d_status = nullptr;
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
      if (b.threads != 64) return fail(FJMI_EINVAL, "tiled rows run 64 threads");
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
                       int32_t *nj, unsigned char *pj, hipStream_t stream) {
  /* engine forced to tiled: the table's tiled row for n, else tile side R with the 512-tile variant */
  const fjmi_bucket *b = bucket_for(ctx, n, engine);
  if (!b && engine == FJMI_ENGINE_TILED) b = &kTiledRow;
  fjmi_bucket forced;
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
  if (tiled) {
    void *params[] = {&a, &ts};
    HIPCHK(hipLaunchKernel(fjmi::kTiledKernels[b->p4_global], dim3(1), dim3(64), params, 0, stream));
  } else if ((rc = launch_dense(*kd, 1, dyn_lds_bytes(*kd), a, stream))) {
    return rc;
  }
  return FJMI_OK;
}

/* The fjmi_status of a kernel status word. */
int status_error(uint32_t status) {
  if (status & fjmi::STATUS_INTERNAL)
    return fail(FJMI_EINTERNAL, "a clustering kernel found an inconsistent state (status 0x%x)", status);
  return fail(FJMI_ECAPACITY, "a clustering kernel reported an event beyond its capacity (status 0x%x)", status);
}

int check_engine(int engine) {
  if (engine != FJMI_ENGINE_AUTO && engine != FJMI_ENGINE_DENSE && engine != FJMI_ENGINE_TILED)
    return fail(FJMI_EUNSUPPORTED, "engine %d not available for this entry point", engine);
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
  return fjmi::kNumDenseKernels + fjmi::kNumTiledKernels;
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
  delete ctx;
  return FJMI_OK;
}

int fjmi_set_buckets(fjmi_context *ctx, const fjmi_bucket *rows, int n_rows) {
  if (!ctx || !rows) return fail(FJMI_EINVAL, "null argument");
  int rc = check_table(rows, n_rows);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(ctx->mtx);
  ctx->buckets.assign(rows, rows + n_rows);
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
  /* the table's tiled row for n, else tile side R with the 512-tile variant */
  const fjmi_bucket *b = bucket_for(ctx, n, FJMI_ENGINE_TILED);
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
  void *params[] = {&d_offsets, &n_single, &d_p4, &d_hist_p1, &d_hist_p2, &d_hist_dij, &mode, &param, &d_work,
                    &d_out_jets, &d_out_njets};
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

int fjmi_cluster_event(fjmi_context *ctx, const double *p4, int n, int p, double R, int engine, int32_t *h1,
                       int32_t *h2, double *hd, double *jets, int32_t *nj) {
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
      (rc = st->hdij.ensure(8 * nn)) || (rc = st->jets.ensure(32 * nn)) || (rc = st->njets.ensure(4)))
    return rc;
  HIPCHK(hipMemcpy(st->p4in.ptr, p4, 32 * nn, hipMemcpyHostToDevice));
  if ((rc = cluster_one_device(ctx, *st, engine, st->p4in.as<double>(), 4, 0, n, p, R, st->hist1.as<int32_t>(),
                               st->hist2.as<int32_t>(), st->hdij.as<double>(), st->jets.as<double>(),
                               st->njets.as<int32_t>(), nullptr, nullptr)))
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
  return FJMI_OK;
}

/* ---- batched path ---- */

int fjmi_batch_create(fjmi_context *ctx, int n_events, const int64_t *h_offsets, const double *d_p4,
                      const int64_t *d_offsets, const int32_t *d_p, const double *d_R, int p_uni, double R_uni,
                      int32_t *d_h1, int32_t *d_h2, double *d_hd, double *d_jets, int32_t *d_nj, fjmi_batch **out) {
  if (!ctx || !out || !h_offsets || n_events < 0) return fail(FJMI_EINVAL, "bad argument");
  *out = nullptr;
  if (!d_p && (p_uni < -1 || p_uni > 1)) return fail(FJMI_EINVAL, "exponent p must be -1, 0 or 1");
  if (!d_R && !(R_uni > 0.0)) return fail(FJMI_EINVAL, "radius must be positive");
  std::lock_guard<std::mutex> lk(ctx->mtx);
  int rc;
  if ((rc = set_device(ctx))) return rc;
  const int nb = (int)ctx->buckets.size();
  std::vector<std::vector<int32_t>> lists(nb);
  int64_t max_p4g_total = 0;
  bool any_tiled = false;
  for (int e = 0; e < n_events; ++e) {
    const int64_t n = h_offsets[e + 1] - h_offsets[e];
    if (n < 0 || h_offsets[e] < 0) return fail(FJMI_EINVAL, "offsets not monotonic at event %d", e);
    int b = 0;
    while (b < nb && n > ctx->buckets[b].max_n) ++b;
    if (b == nb) return fail(FJMI_ECAPACITY, "event %d has %lld particles, above the largest bucket", e, (long long)n);
    if (ctx->buckets[b].engine == FJMI_ENGINE_TILED && n > fjmi::kTiledMaxN)
      return fail(FJMI_ECAPACITY, "event %d has %lld particles: the tiled engine takes at most %d", e, (long long)n,
                  fjmi::kTiledMaxN);
    lists[b].push_back(e);
    if (ctx->buckets[b].engine == FJMI_ENGINE_TILED)
      any_tiled = true;
    else if (ctx->buckets[b].p4_global)
      max_p4g_total = h_offsets[n_events];
  }
  auto *bt = new fjmi_batch();
  bt->ctx = ctx;
  bt->n_events = n_events;
  auto cleanup = [&](int code) {
    fjmi_batch_destroy(bt);
    return code;
  };
  /* cost-ordered queues (longest first): the last events to start are the cheapest */
  std::vector<int32_t> flat;
  std::vector<size_t> start(nb);
  for (int b = 0; b < nb; ++b) {
    std::stable_sort(lists[b].begin(), lists[b].end(), [&](int32_t x, int32_t y) {
      return h_offsets[x + 1] - h_offsets[x] > h_offsets[y + 1] - h_offsets[y];
    });
    start[b] = flat.size();
    flat.insert(flat.end(), lists[b].begin(), lists[b].end());
  }
  if ((rc = bt->queue.ensure(sizeof(int32_t) * std::max<size_t>(flat.size(), 1)))) return cleanup(rc);
  if ((rc = bt->heads.ensure(sizeof(uint32_t) * (nb + 1)))) return cleanup(rc);
  if ((rc = bt->times.ensure(sizeof(unsigned long long) * (nb + 1)))) return cleanup(rc);
  if (max_p4g_total > 0) {
    if ((rc = bt->p4s.ensure(32 * (size_t)max_p4g_total))) return cleanup(rc);
    if ((rc = bt->coords.ensure(32 * (size_t)max_p4g_total))) return cleanup(rc);
  }
  if (any_tiled) {
    const int64_t P = h_offsets[n_events];
    if ((rc = bt->tscratch.ensure(fjmi::tiled_scratch_bytes(P, n_events, fjmi::kTiledPoolPerEvent))))
      return cleanup(rc);
  }
  if (!flat.empty()) {
    hipError_t e = hipMemcpy(bt->queue.ptr, flat.data(), sizeof(int32_t) * flat.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return cleanup(fail(FJMI_EHIP, "queue upload: %s", hipGetErrorString(e)));
  }
  uint32_t *heads = bt->heads.as<uint32_t>();
  bt->d_status = heads + nb;
  for (int b = nb - 1; b >= 0; --b) { /* largest multiplicities first */
    if (lists[b].empty()) continue;
    const fjmi_bucket &bk = ctx->buckets[b];
    BucketRun r;
    r.tiled = bk.engine == FJMI_ENGINE_TILED;
    r.kd = r.tiled ? nullptr : find_kernel(bk.threads, bk.slots, bk.p4_global);
    r.fn = r.tiled ? fjmi::kTiledKernels[bk.p4_global] : r.kd->fn;
    r.threads = r.tiled ? 64 : r.kd->T;
    r.lds = r.tiled ? 0 : dyn_lds_bytes(*r.kd);
    r.ts = r.tiled ? fjmi::tiled_scratch_carve(bt->tscratch.ptr, h_offsets[n_events], fjmi::kTiledPoolPerEvent)
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
    a.queue = bt->queue.as<int32_t>() + start[b];
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
    a.p4_scratch = bt->p4s.as<double>();
    a.coord_scratch = bt->coords.as<double>();
    a.status = bt->d_status;
    a.t_end = bt->times.as<unsigned long long>() + 1 + bt->runs.size();
    a.tile_q = r.tile_q;
    r.args = a;
    bt->runs.push_back(r);
  }
  /* capture: zero the heads and status, then every bucket's persistent kernel on
   * its own forked stream so the buckets run concurrently */
  hipStream_t s0 = nullptr;
  std::vector<hipStream_t> ss(bt->runs.size(), nullptr);
  bt->tev.assign(bt->runs.size() + 1, nullptr);
  hipError_t e = hipStreamCreateWithFlags(&s0, hipStreamNonBlocking);
  for (size_t i = 0; e == hipSuccess && i < ss.size(); ++i) e = hipStreamCreateWithFlags(&ss[i], hipStreamNonBlocking);
  for (size_t i = 0; e == hipSuccess && i < bt->tev.size(); ++i) e = hipEventCreate(&bt->tev[i]);
  if (e == hipSuccess) e = hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    unsigned *hp = heads;
    int cnt = nb + 1;
    unsigned long long *tp = bt->times.as<unsigned long long>();
    int tcnt = (int)bt->runs.size() + 1;
    void *zp[] = {&hp, &cnt, &tp, &tcnt};
    e = hipLaunchKernel(fjmi::kZeroKernel, dim3(1), dim3(64), zp, 0, s0);
    if (e == hipSuccess) e = hipEventRecord(bt->tev[0], s0);
    for (size_t i = 0; e == hipSuccess && i < bt->runs.size(); ++i) {
      e = hipStreamWaitEvent(ss[i], bt->tev[0], 0);
      BucketRun &r = bt->runs[i];
      void *params[] = {&r.args, &r.ts};
      if (e == hipSuccess) e = hipLaunchKernel(r.fn, dim3(r.grid), dim3(r.threads), params, r.lds, ss[i]);
      if (e == hipSuccess) e = hipEventRecord(bt->tev[i + 1], ss[i]);
      if (e == hipSuccess) e = hipStreamWaitEvent(s0, bt->tev[i + 1], 0);
    }
    hipError_t e2 = hipStreamEndCapture(s0, &bt->graph);
    if (e == hipSuccess) e = e2;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0);
  for (hipStream_t s : ss)
    if (s) (void)hipStreamDestroy(s);
  if (s0) (void)hipStreamDestroy(s0);
  if (e != hipSuccess) return cleanup(fail(FJMI_EHIP, "graph capture: %s", hipGetErrorString(e)));
  *out = bt;
  return FJMI_OK;
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
  HIPCHK(hipMemcpy(&st, b->d_status, sizeof st, hipMemcpyDeviceToHost));
  if (st) return status_error(st);
  return FJMI_OK;
}

int fjmi_batch_stats(fjmi_batch *b, fjmi_bucket_stats *out, int cap, int *n) {
  if (!b || !n) return fail(FJMI_EINVAL, "null argument");
  HIPCHK(hipSetDevice(b->ctx->device));
  *n = (int)b->runs.size();
  for (int i = 0; i < *n && i < cap; ++i) {
    const BucketRun &r = b->runs[i];
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
  std::vector<unsigned long long> t(b->runs.size() + 1, 0);
  HIPCHK(hipMemcpy(t.data(), b->times.ptr, sizeof(unsigned long long) * t.size(), hipMemcpyDeviceToHost));
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
  b->queue.release();
  b->heads.release();
  b->p4s.release();
  b->coords.release();
  b->tscratch.release();
  b->times.release();
  delete b;
  return FJMI_OK;
}

int fjmi_cluster_batch(fjmi_context *ctx, int n_events, const int64_t *offsets, const double *p4, const int32_t *p,
                       const double *R, int p_uni, double R_uni, int32_t *h1, int32_t *h2, double *hd, double *jets,
                       int32_t *nj) {
  if (!ctx || !offsets || n_events < 0) return fail(FJMI_EINVAL, "bad argument");
  if (n_events == 0) return FJMI_OK;
  const size_t total = (size_t)offsets[n_events];
  if (total > 0 && (!p4 || !h1 || !h2 || !hd || !jets)) return fail(FJMI_EINVAL, "null buffer");
  if (!nj) return fail(FJMI_EINVAL, "null njets");
  HIPCHK(hipSetDevice(ctx->device));
  DevBuf dp4, doff, dpp, dR, d1, d2, dd, dj, dn;
  auto rel = [&]() {
    for (DevBuf *x : {&dp4, &doff, &dpp, &dR, &d1, &d2, &dd, &dj, &dn}) x->release();
  };
  int rc = 0;
  const size_t tb = std::max<size_t>(total, 1);
  if ((rc = dp4.ensure(32 * tb)) || (rc = doff.ensure(8 * (size_t)(n_events + 1))) || (rc = d1.ensure(4 * tb)) ||
      (rc = d2.ensure(4 * tb)) || (rc = dd.ensure(8 * tb)) || (rc = dj.ensure(32 * tb)) ||
      (rc = dn.ensure(4 * (size_t)n_events)) || (p && (rc = dpp.ensure(4 * (size_t)n_events))) ||
      (R && (rc = dR.ensure(8 * (size_t)n_events)))) {
    rel();
    return rc;
  }
  hipError_t e = hipSuccess;
  if (total) e = hipMemcpy(dp4.ptr, p4, 32 * total, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(doff.ptr, offsets, 8 * (size_t)(n_events + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess && p) e = hipMemcpy(dpp.ptr, p, 4 * (size_t)n_events, hipMemcpyHostToDevice);
  if (e == hipSuccess && R) e = hipMemcpy(dR.ptr, R, 8 * (size_t)n_events, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    rel();
    return fail(FJMI_EHIP, "upload: %s", hipGetErrorString(e));
  }
  fjmi_batch *b = nullptr;
  rc = fjmi_batch_create(ctx, n_events, offsets, dp4.as<double>(), doff.as<int64_t>(), p ? dpp.as<int32_t>() : nullptr,
                         R ? dR.as<double>() : nullptr, p_uni, R_uni, d1.as<int32_t>(), d2.as<int32_t>(),
                         dd.as<double>(), dj.as<double>(), dn.as<int32_t>(), &b);
  if (!rc) rc = fjmi_batch_launch(b, nullptr);
  if (!rc) rc = fjmi_batch_sync(b, nullptr);
  if (b) fjmi_batch_destroy(b);
  if (!rc) {
    e = hipSuccess;
    if (total) {
      e = hipMemcpy(h1, d1.ptr, 4 * total, hipMemcpyDeviceToHost);
      if (e == hipSuccess) e = hipMemcpy(h2, d2.ptr, 4 * total, hipMemcpyDeviceToHost);
      if (e == hipSuccess) e = hipMemcpy(hd, dd.ptr, 8 * total, hipMemcpyDeviceToHost);
      if (e == hipSuccess) e = hipMemcpy(jets, dj.ptr, 32 * total, hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess) e = hipMemcpy(nj, dn.ptr, 4 * (size_t)n_events, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(FJMI_EHIP, "download: %s", hipGetErrorString(e));
  }
  rel();
  return rc;
}

}  // extern "C"
