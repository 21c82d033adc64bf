/*
 * serve.hip — the serving loop of include/fjmi.h (fjmi_server_*): continuous
 * batching of arriving events under a fixed memory budget.
 *
 * Memory (allocated once, at create):
 *   device arena   per particle: tiled-engine scratch (a 64-byte record, the
 *                  four-momentum and two indices), history (16 B) and jets
 *                  (32 B); per event: its tile-segment pool range (3 entries
 *                  per particle + 2048);
 *   pinned host    per particle: input four-momenta (32 B), result history
 *                  (16 B) and jets (32 B); per slot: descriptor ring, completion
 *                  ring and njets (fine-grained, so the GPU reads the host's
 *                  stores and the host reads the GPU's without copies).
 * Particle ranges and pool ranges are handed out best-fit and coalesced on
 * release; an event is admitted when its particles, its pool range and a slot
 * are all free, otherwise submit() says FJMI_EAGAIN.
 *
 * Protocol (one producer thread on the host, persistent workgroups on the GPU):
 *   submit   copy the input into pinned memory, write the descriptor at
 *            ring[published % slots], then release-store published + 1 into
 *            ServeHost::tail;
 *   GPU      idle workgroups claim descriptors in order (device-memory head,
 *            device mirror of the tail refreshed by one wave at a time), cluster
 *            the event from the arenas, copy its results to pinned memory and
 *            push slot + 1 onto the completion ring after a system-scope fence;
 *   poll     read the completion ring in order (acquire), hand out results,
 *            clear the entries; relaunch the kernel if it has left (idle
 *            timeout) while published events are still unfinished;
 *   release  return the event's ranges and slot.
 * At most `slots` events are admitted at once, so the completion ring cannot
 * wrap onto an entry still in use.  The descriptor ring could, if a workgroup
 * stalled between claiming descriptor h and copying it while later events
 * completed and were released: each descriptor therefore carries a sequence tag
 * that the GPU clears after its copy, and submit() reuses a cell only once it
 * has been cleared (FJMI_EAGAIN until then).
 */
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cstring>
#include <map>
#include <set>
#include <vector>

#include "fjmi_internal.h"
#include "fjmi_kernels.h"
#include "fjmi_runtime.h"

using fjmi::rt::fail;

namespace {

/* Best-fit allocator of ranges within [0, cap), coalescing on release. */
class RangeAlloc {
 public:
  void reset(int64_t cap) {
    by_start_.clear();
    by_size_.clear();
    if (cap > 0) insert(0, cap);
  }
  int64_t alloc(int64_t n) {
    auto it = by_size_.lower_bound({n, INT64_MIN});
    if (it == by_size_.end()) return -1;
    const int64_t len = it->first, start = it->second;
    by_size_.erase(it);
    by_start_.erase(start);
    if (len > n) insert(start + n, len - n);
    return start;
  }
  void release(int64_t start, int64_t n) {
    auto next = by_start_.lower_bound(start);
    if (next != by_start_.end() && start + n == next->first) {
      n += next->second;
      by_size_.erase({next->second, next->first});
      next = by_start_.erase(next);
    }
    if (next != by_start_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == start) {
        start = prev->first;
        n += prev->second;
        by_size_.erase({prev->second, prev->first});
        by_start_.erase(prev);
      }
    }
    insert(start, n);
  }

 private:
  void insert(int64_t start, int64_t n) {
    by_start_[start] = n;
    by_size_.insert({n, start});
  }
  std::map<int64_t, int64_t> by_start_;
  std::set<std::pair<int64_t, int64_t>> by_size_;
};

constexpr int64_t kPoolPerParticle = (int64_t)fjmi::kTiledPoolPerParticle; /* tile-segment pool entries */
constexpr int64_t kPoolEntryBytes = (int64_t)fjmi::kTiledPoolEntryBytes;
constexpr int64_t kDeviceBytesPerParticle = /* tiled scratch less its pool + history + jets */
    (int64_t)fjmi::kTiledBytesPerParticle - kPoolPerParticle * kPoolEntryBytes + 16 + 32;
constexpr int64_t kHostBytesPerParticle = 32 + 16 + 32;    /* input + result history + result jets */
constexpr int64_t kHostBytesPerSlot = sizeof(fjmi::ServeDesc) + 4 + 4;

struct Slot {
  uint64_t tag = 0;
  int64_t base = 0, pool_off = 0;
  int32_t n = 0;
  int32_t state = 0; /* 0 free, 1 admitted, 2 completed (handed out) */
};

int64_t alloc_len(int32_t n) { return n > 0 ? n : 1; }
int64_t pool_len(int32_t n) { return kPoolPerParticle * n + fjmi::kTiledPoolPerEvent; }

}  // namespace

struct fjmi_server {
  fjmi_context *ctx = nullptr;
  int64_t P = 0; /* particle capacity */
  int64_t Q = 0; /* pool entries */
  uint32_t E = 0; /* slots */
  int variant = 0;
  /* pinned host memory (fine-grained) */
  void *host_block = nullptr;
  size_t host_bytes = 0;
  fjmi::ServeHost *host = nullptr;
  fjmi::ServeDesc *ring = nullptr;
  int32_t *comp = nullptr;
  int32_t *res_nj = nullptr;
  double *in_p4 = nullptr;
  int32_t *res_p1 = nullptr, *res_p2 = nullptr;
  double *res_dij = nullptr, *res_jets = nullptr;
  /* device memory */
  void *dev_block = nullptr;
  size_t dev_bytes = 0;
  /* host-side bookkeeping */
  RangeAlloc parts, pool;
  std::vector<int32_t> free_slots;
  std::vector<Slot> slots;
  uint64_t published = 0, comp_next = 0;
  int64_t in_flight = 0, particles_in_flight = 0, peak_particles = 0, completed = 0, launches = 0;
  /* launch */
  hipStream_t stream = nullptr, copy_stream = nullptr;
  hipEvent_t done_ev = nullptr;
  unsigned long long *dctl = nullptr; /* the kernels' device control words */
  bool running = false;
  std::chrono::steady_clock::time_point last_check;
  const void *fn = nullptr;
  int grid = 0;
  fjmi::LaunchArgs args{};
  fjmi::TiledScratch ts{};
};

namespace {

int host_ptr(void *h, void **d) {
  HIPCHK(hipHostGetDevicePointer(d, h, 0));
  return FJMI_OK;
}

void free_server(fjmi_server *s) {
  if (s->running && s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->done_ev) (void)hipEventDestroy(s->done_ev);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
  if (s->dev_block) (void)hipFree(s->dev_block);
  if (s->host_block) (void)hipHostFree(s->host_block);
  delete s;
}

/* Starts a serving launch when none is running and published events are unfinished. */
int ensure_running(fjmi_server *s, bool force_check) {
  if (s->running) {
    const auto now = std::chrono::steady_clock::now();
    if (!force_check && now - s->last_check < std::chrono::microseconds(200)) return FJMI_OK;
    s->last_check = now;
    const hipError_t q = hipEventQuery(s->done_ev);
    if (q == hipErrorNotReady) return FJMI_OK;
    if (q != hipSuccess) return fail(FJMI_EHIP, "serving kernel: %s", hipGetErrorString(q));
    s->running = false;
  }
  if ((int64_t)(s->published - s->comp_next) <= 0) return FJMI_OK;
  s->args.serve.epoch = (unsigned long long)s->launches + 1;
  void *params[] = {&s->args, &s->ts};
  HIPCHK(hipLaunchKernel(s->fn, dim3(s->grid), dim3(64), params, 0, s->stream));
  HIPCHK(hipEventRecord(s->done_ev, s->stream));
  s->running = true;
  s->launches += 1;
  s->last_check = std::chrono::steady_clock::now();
  return FJMI_OK;
}

}  // namespace

extern "C" {

int fjmi_server_create(fjmi_context *ctx, const fjmi_server_config *cfg_in, fjmi_server **out) {
  if (!ctx || !out) return fail(FJMI_EINVAL, "null argument");
  *out = nullptr;
  fjmi_server_config cfg{};
  if (cfg_in) cfg = *cfg_in;
  const int64_t dev_budget = cfg.device_budget > 0 ? cfg.device_budget : (int64_t)16 << 30;
  const int64_t host_budget = cfg.host_budget > 0 ? cfg.host_budget : (int64_t)8 << 30;
  const int64_t E = cfg.max_events > 0 ? cfg.max_events : 65536;
  if (cfg.variant < 0 || cfg.variant >= fjmi::kNumTiledKernels)
    return fail(FJMI_EINVAL, "variant must be 0..%d", fjmi::kNumTiledKernels - 1);
  if (cfg.tile_q != 0 && (cfg.tile_q < 4 || cfg.tile_q > 16))
    return fail(FJMI_EINVAL, "tile_q is the tile side in quarters of R, 4..16 (0 = 4)");
  if (E > (1 << 30)) return fail(FJMI_EINVAL, "max_events too large");
  const int64_t ppe = fjmi::kTiledPoolPerEvent;
  const int64_t fixed_dev = E * ppe * kPoolEntryBytes + 4096;
  const int64_t fixed_host = E * kHostBytesPerSlot + 4096;
  const int64_t P_dev = (dev_budget - fixed_dev) / (kDeviceBytesPerParticle + kPoolEntryBytes * kPoolPerParticle);
  const int64_t P_host = (host_budget - fixed_host) / kHostBytesPerParticle;
  const int64_t P = std::min(P_dev, P_host);
  if (P < 64) return fail(FJMI_EINVAL, "budgets too small for %lld event slots", (long long)E);

  auto *s = new fjmi_server();
  s->ctx = ctx;
  s->P = P;
  s->Q = kPoolPerParticle * P + E * ppe;
  s->E = (uint32_t)E;
  s->variant = cfg.variant;
  int rc = FJMI_OK;
  auto bail = [&](int code) {
    free_server(s);
    return code;
  };
  if (hipSetDevice(ctx->device) != hipSuccess) return bail(fail(FJMI_EHIP, "hipSetDevice failed"));

  /* pinned host block: host words | ring | comp | res_nj | in_p4 | res_p1 | res_p2 | res_dij | res_jets */
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t at = (off + 255) & ~(size_t)255;
    off = at + bytes;
    return at;
  };
  const size_t o_host = carve(sizeof(fjmi::ServeHost)), o_ring = carve(sizeof(fjmi::ServeDesc) * E),
               o_comp = carve(4 * E), o_nj = carve(4 * E), o_in = carve(32 * (size_t)P),
               o_p1 = carve(4 * (size_t)P), o_p2 = carve(4 * (size_t)P), o_dij = carve(8 * (size_t)P),
               o_jets = carve(32 * (size_t)P);
  s->host_bytes = off;
  hipError_t e = hipHostMalloc(&s->host_block, s->host_bytes, hipHostMallocCoherent | hipHostMallocMapped);
  if (e != hipSuccess) {
    s->host_block = nullptr;
    return bail(fail(FJMI_ENOMEM, "pinned host memory (%zu bytes): %s", s->host_bytes, hipGetErrorString(e)));
  }
  char *hb = static_cast<char *>(s->host_block);
  std::memset(hb, 0, o_in); /* control words, rings and njets start at zero */
  s->host = reinterpret_cast<fjmi::ServeHost *>(hb + o_host);
  s->ring = reinterpret_cast<fjmi::ServeDesc *>(hb + o_ring);
  s->comp = reinterpret_cast<int32_t *>(hb + o_comp);
  s->res_nj = reinterpret_cast<int32_t *>(hb + o_nj);
  s->in_p4 = reinterpret_cast<double *>(hb + o_in);
  s->res_p1 = reinterpret_cast<int32_t *>(hb + o_p1);
  s->res_p2 = reinterpret_cast<int32_t *>(hb + o_p2);
  s->res_dij = reinterpret_cast<double *>(hb + o_dij);
  s->res_jets = reinterpret_cast<double *>(hb + o_jets);
  void *hd = nullptr;
  if ((rc = host_ptr(s->host_block, &hd))) return bail(rc);
  char *hdb = static_cast<char *>(hd);

  /* device block: control words | status | history | jets | tiled scratch */
  const size_t d_ctl = 0, d_status = 8 * (size_t)fjmi::kServeCtlWords, d_p1 = (d_status + 64 + 255) & ~(size_t)255,
               d_p2 = d_p1 + ((4 * (size_t)P + 255) & ~(size_t)255),
               d_dij = d_p2 + ((4 * (size_t)P + 255) & ~(size_t)255),
               d_jets = d_dij + ((8 * (size_t)P + 255) & ~(size_t)255),
               d_ts = d_jets + ((32 * (size_t)P + 255) & ~(size_t)255);
  s->dev_bytes = d_ts + fjmi::tiled_scratch_bytes(P, E, (int)ppe);
  e = hipMalloc(&s->dev_block, s->dev_bytes);
  if (e != hipSuccess) {
    s->dev_block = nullptr;
    return bail(fail(FJMI_ENOMEM, "device arena (%zu bytes): %s", s->dev_bytes, hipGetErrorString(e)));
  }
  char *db = static_cast<char *>(s->dev_block);
  if (hipMemset(db, 0, d_p1) != hipSuccess) return bail(fail(FJMI_EHIP, "hipMemset failed"));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&s->done_ev, hipEventDisableTiming) != hipSuccess)
    return bail(fail(FJMI_EHIP, "stream/event creation failed"));

  s->fn = fjmi::kTiledServeKernels[cfg.variant];
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, s->fn, 64, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  s->grid = per_cu * ctx->n_cu;
  s->ts = fjmi::tiled_scratch_carve(db + d_ts, P, (int)ppe);

  fjmi::LaunchArgs &a = s->args;
  std::memset(&a, 0, sizeof a);
  a.in = reinterpret_cast<const double *>(hdb + o_in);
  a.in_stride = 4;
  a.in_off = 0;
  a.hist_p1 = reinterpret_cast<int32_t *>(db + d_p1);
  a.hist_p2 = reinterpret_cast<int32_t *>(db + d_p2);
  a.hist_dij = reinterpret_cast<double *>(db + d_dij);
  a.jets = reinterpret_cast<double *>(db + d_jets);
  a.status = reinterpret_cast<unsigned int *>(db + d_status);
  a.stage_chunk = fjmi::kStageChunk;
  a.tile_q = cfg.tile_q;
  fjmi::ServeArgs &sv = a.serve;
  sv.host = reinterpret_cast<fjmi::ServeHost *>(hdb + o_host);
  sv.ring = reinterpret_cast<const fjmi::ServeDesc *>(hdb + o_ring);
  sv.comp = reinterpret_cast<int32_t *>(hdb + o_comp);
  sv.res_nj = reinterpret_cast<int32_t *>(hdb + o_nj);
  sv.res_p1 = reinterpret_cast<int32_t *>(hdb + o_p1);
  sv.res_p2 = reinterpret_cast<int32_t *>(hdb + o_p2);
  sv.res_dij = reinterpret_cast<double *>(hdb + o_dij);
  sv.res_jets = reinterpret_cast<double *>(hdb + o_jets);
  sv.dev = reinterpret_cast<unsigned long long *>(db + d_ctl);
  s->dctl = sv.dev;
  sv.ring_cap = (unsigned)E;
  sv.idle_ticks = (unsigned long long)(cfg.idle_us > 0 ? cfg.idle_us : 20000) * 100ull; /* 100 MHz clock */

  s->parts.reset(P);
  s->pool.reset(s->Q);
  s->slots.assign(E, Slot{});
  s->free_slots.resize(E);
  for (int64_t i = 0; i < E; ++i) s->free_slots[i] = (int32_t)(E - 1 - i);
  *out = s;
  return FJMI_OK;
}

int fjmi_server_submit(fjmi_server *s, const double *p4, int n, int p, double R, uint64_t tag) {
  if (!s || (n > 0 && !p4)) return fail(FJMI_EINVAL, "null argument");
  if (n < 0) return fail(FJMI_EINVAL, "negative particle count");
  if (p < -1 || p > 1) return fail(FJMI_EINVAL, "exponent p must be -1, 0 or 1");
  if (!(R > 0.0) || !(R < 1e150)) return fail(FJMI_EINVAL, "radius must be positive and finite");
  if (n > s->P) return fail(FJMI_ECAPACITY, "event of %d particles exceeds the server's %lld", n, (long long)s->P);
  if (n > fjmi::kTiledMaxN)
    return fail(FJMI_ECAPACITY, "event of %d particles: the tiled engine takes at most %d", n, fjmi::kTiledMaxN);
  if (s->free_slots.empty()) return fail(FJMI_EAGAIN, "all %u event slots in use", s->E);
  fjmi::ServeDesc &d = s->ring[s->published % s->E];
  if (__atomic_load_n(&d.seq, __ATOMIC_ACQUIRE) != 0u) /* the GPU has not yet copied the descriptor E back */
    return fail(FJMI_EAGAIN, "descriptor ring full");
  const int64_t base = s->parts.alloc(alloc_len(n));
  if (base < 0) return fail(FJMI_EAGAIN, "particle arena full");
  const int64_t pool_off = s->pool.alloc(pool_len(n));
  if (pool_off < 0) {
    s->parts.release(base, alloc_len(n));
    return fail(FJMI_EAGAIN, "pool arena full");
  }
  const int32_t slot = s->free_slots.back();
  s->free_slots.pop_back();
  Slot &sl = s->slots[slot];
  sl.tag = tag;
  sl.base = base;
  sl.pool_off = pool_off;
  sl.n = n;
  sl.state = 1;
  if (n > 0) std::memcpy(s->in_p4 + 4 * base, p4, 32 * (size_t)n);
  d.base = base;
  d.pool_off = pool_off;
  d.n = n;
  d.p = p;
  d.R = R;
  d.slot = slot;
  __atomic_store_n(&d.seq, fjmi::serve_seq(s->published), __ATOMIC_RELAXED); /* ordered by the release below */
  s->published += 1;
  __atomic_store_n(&s->host->tail, (unsigned long long)s->published, __ATOMIC_RELEASE);
  s->in_flight += 1;
  s->particles_in_flight += n;
  s->peak_particles = std::max(s->peak_particles, s->particles_in_flight);
  return ensure_running(s, !s->running);
}

int fjmi_server_poll(fjmi_server *s, fjmi_server_result *out, int cap, int *n_out) {
  if (!s || !n_out || (cap > 0 && !out)) return fail(FJMI_EINVAL, "null argument");
  int k = 0;
  while (k < cap) {
    int32_t *cell = s->comp + (s->comp_next % s->E);
    const int32_t v = __atomic_load_n(cell, __ATOMIC_ACQUIRE);
    if (v == 0) break;
    *cell = 0;
    s->comp_next += 1;
    const int32_t slot = v - 1;
    if (slot < 0 || slot >= (int32_t)s->E || s->slots[slot].state != 1)
      return fail(FJMI_EHIP, "serving: completion for slot %d that is not in flight", slot);
    Slot &sl = s->slots[slot];
    sl.state = 2;
    fjmi_server_result &r = out[k++];
    r.tag = sl.tag;
    r.handle = slot;
    r.n = sl.n;
    r.njets = __atomic_load_n(s->res_nj + slot, __ATOMIC_ACQUIRE);
    r.pad = 0;
    r.hist_p1 = s->res_p1 + sl.base;
    r.hist_p2 = s->res_p2 + sl.base;
    r.hist_dij = s->res_dij + sl.base;
    r.jets = s->res_jets + 4 * sl.base;
    s->completed += 1;
  }
  *n_out = k;
  return ensure_running(s, false);
}

int fjmi_server_release(fjmi_server *s, int32_t handle) {
  if (!s) return fail(FJMI_EINVAL, "null argument");
  if (handle < 0 || handle >= (int32_t)s->E || s->slots[handle].state != 2)
    return fail(FJMI_EINVAL, "handle %d is not a completed event", handle);
  Slot &sl = s->slots[handle];
  s->parts.release(sl.base, alloc_len(sl.n));
  s->pool.release(sl.pool_off, pool_len(sl.n));
  s->in_flight -= 1;
  s->particles_in_flight -= sl.n;
  sl.state = 0;
  s->free_slots.push_back(handle);
  return FJMI_OK;
}

int fjmi_server_info_get(fjmi_server *s, fjmi_server_info *o) {
  if (!s || !o) return fail(FJMI_EINVAL, "null argument");
  o->particle_capacity = s->P;
  o->event_capacity = (int32_t)s->E;
  o->grid = s->grid;
  o->device_bytes = (int64_t)s->dev_bytes;
  o->host_bytes = (int64_t)s->host_bytes;
  o->submitted = (int64_t)s->published;
  o->completed = s->completed;
  o->in_flight = s->in_flight;
  o->particles_in_flight = s->particles_in_flight;
  o->peak_particles_in_flight = s->peak_particles;
  o->launches = s->launches;
  /* read while the kernels may run: a copy on its own stream */
  unsigned long long peak = 0;
  HIPCHK(hipMemcpyAsync(&peak, s->dctl + fjmi::kServePeakBusy, sizeof peak, hipMemcpyDeviceToHost,
                         s->copy_stream));
  HIPCHK(hipStreamSynchronize(s->copy_stream));
  o->gpu_peak_busy = (int64_t)peak;
  return FJMI_OK;
}

int fjmi_server_destroy(fjmi_server *s) {
  if (!s) return FJMI_OK;
  int rc = FJMI_OK;
  __atomic_store_n(&s->host->stop, 1u, __ATOMIC_RELEASE);
  if (s->running) {
    const hipError_t e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) rc = fail(FJMI_EHIP, "serving kernel: %s", hipGetErrorString(e));
    s->running = false;
  }
  free_server(s);
  return rc;
}

}  // extern "C"
