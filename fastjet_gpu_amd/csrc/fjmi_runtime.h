/*
 * fjmi_runtime.h — host-side pieces shared by the library's translation units
 * (runtime.hip: contexts, single-event and batched paths; serve.hip: the
 * serving loop): the error reporting behind fjmi_last_error(), a growable
 * device buffer, and the per-device context.
 */
#ifndef FJMI_RUNTIME_H
#define FJMI_RUNTIME_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/fjmi.h"

namespace fjmi {
namespace rt {

/* Records a printf-style message for fjmi_last_error() (per thread) and returns code. */
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

#define HIPCHK(call)                                                                                   \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess)                                                                              \
      return ::fjmi::rt::fail(FJMI_EHIP, "%s:%d: %s: %s: %s", __FILE__, __LINE__, #call,              \
                              hipGetErrorName(e_), hipGetErrorString(e_));                             \
  } while (0)

/* Growable device buffer. */
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t need) {
    if (need <= bytes) return FJMI_OK;
    if (ptr) HIPCHK(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
    size_t b = std::max(need, (size_t)256);
    HIPCHK(hipMalloc(&ptr, b));
    bytes = b;
    return FJMI_OK;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(ptr);
  }
};

}  // namespace rt
}  // namespace fjmi

namespace fjmi {
namespace rt {

/* Device state of the one-event entry points on one stream: their scratch and
 * outputs, and the status word the kernels raise on failure.  Launches on one
 * stream are ordered, so they may share it; launches on different streams each
 * get their own. */
struct StreamState {
  DevBuf status; /* 4 bytes, zeroed at creation and by fjmi_stream_check */
  DevBuf hist1, hist2, hdij, jets, njets, p4s, coords, p4in, tscratch, hjets;
  void release_all() {
    for (DevBuf *b : {&status, &hist1, &hist2, &hdij, &jets, &njets, &p4s, &coords, &p4in, &tscratch, &hjets})
      b->release();
  }
};

}  // namespace rt
}  // namespace fjmi

struct HostPath; /* host-memory batch pipeline (runtime.hip) */

/* Per-device library context. */
struct fjmi_context {
  int device = 0;
  int n_cu = 0;
  std::vector<fjmi_bucket> buckets;
  /* set by fjmi_set_buckets: the caller's table then also decides the engine
   * of one-event calls with FJMI_ENGINE_AUTO (else the latency rule does) */
  bool caller_table = false;
  std::mutex mtx;
  /* one-event paths: per-stream state (key: the hipStream_t, nullptr for the null stream) */
  std::map<void *, fjmi::rt::StreamState> streams;
  int max_lds = 0;
  /* device address of this device's copy of the library's static scratch array
   * (fjmi_cluster_event_device_noalloc), from hipGetSymbolAddress at creation */
  unsigned char *static_scratch = nullptr;
  HostPath *host = nullptr; /* fjmi_cluster_batch's staging, created on first use */
};

#endif
