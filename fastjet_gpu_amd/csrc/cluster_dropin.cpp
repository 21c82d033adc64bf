/*
 * cluster_dropin.cpp — the drop-in engine symbol of include/cluster.h:
 *
 *   void cluster(PseudoJet *particles, int size, Algorithm algo, double r);
 *
 * Same contract as the reference's cluster() (cluster.h:6, implemented per engine
 * at tri_matrix.cu:286, n_array.cu:229, grid.cu:580, tri_matrix_empty.cu:257):
 * `particles` is a device pointer, the clustering is done in place, records with
 * isJet hold the jets, and errors abort after printing "file, line:" and the
 * failing call (the reference's cudaCheck, cudaCheck.h:11-47).
 *
 * The engine is chosen when this file is compiled, as the reference chooses it
 * at link time (Makefile:31-50): FJMI_LINKED_ENGINE is one of the fjmi_engine
 * values.  The call runs on the null stream and returns once the kernel has
 * finished, so the caller's next hipMemcpy sees the result.  One library context
 * per process and device, created on first use.
 */
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>

#include "cluster.h"
#include "fjmi.h"
#include "fjmi_dropin.h"

#ifndef FJMI_LINKED_ENGINE
#define FJMI_LINKED_ENGINE FJMI_ENGINE_AUTO
#endif

namespace {

fjmi_context *g_ctx = nullptr;
int g_ctx_device = -1;

[[noreturn]] void die(const char *file, int line, const char *call, const char *msg) {
  std::fprintf(stderr, "\n%s, line %d:\n%s;\n%s\n\n", file, line, call, msg);
  std::abort();
}

#define FJMI_CHECK(call)                                             \
  do {                                                               \
    if ((call) != 0) die(__FILE__, __LINE__, #call, fjmi_last_error()); \
  } while (0)

/* The library context of the current device, created on first use. */
fjmi_context *context() {
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) device = 0;
  if (!g_ctx || g_ctx_device != device) {
    if (g_ctx) fjmi_destroy(g_ctx);
    g_ctx = nullptr;
    FJMI_CHECK(fjmi_create(device, &g_ctx));
    g_ctx_device = device;
  }
  return g_ctx;
}

}  // namespace

void inclusive_jets(const PseudoJet *particles, int size, double ptmin, PseudoJet *jets, int *count) {
  fjmi_context *ctx = context();
  FJMI_CHECK(fjmi_inclusive_pseudojets(ctx, particles, size, ptmin, jets, count, nullptr));
}

void cluster(PseudoJet *particles, int size, Algorithm algo, double r) {
  fjmi_context *ctx = context();
  FJMI_CHECK(fjmi_cluster_pseudojets(ctx, particles, size, static_cast<int>(algo), r, FJMI_LINKED_ENGINE, nullptr));
  /* waits for the kernel, as the reference's closing cudaFree calls do (tri_matrix.cu:332-335),
   * and aborts if it reported a failure */
  FJMI_CHECK(fjmi_stream_check(ctx, nullptr));
}
