/*
 * event_file.cpp — the binary event-batch file of include/fjmi.h
 * (fjmi_events_write / _info / _read): many events' four-momenta and offsets,
 * with optional per-event algorithm exponent and radius, for throughput runs
 * and serving replays.  Not in the reference, which reads text only (one
 * "px py pz E" line per particle, events separated by '#' or blank lines,
 * main.cc:34-85).
 *
 * Layout, little-endian, no padding:
 *   bytes  0..7   magic "FJMIEVT1"
 *          8..15  uint64 number of events E
 *         16..23  uint64 number of particles P
 *         24..27  uint32 flags: bit 0 per-event p present, bit 1 per-event R present
 *         28..31  uint32 zero
 *   then  int64 offsets[E + 1] (offsets[0] = 0, offsets[E] = P, non-decreasing),
 *         double p4[4 P] (px, py, pz, E per particle),
 *         int32 p[E] if flag 0, then double R[E] if flag 1.
 * The reader checks the magic, the flags, the offsets and the file size.
 */
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fjmi.h"
#include "fjmi_runtime.h"

using fjmi::rt::fail;

namespace {

constexpr char kMagic[8] = {'F', 'J', 'M', 'I', 'E', 'V', 'T', '1'};

struct Header {
  uint64_t n_events, n_particles;
  uint32_t flags, zero;
};

struct File {
  std::FILE *f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

int read_header(std::FILE *f, const char *path, Header *h) {
  char magic[8];
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kMagic, 8) != 0)
    return fail(FJMI_EINVAL, "%s: not an event file (bad magic)", path);
  if (std::fread(h, sizeof *h, 1, f) != 1) return fail(FJMI_EINVAL, "%s: truncated header", path);
  if ((h->flags & ~3u) != 0 || h->zero != 0) return fail(FJMI_EINVAL, "%s: unknown flags", path);
  if (h->n_events > (uint64_t)INT32_MAX - 1) return fail(FJMI_EINVAL, "%s: too many events", path);
  if (h->n_particles > ((uint64_t)1 << 40)) return fail(FJMI_EINVAL, "%s: too many particles", path);
  /* the file's size must match the header */
  const long here = std::ftell(f);
  if (std::fseek(f, 0, SEEK_END) != 0) return fail(FJMI_EINVAL, "%s: not seekable", path);
  const uint64_t size = (uint64_t)std::ftell(f);
  std::fseek(f, here, SEEK_SET);
  const uint64_t want = 32 + 8 * (h->n_events + 1) + 32 * h->n_particles + ((h->flags & 1) ? 4 * h->n_events : 0) +
                        ((h->flags & 2) ? 8 * h->n_events : 0);
  if (size != want)
    return fail(FJMI_EINVAL, "%s: %llu bytes where the header implies %llu", path, (unsigned long long)size,
                (unsigned long long)want);
  return FJMI_OK;
}

}  // namespace

extern "C" {

int fjmi_events_write(const char *path, int n_events, const int64_t *offsets, const double *p4, const int32_t *p,
                      const double *R) {
  if (!path || n_events < 0 || !offsets) return fail(FJMI_EINVAL, "bad argument");
  if (offsets[0] != 0) return fail(FJMI_EINVAL, "offsets[0] must be 0");
  for (int e = 0; e < n_events; ++e)
    if (offsets[e + 1] < offsets[e]) return fail(FJMI_EINVAL, "offsets not monotonic at event %d", e);
  const uint64_t P = (uint64_t)offsets[n_events];
  if (P > 0 && !p4) return fail(FJMI_EINVAL, "null four-momenta");
  const std::string tmp = std::string(path) + ".tmp";
  File out;
  out.f = std::fopen(tmp.c_str(), "wb");
  if (!out.f) return fail(FJMI_EINVAL, "%s: cannot open for writing", tmp.c_str());
  const Header h{(uint64_t)n_events, P, (p ? 1u : 0u) | (R ? 2u : 0u), 0u};
  bool ok = std::fwrite(kMagic, 1, 8, out.f) == 8 && std::fwrite(&h, sizeof h, 1, out.f) == 1 &&
            std::fwrite(offsets, 8, (size_t)n_events + 1, out.f) == (size_t)n_events + 1 &&
            std::fwrite(p4, 32, (size_t)P, out.f) == (size_t)P;
  if (ok && p) ok = std::fwrite(p, 4, (size_t)n_events, out.f) == (size_t)n_events;
  if (ok && R) ok = std::fwrite(R, 8, (size_t)n_events, out.f) == (size_t)n_events;
  ok = (std::fclose(out.f) == 0) && ok;
  out.f = nullptr;
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return fail(FJMI_EINVAL, "%s: write failed", path);
  }
  return FJMI_OK;
}

int fjmi_events_info(const char *path, int *n_events, int64_t *n_particles, int *has_p, int *has_R) {
  if (!path || !n_events || !n_particles) return fail(FJMI_EINVAL, "bad argument");
  File in;
  in.f = std::fopen(path, "rb");
  if (!in.f) return fail(FJMI_EINVAL, "%s: cannot open", path);
  Header h;
  int rc = read_header(in.f, path, &h);
  if (rc) return rc;
  *n_events = (int)h.n_events;
  *n_particles = (int64_t)h.n_particles;
  if (has_p) *has_p = (h.flags & 1) ? 1 : 0;
  if (has_R) *has_R = (h.flags & 2) ? 1 : 0;
  return FJMI_OK;
}

int fjmi_events_read(const char *path, int64_t *offsets, double *p4, int32_t *p, double *R) {
  if (!path || !offsets) return fail(FJMI_EINVAL, "bad argument");
  File in;
  in.f = std::fopen(path, "rb");
  if (!in.f) return fail(FJMI_EINVAL, "%s: cannot open", path);
  Header h;
  int rc = read_header(in.f, path, &h);
  if (rc) return rc;
  const size_t E = (size_t)h.n_events, P = (size_t)h.n_particles;
  if (P > 0 && !p4) return fail(FJMI_EINVAL, "null four-momenta buffer");
  if (std::fread(offsets, 8, E + 1, in.f) != E + 1) return fail(FJMI_EINVAL, "%s: truncated offsets", path);
  if (offsets[0] != 0 || (uint64_t)offsets[E] != h.n_particles) return fail(FJMI_EINVAL, "%s: bad offsets", path);
  for (size_t e = 0; e < E; ++e)
    if (offsets[e + 1] < offsets[e]) return fail(FJMI_EINVAL, "%s: offsets not monotonic at event %zu", path, e);
  if (std::fread(p4, 32, P, in.f) != P) return fail(FJMI_EINVAL, "%s: truncated four-momenta", path);
  std::vector<int32_t> pbuf;
  if (h.flags & 1) {
    int32_t *dst = p;
    if (!dst) {
      pbuf.resize(E);
      dst = pbuf.data();
    }
    if (std::fread(dst, 4, E, in.f) != E) return fail(FJMI_EINVAL, "%s: truncated exponents", path);
    for (size_t e = 0; e < E; ++e)
      if (dst[e] < -1 || dst[e] > 1) return fail(FJMI_EINVAL, "%s: exponent %d of event %zu", path, dst[e], e);
  }
  if (h.flags & 2) {
    std::vector<double> rbuf;
    double *dst = R;
    if (!dst) {
      rbuf.resize(E);
      dst = rbuf.data();
    }
    if (std::fread(dst, 8, E, in.f) != E) return fail(FJMI_EINVAL, "%s: truncated radii", path);
    for (size_t e = 0; e < E; ++e)
      if (!(dst[e] > 0.0)) return fail(FJMI_EINVAL, "%s: radius of event %zu is not positive", path, e);
  }
  return FJMI_OK;
}

}  // extern "C"
