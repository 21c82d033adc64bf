// Runs the tiled kernel's source (fastjet_gpu_amd/csrc/tiled.hip) on the CPU as
// one emulated wavefront and checks every result bit-exactly against the oracle
// (N2Plain, or its tiled form above 3000 particles).  Test infrastructure only.
//   tiled_emu WORKLOAD E0 E1 [VARIANT P R TILE_Q NMAX POOL_EXTRA POOL_PER_N]   generated events
//   tiled_emu file PATH VARIANT P R                       raw little-endian px,py,pz,E doubles
//   tiled_emu qbound                                      queue indices outside [0, n_queue) leave at once
//   tiled_emu gen WORKLOAD E0 E1 VARIANT P SCHEME R       general jet definition (real p, scheme)
// FJMI_EMU_HEADLESS=1 launches without a queue head (the one-event form).
// FJMI_EMU_COOP=1 runs tiled_coop_kernel (8 wavefronts per event) instead.
// The tile pool of an event holds POOL_PER_N * N + POOL_EXTRA entries (default
// 3 N + kTiledPoolPerEvent, as the library allocates); a small pool, with less
// segment slack (FJMI_EMU_SLACK_SHIFT=30 FJMI_EMU_SLACK_ADD=0), makes segments
// relocate until the pool is exhausted and exercises the kernel's segment
// rebuild (counted and printed).  FJMI_EMU_STAMP_LIMIT=n restarts the gather
// counter every n gathers.
static long g_rebuilds = 0;
#define FJMI_TILED_ON_REBUILD() (++g_rebuilds)
/* The kernel's gather-counter limit and segment slack, set at run time here
 * (FJMI_EMU_STAMP_LIMIT, FJMI_EMU_SLACK_SHIFT, FJMI_EMU_SLACK_ADD) so that one
 * build serves every test; the library's defaults otherwise. */
namespace emu_knobs {
static unsigned stamp_limit = 0xFFFFFF00u;
static int slack_shift = 1, slack_add = 2;
}  // namespace emu_knobs
#define FJMI_TILED_STAMP_LIMIT (emu_knobs::stamp_limit)
#define FJMI_TILED_SLACK_SHIFT (emu_knobs::slack_shift)
#define FJMI_TILED_SLACK_ADD (emu_knobs::slack_add)
#include "tiled.hip"

#include <vector>

#include "fj_oracle.h"
#include "fjmi_gen.h"

#include <cmath>
#include <cstdint>
#include <functional>
namespace emu {
void run_wave(const std::function<void()> &fn);
void run_block(int nwaves, const std::function<void()> &fn);
}

static int g_pool_extra = fjmi::kTiledPoolPerEvent, g_pool_per_n = 3;
static bool g_headless = false;
static bool g_coop = false; /* FJMI_EMU_COOP=1: tiled_coop_kernel, 8 waves per event (variants 0 and 1) */

/* def: a general jet definition (tiled_gen_kernel, checked against the
 * oracle's *_def restatements), or null for p in {1, 0, -1} with the E-scheme */
template <int MAXT, int WPE>
static int run_event(const double *p4, int N, int p, double R, int tile_q, bool verbose,
                     const fjmi::JetDef *def = nullptr) {
  using namespace fjmi;
  std::vector<JetRec> rec(N > 0 ? N : 1);
  std::vector<double> p4s(4 * (size_t)(N > 0 ? N : 1));
  std::vector<int> hist(N > 0 ? N : 1);
  std::vector<unsigned short> lab2(N > 0 ? N : 1), pool((size_t)g_pool_per_n * N + g_pool_extra);
  TiledScratch ts{rec.data(), p4s.data(), hist.data(), lab2.data(), pool.data(), (g_pool_per_n - 3) * N + g_pool_extra};
  std::vector<int32_t> h1(N > 0 ? N : 1, -7), h2(N > 0 ? N : 1, -7);
  std::vector<double> hd(N > 0 ? N : 1, -7.0), jets(4 * (size_t)(N > 0 ? N : 1), 0.0);
  int32_t nj = -1;
  unsigned head = 0, status = 0;
  int64_t offs[2] = {0, N};
  LaunchArgs a;
  std::memset(&a, 0, sizeof a);
  a.in = p4;
  a.in_stride = 4;
  a.in_off = 0;
  a.offsets = offs;
  a.n_single = N;
  a.queue = nullptr;
  a.n_queue = 1;
  a.head = g_headless ? nullptr : &head;
  a.head_base = 0;
  a.p_uni = p;
  a.R_uni = R;
  a.hist_p1 = h1.data();
  a.hist_p2 = h2.data();
  a.hist_dij = hd.data();
  a.jets = jets.data();
  a.njets = &nj;
  a.status = &status;
  a.stage_chunk = kStageChunk;
  a.tile_q = tile_q;
  if (def) {
    a.def_uni = *def;
    emu::run_wave([&] { tiled_gen_kernel<MAXT, WPE>(a, ts); });
  } else if (g_coop) {
    emu::run_block(8, [&] { tiled_coop_kernel<MAXT == 512 ? 512 : 384, 8>(a, ts); });
  } else {
    emu::run_wave([&] { tiled_kernel<MAXT, WPE, false>(a, ts); });
  }
  std::vector<int32_t> w1(N > 0 ? N : 1), w2(N > 0 ? N : 1);
  std::vector<double> wd(N > 0 ? N : 1), wj(4 * (size_t)(N > 0 ? N : 1));
  int32_t wn = 0;
  if (def && N <= 3000)
    fjo_cluster_n2plain_def(p4, N, def->p, def->scheme, R, 0, w1.data(), w2.data(), wd.data(), wj.data(), &wn);
  else if (def)
    fjo_cluster_tiled_def(p4, N, def->p, def->scheme, R, 0, w1.data(), w2.data(), wd.data(), wj.data(), &wn);
  else if (N <= 3000)
    fjo_cluster_n2plain(p4, N, p, R, 0, w1.data(), w2.data(), wd.data(), wj.data(), &wn);
  else
    fjo_cluster_tiled(p4, N, p, R, 0, w1.data(), w2.data(), wd.data(), wj.data(), &wn);
  int bad = status != 0 || nj != wn;
  int first = -1;
  for (int s = 0; s < N && !bad; ++s)
    if (h1[s] != w1[s] || h2[s] != w2[s] || std::memcmp(&hd[s], &wd[s], 8) != 0) {
      bad = 1;
      first = s;
    }
  if (!bad && std::memcmp(jets.data(), wj.data(), 32 * (size_t)wn) != 0) bad = 1;
  if (bad || verbose)
    std::printf("N=%d p=%g scheme=%d R=%g status=%u nj=%d/%d first_diff_step=%d %s\n", N, def ? def->p : (double)p,
                def ? def->scheme : 0, R, status, nj, wn, first, bad ? "MISMATCH" : "ok");
  return bad;
}

static int run_file(const char *path, int variant, int p, double R) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return 2;
  std::vector<double> p4;
  double x;
  while (std::fread(&x, 8, 1, f) == 1) p4.push_back(x);
  std::fclose(f);
  const int n = (int)(p4.size() / 4);
  switch (variant) {
    case 0: return run_event<512, 2>(p4.data(), n, p, R, 0, true);
    case 1: return run_event<384, 3>(p4.data(), n, p, R, 0, true);
    case 2: return run_event<320, 3>(p4.data(), n, p, R, 0, true);
    default: return run_event<256, 3>(p4.data(), n, p, R, 0, true);
  }
}

/* The kernel's kinematics (its straight-line common path and the general
 * one) against the shared scalar functions, bit for bit, on edge cases and on
 * random momenta spread over many decades. */
static int run_kin(long n, unsigned seed) {
  const double edge[] = {0.0, -0.0, 1e-320, -1e-320, 1e-200, -1e-200, 1e-160, 1e-30, 0.5, 1.0, -1.0, 3.0, 1e30,
                         1e160, -1e160, 1e300, -1e300, __builtin_inf(), -__builtin_inf(), __builtin_nan("")};
  const int ne = (int)(sizeof edge / sizeof edge[0]);
  uint64_t x = 0x9E3779B97F4A7C15ull ^ seed;
  auto rnd = [&]() {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return x;
  };
  auto pick = [&]() -> double {
    const uint64_t r = rnd();
    if ((r & 7) == 0) return edge[(r >> 3) % ne];
    const double mant = (double)((r >> 11) & 0xFFFFFFFFFFFFFull) * 0x1p-52 + 1.0; /* [1, 2) */
    const int ex = (int)((r >> 3) % 160) - 80;
    return ((r >> 63) ? -1.0 : 1.0) * std::ldexp(mant, ex);
  };
  long bad = 0;
  for (long i = 0; i < n; ++i) {
    double px = pick(), py = pick(), pz = pick(), E = pick();
    if (rnd() % 4 == 0) E = std::fabs(pz); /* E == |pz| */
    if (rnd() % 4 == 0) E = std::sqrt(px * px + py * py + pz * pz); /* massless */
    for (int p = -1; p <= 1; ++p) {
      const double4 c = fjmi::tiled::kin_coords_t(px, py, pz, E, p);
      const double kt2 = px * px + py * py;
      const double r = fjmi_rap(pz, E, kt2), f = fjmi_phi(px, py, kt2), w = fjmi_weight(kt2, p);
      if (std::memcmp(&c.x, &r, 8) != 0 || std::memcmp(&c.y, &f, 8) != 0 || std::memcmp(&c.z, &w, 8) != 0) {
        if (bad < 5)
          std::printf("kinematics differ: p4=(%a %a %a %a) p=%d: (%a %a %a) vs (%a %a %a)\n", px, py, pz, E, p, c.x, c.y,
                      c.z, r, f, w);
        ++bad;
      }
    }
  }
  std::printf("kinematics: %ld of %ld differ\n", bad, 3 * n);
  return bad ? 1 : 0;
}

/* Launches whose queue head yields an index below 0 or at/after n_queue must
 * leave without touching any output. */
static int run_qbound() {
  using namespace fjmi;
  const double p4[8] = {1.0, 0.0, 0.0, 1.0, 1.0, 0.01, 0.0, 1.0001};
  std::vector<JetRec> rec(2);
  std::vector<double> p4s(8);
  std::vector<int> hist(2);
  std::vector<unsigned short> lab2(2), pool(6 + kTiledPoolPerEvent);
  TiledScratch ts{rec.data(), p4s.data(), hist.data(), lab2.data(), pool.data(), kTiledPoolPerEvent};
  int fails = 0;
  const unsigned cases[][2] = {{0u, 3u}, {5u, 0u}, {0x80000000u, 0u}};
  for (const auto &c : cases) {
    int32_t h1[2] = {-7, -7}, h2[2] = {-7, -7}, nj = -1;
    double hd[2] = {-7.0, -7.0}, jets[8] = {0};
    unsigned head = c[0], status = 0;
    int64_t offs[2] = {0, 2};
    LaunchArgs a;
    std::memset(&a, 0, sizeof a);
    a.in = p4;
    a.in_stride = 4;
    a.offsets = offs;
    a.n_single = 2;
    a.n_queue = 1;
    a.head = &head;
    a.head_base = c[1];
    a.p_uni = -1;
    a.R_uni = 0.4;
    a.hist_p1 = h1;
    a.hist_p2 = h2;
    a.hist_dij = hd;
    a.jets = jets;
    a.njets = &nj;
    a.status = &status;
    a.stage_chunk = kStageChunk;
    emu::run_wave([&] { tiled_kernel<256, 3, false>(a, ts); });
    const bool ok = nj == -1 && h1[0] == -7 && h1[1] == -7 && h2[0] == -7 && status == 0;
    std::printf("qbound head=%u base=%u: %s\n", c[0], c[1], ok ? "ok" : "TOUCHED");
    fails += !ok;
  }
  std::printf("done: %d mismatches\n", fails);
  return fails ? 1 : 0;
}

/* Generated events under a general jet definition: tiled_gen_kernel against
 * the oracle's *_def restatements. */
static int run_gen(int workload, long e0, long e1, int variant, double pd, int scheme, double R) {
  const fjmi::JetDef def{pd, scheme, 0};
  int fails = 0;
  for (long e = e0; e < e1; ++e) {
    const int n = fjmi_gen_event_size(workload, e);
    std::vector<double> p4(4 * (size_t)(n > 0 ? n : 1));
    fjmi_gen_event(workload, e, p4.data(), n);
    std::printf("[event %ld N=%d]\n", e, n);
    std::fflush(stdout);
    switch (variant) {
      case 0: fails += run_event<512, 2>(p4.data(), n, 0, R, 0, true, &def); break;
      case 1: fails += run_event<384, 3>(p4.data(), n, 0, R, 0, true, &def); break;
      case 2: fails += run_event<320, 3>(p4.data(), n, 0, R, 0, true, &def); break;
      default: fails += run_event<256, 3>(p4.data(), n, 0, R, 0, true, &def); break;
    }
  }
  std::printf("done: %d mismatches\n", fails);
  return fails ? 1 : 0;
}

int main(int argc, char **argv) {
  if (const char *v = std::getenv("FJMI_EMU_STAMP_LIMIT")) emu_knobs::stamp_limit = (unsigned)std::atol(v);
  if (const char *v = std::getenv("FJMI_EMU_SLACK_SHIFT")) emu_knobs::slack_shift = std::atoi(v);
  if (const char *v = std::getenv("FJMI_EMU_SLACK_ADD")) emu_knobs::slack_add = std::atoi(v);
  if (const char *h = std::getenv("FJMI_EMU_HEADLESS")) g_headless = std::atoi(h) != 0;
  if (const char *c = std::getenv("FJMI_EMU_COOP")) g_coop = std::atoi(c) != 0;
  if (argc >= 9 && std::strcmp(argv[1], "gen") == 0)
    return run_gen(std::atoi(argv[2]), std::atol(argv[3]), std::atol(argv[4]), std::atoi(argv[5]), std::atof(argv[6]),
                   std::atoi(argv[7]), std::atof(argv[8]));
  if (argc >= 2 && std::strcmp(argv[1], "qbound") == 0) return run_qbound();
  if (argc >= 2 && std::strcmp(argv[1], "kin") == 0) return run_kin(std::atol(argv[2]), (unsigned)std::atol(argv[3]));
  if (argc >= 2 && std::strcmp(argv[1], "file") == 0)
    return run_file(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), std::atof(argv[5]));
  if (argc < 4) {
    std::fprintf(stderr, "usage: emu workload e0 e1 [variant p R tile_q]\n");
    return 2;
  }
  const int workload = std::atoi(argv[1]);
  const long e0 = std::atol(argv[2]), e1 = std::atol(argv[3]);
  const int variant = argc > 4 ? std::atoi(argv[4]) : 0;
  const int p = argc > 5 ? std::atoi(argv[5]) : -1;
  const double R = argc > 6 ? std::atof(argv[6]) : 0.4;
  const int tq = argc > 7 ? std::atoi(argv[7]) : 0;
  const int nmax = argc > 8 ? std::atoi(argv[8]) : 1 << 30;
  if (argc > 9) g_pool_extra = std::atoi(argv[9]);
  if (argc > 10) g_pool_per_n = std::atoi(argv[10]);
  int fails = 0;
  for (long e = e0; e < e1; ++e) {
    const int n = fjmi_gen_event_size(workload, e);
    if (n > nmax) continue;
    std::vector<double> p4(4 * (size_t)(n > 0 ? n : 1));
    fjmi_gen_event(workload, e, p4.data(), n);
    std::printf("[event %ld N=%d]\n", e, n);
    std::fflush(stdout);
    int b = 0;
    switch (variant) {
      case 0: b = run_event<512, 2>(p4.data(), n, p, R, tq, false); break;
      case 1: b = run_event<384, 3>(p4.data(), n, p, R, tq, false); break;
      case 2: b = run_event<320, 3>(p4.data(), n, p, R, tq, false); break;
      default: b = run_event<256, 3>(p4.data(), n, p, R, tq, false); break;
    }
    fails += b;
    std::fflush(stdout);
  }
  std::printf("rebuilds: %ld\n", g_rebuilds);
  std::printf("done: %d mismatches\n", fails);
  return fails ? 1 : 0;
}
