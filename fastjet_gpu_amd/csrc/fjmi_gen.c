/*
 * fjmi_gen.c — seeded synthetic event generator for the benchmark and test
 * workloads (host C, no GPU).  Declared in include/fjmi_gen.h.
 *
 * RNG: xoshiro256** (Blackman & Vigna 2018) seeded through SplitMix64; uniform
 * variates U = ((x >> 11) + 1) * 2^-53 in (0, 1]; normals by Box-Muller;
 * exponentials by -mean * log(U).  No <random> distributions are used, so the
 * same seed gives the same bytes with any C library that rounds log/exp/sin/cos
 * the same way (the generated bytes are hashed by the callers to prove it).
 *
 * Particles are massless: px = pT cos(phi), py = pT sin(phi), pz = pT sinh(eta),
 * E = pT cosh(eta), phi wrapped into [0, 2pi).
 *
 * Workloads (one RNG stream per event: seed = SplitMix64(base ^ event), base =
 * 0x5EED0000 + workload):
 *  1  N = 354: 4 hard axes (eta~U[-2.5,2.5], phi~U[0,2pi)) x 25 particles with
 *     pT = 50 U^(-1/3) at dR = |N(0,0.15)| around the axis, plus 254 soft
 *     particles pT = 0.3 + Exp(1.0), eta~U[-5,5].
 *  2  N = 808: 6 axes x 30 hard + 628 soft, as above.
 *  3  N ~ U{50..2000}: U{2..8} axes of U{15..40} hard particles each (hard total
 *     truncated to N), the rest soft as above.
 *  4  N ~ U{5000..20000}: 10 axes x 40 hard + soft pT = 0.3 + Exp(0.8),
 *     eta~U[-4,4].
 *  5  N ~ lognormal(median 1500, sigma 0.6) capped to [2, 20000]; hard content
 *     as workload 3 (axes U{2..8} x U{15..40}), soft as workload 1.
 */
#include "../../include/fjmi_gen.h"

#include <math.h>
#include <stdint.h>

typedef struct {
  uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t next_u64(rng_t *r) {
  uint64_t *s = r->s;
  const uint64_t result = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}

static void rng_seed(rng_t *r, uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&x);
}

static double unif(rng_t *r) { return (double)((next_u64(r) >> 11) + 1) * 0x1p-53; }
static double unif_ab(rng_t *r, double a, double b) { return a + (b - a) * unif(r); }
static int unif_int(rng_t *r, int lo, int hi) { /* inclusive */
  return lo + (int)(next_u64(r) % (uint64_t)(hi - lo + 1));
}
static double normal(rng_t *r) {
  double u1 = unif(r), u2 = unif(r);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
static double expo(rng_t *r, double mean) { return -mean * log(unif(r)); }

static void emit(double *out, double pt, double eta, double phi) {
  const double twopi = 6.283185307179586;
  while (phi < 0) phi += twopi;
  while (phi >= twopi) phi -= twopi;
  out[0] = pt * cos(phi);
  out[1] = pt * sin(phi);
  out[2] = pt * sinh(eta);
  out[3] = pt * cosh(eta);
}

static void event_rng(rng_t *r, int workload, int64_t event) {
  uint64_t base = 0x5EED0000ULL + (uint64_t)workload;
  uint64_t x = base ^ (uint64_t)event;
  rng_seed(r, splitmix64(&x));
}

/* Draws the multiplicity (consumes the first variates of the stream). */
static int draw_n(rng_t *r, int workload) {
  switch (workload) {
    case 1: return 354;
    case 2: return 808;
    case 3: return unif_int(r, 50, 2000);
    case 4: return unif_int(r, 5000, 20000);
    case 5: {
      double n = exp(log(1500.0) + 0.6 * normal(r));
      if (n > 20000.0) n = 20000.0;
      if (n < 2.0) n = 2.0;
      return (int)n;
    }
    default: return -1;
  }
}

int fjmi_gen_event_size(int workload, int64_t event) {
  rng_t r;
  event_rng(&r, workload, event);
  return draw_n(&r, workload);
}

int fjmi_gen_event(int workload, int64_t event, double *p4, int capacity) {
  rng_t r;
  event_rng(&r, workload, event);
  const int n = draw_n(&r, workload);
  if (n < 0) return -1;
  if (n > capacity) return -2;
  int axes, per_axis_fixed = 0;
  double soft_mean = 1.0, soft_eta = 5.0;
  switch (workload) {
    case 1: axes = 4; per_axis_fixed = 25; break;
    case 2: axes = 6; per_axis_fixed = 30; break;
    case 4: axes = 10; per_axis_fixed = 40; soft_mean = 0.8; soft_eta = 4.0; break;
    default: axes = unif_int(&r, 2, 8); break;
  }
  int k = 0;
  for (int a = 0; a < axes && k < n; ++a) {
    const double aeta = unif_ab(&r, -2.5, 2.5);
    const double aphi = unif_ab(&r, 0.0, 6.283185307179586);
    const int m = per_axis_fixed ? per_axis_fixed : unif_int(&r, 15, 40);
    for (int j = 0; j < m && k < n; ++j) {
      const double pt = 50.0 * pow(unif(&r), -1.0 / 3.0);
      const double dr = fabs(0.15 * normal(&r));
      const double th = unif_ab(&r, 0.0, 6.283185307179586);
      emit(p4 + 4 * k, pt, aeta + dr * cos(th), aphi + dr * sin(th));
      ++k;
    }
  }
  for (; k < n; ++k) {
    const double pt = 0.3 + expo(&r, soft_mean);
    const double eta = unif_ab(&r, -soft_eta, soft_eta);
    const double phi = unif_ab(&r, 0.0, 6.283185307179586);
    emit(p4 + 4 * k, pt, eta, phi);
  }
  return n;
}

/* Serving-stream attributes of an event: algorithm/R choice in {anti-kt 0.4,
 * anti-kt 0.8, C/A 0.8, kt 0.4} and an exponential inter-arrival gap of mean 1
 * (scaled by 1/lambda by the caller).  Uses a stream independent of the
 * particles' stream. */
int fjmi_gen_serving_attrs(int64_t event, int *p, double *R, double *gap) {
  rng_t r;
  uint64_t x = 0x5E4E0000ULL ^ (uint64_t)event;
  rng_seed(&r, splitmix64(&x));
  int c = unif_int(&r, 0, 3);
  static const int ps[4] = {-1, -1, 0, 1};
  static const double Rs[4] = {0.4, 0.8, 0.8, 0.4};
  *p = ps[c];
  *R = Rs[c];
  *gap = expo(&r, 1.0);
  return 0;
}
