/*
 * fjmi_math.h — deterministic FP64 kinematics shared by the HIP kernels (device)
 * and by host code (C or C++), so that both sides produce bit-identical rapidity,
 * azimuth and clustering weights for the same four-momentum.
 *
 * Why a private atan2/log: the clustering decisions are compared bit for bit
 * between the GPU and the CPU.  ROCm's device libm and glibc are different
 * implementations and may differ in the last ulp, which could flip a near-tie.
 * Everything here uses only IEEE-754 basic operations (+ - * /) and explicit
 * fused multiply-add (fma), which are correctly rounded on x86-64 SSE2 and on
 * gfx950, so the results are identical wherever this header is compiled —
 * provided the compiler does not contract a*b+c on its own: every translation
 * unit that includes this header must be built with -ffp-contract=off.
 *
 * Accuracy: log and atan2 are table driven with double-double range reduction
 * (~2^-68 relative error before the final rounding), so the returned double is
 * the correctly rounded result except in rare cases (measured: about 1 in 10^5).  That makes them agree
 * with glibc's log/atan2 (which FastJet uses) on all but a tiny fraction of
 * inputs; the test-suite measures that fraction against high-precision decimal
 * values and against glibc.
 *
 * The kinematic conventions follow FastJet's PseudoJet (rapidity with the
 * effective-mass clamp, phi in [0, 2pi), MaxRap = 1e5 for zero-pt momenta) as
 * restated by the reference: tri_matrix.cu:56-100 (constants and _set_jet),
 * grid.cu:110-166 (weights per algorithm), tri_matrix.cu:102-113 (distance).
 */
#ifndef FJMI_MATH_H
#define FJMI_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define FJMI_HD __host__ __device__ static inline
#else
#define FJMI_HD static inline
#endif

/* Constants (FastJet conventions; tri_matrix.cu:56-58). */
#define FJMI_PI 0x1.921fb54442d18p+1
#define FJMI_TWOPI 0x1.921fb54442d18p+2
#define FJMI_MAXRAP 1e5

#if defined(__HIP_DEVICE_COMPILE__)
/* constexpr, not __constant__: a __constant__ variable stays visible to the
 * host runtime, so device code reaches it through the GOT (one more dependent
 * scalar load per table lookup); a constexpr table is a private read-only
 * global addressed relative to the program counter */
#define FJMI_TABLE static constexpr
#else
#define FJMI_TABLE static
#endif
#include "fjmi_math_tables.h"

typedef struct {
  double hi, lo;
} fjmi_dd;

FJMI_HD uint64_t fjmi_d2u(double x) {
  uint64_t u;
  __builtin_memcpy(&u, &x, sizeof u);
  return u;
}

FJMI_HD double fjmi_u2d(uint64_t u) {
  double x;
  __builtin_memcpy(&x, &u, sizeof x);
  return x;
}

/* ---- double-double building blocks (Dekker/Knuth error-free transforms) ---- */

FJMI_HD fjmi_dd fjmi_two_sum(double a, double b) {
  fjmi_dd r;
  r.hi = a + b;
  double bb = r.hi - a;
  r.lo = (a - (r.hi - bb)) + (b - bb);
  return r;
}

/* requires |a| >= |b| (or a == 0) */
FJMI_HD fjmi_dd fjmi_fast_two_sum(double a, double b) {
  fjmi_dd r;
  r.hi = a + b;
  r.lo = b - (r.hi - a);
  return r;
}

FJMI_HD fjmi_dd fjmi_two_prod(double a, double b) {
  fjmi_dd r;
  r.hi = a * b;
  r.lo = __builtin_fma(a, b, -r.hi);
  return r;
}

FJMI_HD fjmi_dd fjmi_dd_add(fjmi_dd a, fjmi_dd b) {
  fjmi_dd s = fjmi_two_sum(a.hi, b.hi);
  fjmi_dd t = fjmi_two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = fjmi_fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fjmi_fast_two_sum(s.hi, s.lo);
}

FJMI_HD fjmi_dd fjmi_dd_add_d(fjmi_dd a, double b) {
  fjmi_dd s = fjmi_two_sum(a.hi, b);
  s.lo += a.lo;
  return fjmi_fast_two_sum(s.hi, s.lo);
}

FJMI_HD fjmi_dd fjmi_dd_mul(fjmi_dd a, fjmi_dd b) {
  fjmi_dd p = fjmi_two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return fjmi_fast_two_sum(p.hi, p.lo);
}

FJMI_HD fjmi_dd fjmi_dd_mul_d(fjmi_dd a, double b) {
  fjmi_dd p = fjmi_two_prod(a.hi, b);
  p.lo += a.lo * b;
  return fjmi_fast_two_sum(p.hi, p.lo);
}

FJMI_HD fjmi_dd fjmi_dd_neg(fjmi_dd a) {
  a.hi = -a.hi;
  a.lo = -a.lo;
  return a;
}

FJMI_HD fjmi_dd fjmi_dd_make(double hi, double lo) {
  fjmi_dd r;
  r.hi = hi;
  r.lo = lo;
  return r;
}

/* ---- natural logarithm ----
 * x = 2^e m, m in (sqrt(1/2), sqrt(2)];  j = round(128 m), INV = RN(128/j);
 * r = m INV - 1 is formed exactly in double-double (two_prod, then a Sterbenz
 * subtraction), |r| < 2^-7.4;  log x = e ln2 - log(INV) + log1p(r), with
 * -log(INV) from the table and log1p(r) = r - r^2/2 + r^3 P(r): r and r^2/2 in
 * double-double, the degree-8 tail in double.  No division. */
/* log x for positive normal finite x = 2^e0 * (the argument): the table and
 * polynomial path, written without branches so that a caller's independent
 * work interleaves with it. */
FJMI_HD double fjmi_log_core(double x, int e0) {
  const uint64_t u = fjmi_d2u(x);
  int e = e0 + (int)(u >> 52) - 1023;
  double m = fjmi_u2d((u & 0x000FFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL);
  const int big = m > 0x1.6a09e667f3bcdp+0;
  m = big ? m * 0.5 : m;
  e = big ? e + 1 : e;
  const int j = (int)(m * 128.0 + 0.5); /* 91..181 */
  const double *tab = fjmi_log_tab[j - FJMI_LOG_J0];
  const fjmi_dd pr = fjmi_two_prod(m, tab[0]);
  const fjmi_dd r = fjmi_two_sum(pr.hi - 1.0, pr.lo); /* pr.hi - 1 is exact */
  const double rh = r.hi;
  /* P(r) = 1/3 - r/4 + r^2/5 - ... + r^8/11 */
  double P = __builtin_fma(rh, -1.0 / 11.0, 1.0 / 10.0);
  P = __builtin_fma(rh, -P, 1.0 / 9.0);
  P = __builtin_fma(rh, -P, 1.0 / 8.0);
  P = __builtin_fma(rh, -P, 1.0 / 7.0);
  P = __builtin_fma(rh, -P, 1.0 / 6.0);
  P = __builtin_fma(rh, -P, 1.0 / 5.0);
  P = __builtin_fma(rh, -P, 1.0 / 4.0);
  P = __builtin_fma(rh, -P, 1.0 / 3.0);
  fjmi_dd r2 = fjmi_two_prod(rh, rh);
  r2.lo += 2.0 * rh * r.lo;
  const double tail = (r2.hi * rh) * P;
  fjmi_dd L = fjmi_dd_add(r, fjmi_dd_make(-0.5 * r2.hi, -0.5 * r2.lo));
  L = fjmi_dd_add_d(L, tail);
  const fjmi_dd acc0 = fjmi_dd_make(tab[1], tab[2]);
  const fjmi_dd acc1 = fjmi_dd_add(fjmi_dd_make((double)e * FJMI_LN2_HI, (double)e * FJMI_LN2_LO), acc0);
  /* selects per component: a select of whole structs may be lowered through memory */
  const fjmi_dd acc = fjmi_dd_make(e != 0 ? acc1.hi : acc0.hi, e != 0 ? acc1.lo : acc0.lo);
  return fjmi_dd_add(acc, L).hi;
}

FJMI_HD double fjmi_log(double x) {
  if (!(x > 0.0)) {
    if (x == 0.0) return -__builtin_inf();
    return __builtin_nan(""); /* negative or NaN */
  }
  if (x == __builtin_inf()) return x;
  if (fjmi_d2u(x) < 0x0010000000000000ULL) return fjmi_log_core(x * 0x1p54, -54); /* subnormal */
  return fjmi_log_core(x, 0);
}

/* ---- atan2 ----
 * t = min(|x|,|y|)/max(|x|,|y|) in double-double; j = round(64 t),
 * atan(t) = atan(j/64) + atan(u), u = (t - j/64)/(1 + t j/64) in double-double,
 * |u| <= 2^-7, atan(u) = u + u^3 Q(u^2) with the degree-5 Q in double; then
 * octant/quadrant reconstruction in double-double.  IEEE special cases follow
 * C99 Annex F. */
FJMI_HD double fjmi_atan2_core(double y, double x);

FJMI_HD double fjmi_atan2(double y, double x) {
  if (x != x || y != y) return x + y;
  const int xneg = (int)(fjmi_d2u(x) >> 63);
  const int yneg = (int)(fjmi_d2u(y) >> 63);
  const double ax = __builtin_fabs(x), ay = __builtin_fabs(y);
  const double inf = __builtin_inf();
  if (ay == 0.0) {
    const double r = xneg ? FJMI_PI : 0.0;
    return yneg ? -r : r;
  }
  if (ax == 0.0) return yneg ? -0x1.921fb54442d18p+0 : 0x1.921fb54442d18p+0;
  if (ax == inf || ay == inf) {
    double r;
    if (ax == inf && ay == inf)
      r = xneg ? 0x1.2d97c7f3321d2p+1 /* 3pi/4 */ : 0x1.921fb54442d18p-1 /* pi/4 */;
    else if (ax == inf)
      r = xneg ? FJMI_PI : 0.0;
    else
      r = 0x1.921fb54442d18p+0;
    return yneg ? -r : r;
  }
  return fjmi_atan2_core(y, x);
}

/* atan2(y, x) for finite nonzero x and y: the table and polynomial path,
 * without branches (see fjmi_log_core). */
FJMI_HD double fjmi_atan2_core(double y, double x) {
  const int xneg = (int)(fjmi_d2u(x) >> 63);
  const int yneg = (int)(fjmi_d2u(y) >> 63);
  const double ax = __builtin_fabs(x), ay = __builtin_fabs(y);
  const int swap = ay > ax;
  const double num = swap ? ax : ay;
  const double den = swap ? ay : ax;
  const double th = num / den;
  const double tl = __builtin_fma(-th, den, num) / den;
  const int j = (int)(th * 64.0 + 0.5); /* 0..64 */
  const double c = (double)j * 0.015625;
  const fjmi_dd nu = fjmi_two_sum(th - c, tl); /* th - c is exact (Sterbenz) */
  const fjmi_dd du = fjmi_dd_add_d(fjmi_dd_mul_d(fjmi_dd_make(th, tl), c), 1.0);
  const double uh = nu.hi / du.hi;
  double rr = __builtin_fma(-uh, du.hi, nu.hi);
  rr = rr + nu.lo;
  rr = __builtin_fma(-uh, du.lo, rr);
  const fjmi_dd uu = fjmi_fast_two_sum(uh, rr / du.hi);
  const double w = uu.hi * uu.hi;
  /* Q(w) = -1/3 + w/5 - w^2/7 + w^3/9 - w^4/11 + w^5/13 */
  double Q = __builtin_fma(w, 1.0 / 13.0, -1.0 / 11.0);
  Q = __builtin_fma(w, Q, 1.0 / 9.0);
  Q = __builtin_fma(w, Q, -1.0 / 7.0);
  Q = __builtin_fma(w, Q, 1.0 / 5.0);
  Q = __builtin_fma(w, Q, -1.0 / 3.0);
  const fjmi_dd at = fjmi_dd_add_d(uu, (uu.hi * w) * Q);
  const fjmi_dd th1 = fjmi_dd_add(fjmi_dd_make(fjmi_atan_tab[j][0], fjmi_atan_tab[j][1]), at);
  /* selects per component: a select of whole structs may be lowered through memory */
  fjmi_dd th2 = fjmi_dd_make(j ? th1.hi : at.hi, j ? th1.lo : at.lo);
  const fjmi_dd th3 = fjmi_dd_add(fjmi_dd_make(0x1.921fb54442d18p+0, 0x1.1a62633145c07p-54), fjmi_dd_neg(th2));
  th2 = fjmi_dd_make(swap ? th3.hi : th2.hi, swap ? th3.lo : th2.lo);
  const fjmi_dd th4 = fjmi_dd_add(fjmi_dd_make(0x1.921fb54442d18p+1, 0x1.1a62633145c07p-53), fjmi_dd_neg(th2));
  const double hi = xneg ? th4.hi : th2.hi;
  return yneg ? -hi : hi;
}

/* ---- FastJet PseudoJet kinematics (tri_matrix.cu:61-100, grid.cu:112-166) ---- */

/* Azimuth in [0, 2pi); 0 for zero transverse momentum. */
FJMI_HD double fjmi_phi(double px, double py, double kt2) {
  double phi = (kt2 == 0.0) ? 0.0 : fjmi_atan2(py, px);
  if (phi < 0.0) phi += FJMI_TWOPI;
  if (phi >= FJMI_TWOPI) phi -= FJMI_TWOPI;
  return phi;
}

/* Rapidity with the effective-mass clamp; +-(MaxRap + |pz|) for E == |pz|, kt2 == 0. */
FJMI_HD double fjmi_rap(double pz, double E, double kt2) {
  const double apz = __builtin_fabs(pz);
  if (E == apz && kt2 == 0.0) {
    double maxraphere = FJMI_MAXRAP + apz;
    return (pz >= 0.0) ? maxraphere : -maxraphere;
  }
  double m2 = (E + pz) * (E - pz) - kt2;
  double effective_m2 = m2 > 0.0 ? m2 : 0.0;
  double E_plus_pz = E + apz;
  double rap = 0.5 * fjmi_log((kt2 + effective_m2) / (E_plus_pz * E_plus_pz));
  if (pz > 0) rap = -rap;
  return rap;
}

/* Generalised-kt weight kt^(2p) for p in {1, 0, -1} (cluster.h:4; tri_matrix.cu:40-53, :63). */
FJMI_HD double fjmi_weight(double kt2, int p) {
  if (p > 0) return kt2;
  if (p == 0) return 1.0;
  return kt2 > 1e-300 ? 1.0 / kt2 : 1e300;
}

/* Squared (rapidity, phi) distance with the phi wrap at pi (tri_matrix.cu:102-109). */
FJMI_HD double fjmi_dr2(double rap1, double phi1, double rap2, double phi2) {
  double dphi = __builtin_fabs(phi1 - phi2);
  if (dphi > FJMI_PI) dphi = FJMI_TWOPI - dphi;
  double drap = rap1 - rap2;
  return dphi * dphi + drap * drap;
}

#endif /* FJMI_MATH_H */
