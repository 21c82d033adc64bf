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

/* a + b for double-doubles whose high parts satisfy exponent(a.hi) >=
 * exponent(b.hi), or a.hi == 0: the high parts' sum is exact by Fast2Sum and
 * the low parts are added in double (relative error about 2^-104 when the sum
 * does not cancel); 8 operations against fjmi_dd_add's 20.  Every use below
 * states why its operands are ordered. */
FJMI_HD fjmi_dd fjmi_dd_add_ordered(fjmi_dd a, fjmi_dd b) {
  fjmi_dd s = fjmi_fast_two_sum(a.hi, b.hi);
  s.lo += a.lo + b.lo;
  return fjmi_fast_two_sum(s.hi, s.lo);
}

/* a + b for a double-double and a double with exponent(a.hi) >= exponent(b) */
FJMI_HD fjmi_dd fjmi_dd_add_d_ordered(fjmi_dd a, double b) {
  fjmi_dd s = fjmi_fast_two_sum(a.hi, b);
  s.lo += a.lo;
  return fjmi_fast_two_sum(s.hi, s.lo);
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
FJMI_HD fjmi_dd fjmi_log_core_dd(double x, int e0) {
  const uint64_t u = fjmi_d2u(x);
  int e = e0 + (int)(u >> 52) - 1023;
  double m = fjmi_u2d((u & 0x000FFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL);
  const int big = m > 0x1.6a09e667f3bcdp+0;
  m = big ? m * 0.5 : m;
  e = big ? e + 1 : e;
  const int j = (int)(m * 128.0 + 0.5); /* 91..181 */
  const double *tab = fjmi_log_tab[j - FJMI_LOG_J0];
  const fjmi_dd pr = fjmi_two_prod(m, tab[0]);
  /* pr.hi - 1 is exact, and a nonzero multiple of ulp(pr.hi) >= 2 |pr.lo| */
  const fjmi_dd r = fjmi_fast_two_sum(pr.hi - 1.0, pr.lo);
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
  /* |r| >= 2^8 r^2/2 and >> |tail|; with e != 0, |e ln2| >= 0.69 > 0.35 >= |log(INV)|;
   * and |log(INV)| >= 2^-7 > |L| (|r| <= 1/(2 j)) unless INV = 1, where it is 0 */
  fjmi_dd L = fjmi_dd_add_ordered(r, fjmi_dd_make(-0.5 * r2.hi, -0.5 * r2.lo));
  L = fjmi_dd_add_d_ordered(L, tail);
  const fjmi_dd acc0 = fjmi_dd_make(tab[1], tab[2]);
  const fjmi_dd acc1 = fjmi_dd_add_ordered(fjmi_dd_make((double)e * FJMI_LN2_HI, (double)e * FJMI_LN2_LO), acc0);
  /* selects per component: a select of whole structs may be lowered through memory */
  const fjmi_dd acc = fjmi_dd_make(e != 0 ? acc1.hi : acc0.hi, e != 0 ? acc1.lo : acc0.lo);
  return fjmi_dd_add_ordered(acc, L);
}

/* the rounded logarithm: the high part of the double-double */
FJMI_HD double fjmi_log_core(double x, int e0) { return fjmi_log_core_dd(x, e0).hi; }

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
  /* th - c is exact (Sterbenz), and 0 or a multiple of ulp(th) >= 2 |tl|; 1 >= t c */
  const fjmi_dd nu = fjmi_fast_two_sum(th - c, tl);
  const fjmi_dd du = fjmi_dd_add_ordered(fjmi_dd_make(1.0, 0.0), fjmi_dd_mul_d(fjmi_dd_make(th, tl), c));
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
  /* |u| >> |u^3 Q|; atan(j/64) >= atan(1/64) > 2^-7 >= |atan(u)| for j >= 1 (0 for j = 0);
   * pi/2 > th2 (<= pi/4) and pi > th2 (<= pi/2) */
  const fjmi_dd at = fjmi_dd_add_d_ordered(uu, (uu.hi * w) * Q);
  const fjmi_dd th1 = fjmi_dd_add_ordered(fjmi_dd_make(fjmi_atan_tab[j][0], fjmi_atan_tab[j][1]), at);
  /* selects per component: a select of whole structs may be lowered through memory */
  fjmi_dd th2 = fjmi_dd_make(j ? th1.hi : at.hi, j ? th1.lo : at.lo);
  const fjmi_dd th3 = fjmi_dd_add_ordered(fjmi_dd_make(0x1.921fb54442d18p+0, 0x1.1a62633145c07p-54), fjmi_dd_neg(th2));
  th2 = fjmi_dd_make(swap ? th3.hi : th2.hi, swap ? th3.lo : th2.lo);
  const fjmi_dd th4 = fjmi_dd_add_ordered(fjmi_dd_make(0x1.921fb54442d18p+1, 0x1.1a62633145c07p-53), fjmi_dd_neg(th2));
  const double hi = xneg ? th4.hi : th2.hi;
  return yneg ? -hi : hi;
}

/* ---- exp, pow, sin/cos, sqrt: the functions of general jet definitions ----
 * (the generalised-kt weight kt^(2p) for any real p, and the recombination
 * schemes whose merged jet is built from (pt, y, phi)).  Same rules as above:
 * basic operations and fma only, so host and device round alike. */

/* 2^k as a double for -1022 <= k <= 1023 */
FJMI_HD double fjmi_pow2i(int k) { return fjmi_u2d((uint64_t)(k + 1023) << 52); }

/* x * 2^k for a double-double x in [0.5, 4): the high part, scaled in two steps
 * below the normal range (rounded once more there: such results are subnormal) */
FJMI_HD double fjmi_scale_dd(fjmi_dd x, int k) {
  if (k > 1023) return x.hi * fjmi_pow2i(1023) * fjmi_pow2i(k - 1023);
  if (k >= -1021) return x.hi * fjmi_pow2i(k);
  return (x.hi * fjmi_pow2i(k + 64)) * fjmi_pow2i(-64);
}

/* exp(x) for a double-double x with |x.hi| < 746, as 2^k * (a double-double
 * in [1, 2) up to rounding):  n = round(64 x / ln 2), j = n mod 64, k = (n - j) / 64,
 * r = x - n ln(2)/64 in double-double (|r| <= ln(2)/128), exp(x) = 2^k 2^(j/64)
 * exp(r) with exp(r) - 1 = r + r^2/2 + r^3 P(r) (P to r^4/5040, the next term
 * below 2^-75 relative). */
FJMI_HD fjmi_dd fjmi_exp_core_dd(fjmi_dd x, int *k_out) {
  const double nf = __builtin_rint(x.hi * FJMI_INV_LN2_64);
  const int n = (int)nf;
  const int j = n & 63;
  *k_out = (n - j) / 64;
  fjmi_dd r = fjmi_dd_add_d(x, -nf * FJMI_LN2_64_HI); /* the product is exact */
  r = fjmi_dd_add_d(r, -nf * FJMI_LN2_64_LO);
  const double rh = r.hi;
  double P = __builtin_fma(rh, 1.0 / 5040.0, 1.0 / 720.0);
  P = __builtin_fma(rh, P, 1.0 / 120.0);
  P = __builtin_fma(rh, P, 1.0 / 24.0);
  P = __builtin_fma(rh, P, 1.0 / 6.0);
  fjmi_dd r2 = fjmi_two_prod(rh, rh);
  r2.lo += 2.0 * rh * r.lo;
  fjmi_dd E = fjmi_dd_add(r, fjmi_dd_make(0.5 * r2.hi, 0.5 * r2.lo));
  E = fjmi_dd_add_d(E, (r2.hi * rh) * P);
  const fjmi_dd T = fjmi_dd_make(fjmi_exp2_tab[j][0], fjmi_exp2_tab[j][1]);
  return fjmi_dd_add(T, fjmi_dd_mul(T, E));
}

/* e^x for a double-double argument: +inf above the overflow threshold, 0 below
 * the smallest subnormal's */
FJMI_HD double fjmi_exp_dd(fjmi_dd x) {
  if (x.hi != x.hi) return x.hi;
  if (x.hi > 709.782712893384) return __builtin_inf();
  if (x.hi < -745.1332191019412) return 0.0;
  int k;
  const fjmi_dd v = fjmi_exp_core_dd(x, &k);
  return fjmi_scale_dd(v, k);
}

FJMI_HD double fjmi_exp(double x) { return fjmi_exp_dd(fjmi_dd_make(x, 0.0)); }

/* x^p for finite x > 0 (normal or subnormal) and finite p: exp(p log x) with the
 * logarithm and the product in double-double */
FJMI_HD double fjmi_pow_pos(double x, double p) {
  const fjmi_dd L = fjmi_d2u(x) < 0x0010000000000000ULL ? fjmi_log_core_dd(x * 0x1p54, -54) : fjmi_log_core_dd(x, 0);
  return fjmi_exp_dd(fjmi_dd_mul_d(L, p));
}

/* sqrt(x), correctly rounded: x = m 4^e with m in [1, 4), Newton's iteration
 * y = (y + m/y)/2 from (1 + m)/2 (five steps reach 2^-60), then the neighbour of
 * y whose square is nearest m (the residual m - c^2 is exact in one fma). */
FJMI_HD double fjmi_sqrt(double x) {
  if (x != x || x < 0.0) return __builtin_nan("");
  if (x == 0.0 || x == __builtin_inf()) return x;
  int eb = 0;
  double xs = x;
  if (fjmi_d2u(xs) < 0x0010000000000000ULL) {
    xs *= 0x1p108;
    eb = -108;
  }
  const int ex = (int)(fjmi_d2u(xs) >> 52) - 1023 + eb;
  const int h = (ex >= 0 ? ex : ex - 1) / 2; /* floor(ex / 2) */
  const double m = fjmi_u2d((fjmi_d2u(xs) & 0x000FFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL) *
                   ((ex - 2 * h) ? 2.0 : 1.0); /* [1, 4) */
  double y = 0.5 * (1.0 + m);
  for (int i = 0; i < 5; ++i) y = 0.5 * (y + m / y);
  const double r0 = __builtin_fma(-y, y, m);
  const double up = fjmi_u2d(fjmi_d2u(y) + 1), dn = fjmi_u2d(fjmi_d2u(y) - 1);
  const double c = r0 > 0.0 ? up : dn;
  const double r1 = __builtin_fma(-c, c, m);
  y = (r0 != 0.0 && __builtin_fabs(r1) < __builtin_fabs(r0)) ? c : y;
  return y * fjmi_pow2i(h);
}

/* sin and cos of |x| < 16 pi: r = x - n pi/2 in double-double (|r| <= pi/4);
 * sin r = r + r^3 (-1/6 + z (1/120 + z S(z))) and cos r = 1 - z/2 + z^2 (1/24 +
 * z (-1/720 + z C(z))) with z = r^2, the first two coefficients of each series
 * in double-double and the Taylor tails S, C (to r^17 and r^18, the next terms
 * below 2^-62 relative) in double; then the quadrant of n. */
FJMI_HD void fjmi_sincos(double x, double *s_out, double *c_out) {
  const double nf = __builtin_rint(x * 0x1.45f306dc9c883p-1); /* 2/pi */
  const int n = (int)nf;
  const fjmi_dd r = fjmi_dd_add(fjmi_dd_make(x, 0.0), fjmi_dd_mul_d(fjmi_dd_make(FJMI_PIO2_HI, FJMI_PIO2_LO), -nf));
  const double rh = r.hi;
  fjmi_dd z = fjmi_two_prod(rh, rh);
  z.lo += 2.0 * rh * r.lo;
  z = fjmi_fast_two_sum(z.hi, z.lo);
  const double zh = z.hi;
  double S = __builtin_fma(zh, 1.0 / 355687428096000.0, -1.0 / 1307674368000.0);
  S = __builtin_fma(zh, S, 1.0 / 6227020800.0);
  S = __builtin_fma(zh, S, -1.0 / 39916800.0);
  S = __builtin_fma(zh, S, 1.0 / 362880.0);
  S = __builtin_fma(zh, S, -1.0 / 5040.0);
  fjmi_dd qs = fjmi_dd_add(fjmi_dd_make(0x1.1111111111111p-7, 0x1.1111111111111p-63), fjmi_dd_mul_d(z, S));
  qs = fjmi_dd_add(fjmi_dd_make(-0x1.5555555555555p-3, -0x1.5555555555555p-57), fjmi_dd_mul(z, qs));
  const fjmi_dd sn = fjmi_dd_add(r, fjmi_dd_mul(fjmi_dd_mul(z, r), qs));
  double C = __builtin_fma(zh, -1.0 / 6402373705728000.0, 1.0 / 20922789888000.0);
  C = __builtin_fma(zh, C, -1.0 / 87178291200.0);
  C = __builtin_fma(zh, C, 1.0 / 479001600.0);
  C = __builtin_fma(zh, C, -1.0 / 3628800.0);
  C = __builtin_fma(zh, C, 1.0 / 40320.0);
  fjmi_dd qc = fjmi_dd_add(fjmi_dd_make(-0x1.6c16c16c16c17p-10, 0x1.f49f49f49f49fp-65), fjmi_dd_mul_d(z, C));
  qc = fjmi_dd_add(fjmi_dd_make(0x1.5555555555555p-5, 0x1.5555555555555p-59), fjmi_dd_mul(z, qc));
  fjmi_dd cs = fjmi_dd_add(fjmi_dd_make(1.0, 0.0), fjmi_dd_make(-0.5 * z.hi, -0.5 * z.lo));
  cs = fjmi_dd_add(cs, fjmi_dd_mul(fjmi_dd_mul(z, z), qc));
  const double sv = sn.hi, cv = cs.hi;
  switch (n & 3) {
    case 0: *s_out = sv; *c_out = cv; break;
    case 1: *s_out = cv; *c_out = -sv; break;
    case 2: *s_out = -sv; *c_out = -cv; break;
    default: *s_out = -cv; *c_out = sv; break;
  }
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

/* Generalised-kt weight kt^(2p) for any real p (FastJet's genkt algorithm, its
 * ClusterSequence::jet_scale_for_algorithm: for p <= 0 a kt^2 below 1e-300 is
 * taken as 1e-300).  p = 1, 0 and -1 are the kt, Cambridge/Aachen and anti-kt
 * weights above, bit for bit; other p go through fjmi_pow_pos. */
FJMI_HD double fjmi_weight_gen(double kt2, double p) {
  if (p == 1.0) return kt2;
  if (p == 0.0) return 1.0;
  if (p == -1.0) return kt2 > 1e-300 ? 1.0 / kt2 : 1e300;
  double k = kt2;
  if (p <= 0.0 && k < 1e-300) k = 1e-300;
  if (k != k) return k;
  if (k == 0.0) return 0.0; /* p > 0 */
  if (k == __builtin_inf()) return p > 0.0 ? k : 0.0;
  return fjmi_pow_pos(k, p);
}

/* ---- Recombination schemes (FastJet's JetDefinition::DefaultRecombiner, its
 * RecombinationScheme numbering).  The reference has the E-scheme only
 * (tri_matrix.cu:247-250, grid.cu:453-456). ---- */
#define FJMI_SCHEME_E 0        /* four-momentum sum */
#define FJMI_SCHEME_PT 1       /* massless (pt, y, phi), pt-weighted y and phi; inputs made massless, E = |p| */
#define FJMI_SCHEME_PT2 2      /* as pt, pt^2-weighted */
#define FJMI_SCHEME_ET 3       /* as pt; inputs made massless by rescaling |p| to E */
#define FJMI_SCHEME_ET2 4      /* as Et, pt^2-weighted */
#define FJMI_SCHEME_BIPT 5     /* as pt, inputs unchanged */
#define FJMI_SCHEME_BIPT2 6    /* as pt2, inputs unchanged */
#define FJMI_SCHEME_WTA_PT 7   /* winner takes all: the harder jet's y, phi and mass, pt = pta + ptb */
#define FJMI_SCHEME_WTA_MODP 8 /* winner takes all: the harder jet's direction, |p| = |pa| + |pb| */
#define FJMI_NUM_SCHEMES 9

/* An input particle as the scheme takes it (DefaultRecombiner::preprocess). */
FJMI_HD void fjmi_preprocess(int scheme, double *px, double *py, double *pz, double *E) {
  if (scheme == FJMI_SCHEME_PT || scheme == FJMI_SCHEME_PT2) {
    *E = fjmi_sqrt((*px * *px + *py * *py) + *pz * *pz);
  } else if (scheme == FJMI_SCHEME_ET || scheme == FJMI_SCHEME_ET2) {
    const double rescale = *E / fjmi_sqrt((*px * *px + *py * *py) + *pz * *pz);
    *px *= rescale;
    *py *= rescale;
    *pz *= rescale;
  }
}

/* The momentum of (pt, y, phi, m) (PseudoJet::reset_PtYPhiM) into q[4]. */
FJMI_HD void fjmi_ptyphim(double pt, double y, double phi, double m, double *q) {
  const double ptm = (m == 0.0) ? pt : fjmi_sqrt(pt * pt + m * m);
  const double exprap = fjmi_exp(y);
  const double pminus = ptm / exprap;
  const double pplus = ptm * exprap;
  double s, c;
  fjmi_sincos(phi, &s, &c);
  q[0] = pt * c;
  q[1] = pt * s;
  q[2] = 0.5 * (pplus - pminus);
  q[3] = 0.5 * (pplus + pminus);
}

/* The merged jet of jets a and b (momenta qa, qb; rapidities and azimuths ya,
 * phia, yb, phib) into qo[4] (DefaultRecombiner::recombine).  Returns 1 when
 * the scheme fixes the merged jet's rapidity and azimuth (then in *y and *phi,
 * the azimuth in [0, 2 pi), as PseudoJet::set_cached_rap_phi keeps them), 0
 * when they follow from its momentum. */
FJMI_HD int fjmi_recombine(int scheme, const double *qa, double ya, double phia, const double *qb, double yb,
                           double phib, double *qo, double *y, double *phi) {
  if (scheme == FJMI_SCHEME_E) {
    for (int c = 0; c < 4; ++c) qo[c] = qa[c] + qb[c];
    return 0;
  }
  const double kta = qa[0] * qa[0] + qa[1] * qa[1], ktb = qb[0] * qb[0] + qb[1] * qb[1];
  if (scheme == FJMI_SCHEME_WTA_MODP) {
    const double ma2 = kta + qa[2] * qa[2], mb2 = ktb + qb[2] * qb[2];
    const int a_hard = ma2 >= mb2;
    const double *qh = a_hard ? qa : qb;
    const double mh2 = a_hard ? ma2 : mb2, ms2 = a_hard ? mb2 : ma2;
    if (mh2 == 0.0) {
      for (int c = 0; c < 4; ++c) qo[c] = 0.0;
      return 0;
    }
    const double modp_hard = fjmi_sqrt(mh2);
    const double modp_ab = modp_hard + fjmi_sqrt(ms2);
    const double scale = modp_ab / modp_hard;
    qo[0] = qh[0] * scale;
    qo[1] = qh[1] * scale;
    qo[2] = qh[2] * scale;
    qo[3] = modp_ab;
    return 0;
  }
  if (scheme == FJMI_SCHEME_WTA_PT) {
    const int a_hard = kta >= ktb;
    const double *qh = a_hard ? qa : qb;
    const double kth = a_hard ? kta : ktb;
    const double m2 = (qh[3] + qh[2]) * (qh[3] - qh[2]) - kth;
    const double m = m2 < 0.0 ? -fjmi_sqrt(-m2) : fjmi_sqrt(m2);
    *y = a_hard ? ya : yb;
    *phi = a_hard ? phia : phib;
    fjmi_ptyphim(fjmi_sqrt(kta) + fjmi_sqrt(ktb), *y, *phi, m, qo);
    return 1;
  }
  const int by_pt = scheme == FJMI_SCHEME_PT || scheme == FJMI_SCHEME_ET || scheme == FJMI_SCHEME_BIPT;
  const double pta = fjmi_sqrt(kta), ptb = fjmi_sqrt(ktb);
  const double wa = by_pt ? pta : kta, wb = by_pt ? ptb : ktb;
  const double perp_ab = pta + ptb;
  if (perp_ab == 0.0) {
    for (int c = 0; c < 4; ++c) qo[c] = 0.0;
    return 0;
  }
  const double y_ab = (wa * ya + wb * yb) / (wa + wb);
  double pa = phia, pb = phib;
  if (pa - pb > FJMI_PI) pb += FJMI_TWOPI;
  if (pa - pb < -FJMI_PI) pb -= FJMI_TWOPI;
  const double phi_ab = (wa * pa + wb * pb) / (wa + wb);
  fjmi_ptyphim(perp_ab, y_ab, phi_ab, 0.0, qo);
  double ph = phi_ab;
  if (ph >= FJMI_TWOPI) ph -= FJMI_TWOPI;
  if (ph < 0.0) ph += FJMI_TWOPI;
  *y = y_ab;
  *phi = ph;
  return 1;
}

/* Squared (rapidity, phi) distance with the phi wrap at pi (tri_matrix.cu:102-109). */
FJMI_HD double fjmi_dr2(double rap1, double phi1, double rap2, double phi2) {
  double dphi = __builtin_fabs(phi1 - phi2);
  if (dphi > FJMI_PI) dphi = FJMI_TWOPI - dphi;
  double drap = rap1 - rap2;
  return dphi * dphi + drap * drap;
}

#endif /* FJMI_MATH_H */
