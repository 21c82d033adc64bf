"""Pins the functions of general jet definitions in the shared kinematics header
(fastjet_gpu_amd/csrc/fjmi_math.h), which the HIP kernels and the oracle both use:
exp, pow (the generalised-kt weight kt^(2p) for real p), sqrt, sin and cos
(the recombination schemes that rebuild a jet from pt, y and phi, FastJet's
PseudoJet::reset_PtYPhiM).

- against values from 60-digit decimal arithmetic (independent of every C
  library): sqrt exactly; exp, pow, sin and cos correctly rounded on at least
  99.9 % of the samples (measured: all of them for exp and pow, 99.99 % for sin
  and cos) and never off by more than one ulp;
- their agreement rate with glibc, which FastJet calls, is measured and bounded
  (measured 99.85-99.9 %: glibc's exp, sin and cos are not correctly rounded on
  every input either);
- closed forms: exp(0) = 1, 2^k, x^(1/2) = sqrt(x), x^1 = x, x^2 = x * x,
  sin^2 + cos^2 = 1, symmetries, overflow and underflow;
- the genkt weight reduces to the kt, C/A and anti-kt weights at p = 1, 0, -1
  and applies FastJet's kt^2 >= 1e-300 guard for p <= 0.
"""
import math
import random
from decimal import Decimal as D, getcontext

import pytest

from oracle import oracle as orc

getcontext().prec = 60


def ulp(x):
    return math.ulp(x)


def dec_pi():
    def arctan_inv(k):
        x = D(1) / k
        s, t, n, x2 = D(0), x, 0, x * x
        while True:
            term = t / (2 * n + 1)
            if abs(term) < D(10) ** -58:
                return s
            s += term if n % 2 == 0 else -term
            t *= x2
            n += 1
    return 4 * (4 * arctan_inv(5) - arctan_inv(239))


PI = dec_pi()


def dec_sincos(x):
    """sin and cos of a double to 60 digits (reduction by pi/2 in decimal, then Taylor series)."""
    x = D(x)
    n = int((x / (PI / 2)).to_integral_value())
    r = x - n * (PI / 2)
    s, c, t, k = D(0), D(0), D(1), 0
    while True:
        if k % 2 == 0:
            term = t * (1 if (k // 2) % 2 == 0 else -1)
            c += term
        else:
            term = t * (1 if (k // 2) % 2 == 0 else -1)
            s += term
        k += 1
        t = t * r / k
        if abs(t) < D(10) ** -58:
            break
    return [(s, c), (c, -s), (-s, -c), (-c, s)][n % 4]


def err_ulps(got, want):
    return abs(D(got) - want) / D(ulp(float(want)))


def test_sqrt_correctly_rounded():
    """sqrt is IEEE's correctly rounded square root: equal to the platform's on every input."""
    L = orc.lib()
    rnd = random.Random(5)
    xs = [rnd.uniform(0, 10) for _ in range(20000)] + [math.exp(rnd.uniform(-740, 709)) for _ in range(20000)]
    xs += [float(k * k) for k in range(1, 3000)] + [2.0 ** k for k in range(-1074, 1024, 7)]
    xs += [5e-324, 2.2250738585072014e-308, 2.225073858507201e-308, 1.7976931348623157e308, 1.0, 4.0, 0.25]
    xs += [math.nextafter(1.0, 2.0), math.nextafter(4.0, 0.0), math.nextafter(2.0, 0.0)]
    bad = [x for x in xs if L.fjo_sqrt(x) != math.sqrt(x)]
    assert not bad, bad[:5]
    assert L.fjo_sqrt(0.0) == 0.0 and math.copysign(1, L.fjo_sqrt(-0.0)) == -1
    assert math.isnan(L.fjo_sqrt(-1.0)) and L.fjo_sqrt(math.inf) == math.inf


def test_exp_against_decimal_and_glibc():
    L = orc.lib()
    rnd = random.Random(7)
    xs = [rnd.uniform(-5, 5) for _ in range(4000)] + [rnd.uniform(-708, 709) for _ in range(4000)]
    xs += [rnd.uniform(-1e-8, 1e-8) for _ in range(500)] + [rnd.uniform(-745, -708) for _ in range(500)]
    cr = agree = 0
    worst = D(0)
    for x in xs:
        got = L.fjo_exp(x)
        want = D(x).exp()
        if got == float(want):
            cr += 1
        if got == math.exp(x):
            agree += 1
        if float(want) >= 2.2250738585072014e-308:
            worst = max(worst, err_ulps(got, want))
    assert worst <= 1, worst
    assert cr >= 0.999 * len(xs), cr / len(xs)
    assert agree >= 0.995 * len(xs), agree / len(xs)


def test_exp_closed_forms_and_limits():
    L = orc.lib()
    assert L.fjo_exp(0.0) == 1.0 and L.fjo_exp(-0.0) == 1.0
    xs = sorted(random.Random(29).uniform(-745, 709) for _ in range(20000))
    ys = [L.fjo_exp(x) for x in xs]
    assert all(a <= b for a, b in zip(ys, ys[1:]))  # non-decreasing
    xs = sorted(random.Random(31).uniform(-1, 1) for _ in range(20000))
    ys = [L.fjo_exp(x) for x in xs]
    assert all(a <= b for a, b in zip(ys, ys[1:]))
    assert L.fjo_exp(709.79) == math.inf and L.fjo_exp(1000.0) == math.inf
    assert L.fjo_exp(-746.0) == 0.0 and L.fjo_exp(-745.0) > 0.0
    assert math.isnan(L.fjo_exp(math.nan))
    rnd = random.Random(3)
    for _ in range(2000):  # exp(a) exp(-a) = 1 to two ulps
        a = rnd.uniform(-300, 300)
        assert abs(L.fjo_exp(a) * L.fjo_exp(-a) - 1.0) <= 4.5e-16


def test_pow_against_decimal_and_glibc():
    """x^p for the weights kt^(2p): x from 1e-300 to 1e12, p over [-3, 3]."""
    L = orc.lib()
    rnd = random.Random(9)
    cases = [(math.exp(rnd.uniform(-690, 27)), rnd.uniform(-3, 3)) for _ in range(6000)]
    cases += [(rnd.uniform(0.01, 1e4), rnd.choice([0.5, -0.5, 1.5, -1.5, 0.25, 2.0, -2.0])) for _ in range(2000)]
    cr = agree = n = 0
    worst = D(0)
    for x, p in cases:
        want = (D(p) * D(x).ln()).exp()
        if want > D("1.7e308") or want < D("2.3e-308"):
            continue
        n += 1
        got = L.fjo_pow(x, p)
        if got == float(want):
            cr += 1
        if got == x ** p:
            agree += 1
        worst = max(worst, err_ulps(got, want))
    assert worst <= 1, worst
    assert cr >= 0.999 * n, cr / n
    assert agree >= 0.995 * n, agree / n


def test_pow_closed_forms():
    L = orc.lib()
    rnd = random.Random(13)
    for _ in range(5000):
        x = math.exp(rnd.uniform(-600, 600))
        assert L.fjo_pow(x, 1.0) == x
        assert L.fjo_pow(x, 0.5) == math.sqrt(x)  # sqrt is correctly rounded, and so is this pow here
    for k in range(-60, 61):
        assert L.fjo_pow(2.0, float(k)) == 2.0 ** k
        assert L.fjo_pow(4.0, k / 2) == 2.0 ** k
    for x in (3.0, 7.0, 10.0, 1e-5):  # squares of small integers and products exactly representable
        assert L.fjo_pow(x, 2.0) == x * x
    assert L.fjo_pow(1e-300, -2.0) == math.inf and L.fjo_pow(1e-200, 2.0) == 0.0
    assert L.fjo_pow(5e-324, 0.5) == math.sqrt(5e-324)  # subnormal argument


def test_sincos_within_one_ulp_and_glibc():
    """|x| < 4 pi (the azimuths that the recombination schemes average), and beyond to 16 pi."""
    L = orc.lib()
    rnd = random.Random(17)
    xs = [rnd.uniform(-4 * math.pi, 4 * math.pi) for _ in range(6000)] + [rnd.uniform(-1e-3, 1e-3) for _ in range(500)]
    xs += [rnd.uniform(-16 * math.pi, 16 * math.pi) for _ in range(1000)]
    xs += [k * math.pi / 4 for k in range(-16, 17)] + [math.pi, 2 * math.pi, math.pi / 2]
    cr = agree = 0
    worst = D(0)
    for x in xs:
        s, c = L.fjo_sin(x), L.fjo_cos(x)
        ws, wc = dec_sincos(x)
        for got, want, lib in ((s, ws, math.sin(x)), (c, wc, math.cos(x))):
            if want != 0 and abs(want) > D("1e-300"):
                worst = max(worst, err_ulps(got, want))
            cr += got == float(want)
            agree += got == lib
    assert worst <= 1, worst
    assert cr >= 0.999 * 2 * len(xs), cr / (2 * len(xs))
    assert agree >= 0.995 * 2 * len(xs), agree / (2 * len(xs))


def test_sincos_identities():
    L = orc.lib()
    assert L.fjo_sin(0.0) == 0.0 and L.fjo_cos(0.0) == 1.0
    rnd = random.Random(19)
    for _ in range(3000):
        x = rnd.uniform(-4 * math.pi, 4 * math.pi)
        s, c = L.fjo_sin(x), L.fjo_cos(x)
        assert abs(s * s + c * c - 1.0) <= 4.5e-16
        assert L.fjo_sin(-x) == -s and L.fjo_cos(-x) == c


def test_weight_gen_reduces_to_the_three_algorithms():
    import numpy as np
    L = orc.lib()
    rnd = random.Random(23)
    kt2s = [math.exp(rnd.uniform(-700, 30)) for _ in range(3000)] + [0.0, 1e-300, 5e-324, 1e-301, math.inf]
    for kt2 in kt2s:
        ws = []
        for p in (1, 0, -1):
            rap, phi, w = np.zeros(1), np.zeros(1), np.zeros(1)
            p4 = np.array([math.sqrt(kt2) if kt2 < math.inf else math.inf, 0.0, 0.0, 1.0])
            L.fjo_set_jet(p4.ctypes.data, p, 0, rap.ctypes.data, phi.ctypes.data, w.ctypes.data)
            ws.append(w[0])
        px = math.sqrt(kt2) if kt2 < math.inf else math.inf
        kt2_exact = px * px + 0.0 * 0.0  # as the kinematics form it
        assert [L.fjo_weight_gen(kt2_exact, float(p)) for p in (1, 0, -1)] == ws


def test_weight_gen_genkt_guard_and_limits():
    L = orc.lib()
    assert L.fjo_weight_gen(0.0, 0.5) == 0.0 and L.fjo_weight_gen(0.0, 2.0) == 0.0
    assert L.fjo_weight_gen(0.0, -0.5) == L.fjo_pow(1e-300, -0.5)  # FastJet's guard for p <= 0
    assert L.fjo_weight_gen(1e-310, -2.0) == math.inf  # (1e-300)^-2 overflows
    assert L.fjo_weight_gen(math.inf, 0.5) == math.inf and L.fjo_weight_gen(math.inf, -0.5) == 0.0
    assert L.fjo_weight_gen(4.0, 0.5) == 2.0 and L.fjo_weight_gen(4.0, -0.5) == 0.5
