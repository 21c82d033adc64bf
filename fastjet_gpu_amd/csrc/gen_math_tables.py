"""Generates fjmi_math_tables.h: the constant tables of the table-driven log and
atan2 in fjmi_math.h, computed with 80-digit decimal arithmetic.

  log:   for j in [90, 182]:  INV[j] = RN(128/j), LOGINV[j] = -log(INV[j]) as a
         double-double (hi + lo), where INV[j] is the exact double value.
  atan:  for j in [0, 64]:    ATAN[j] = atan(j/64) as a double-double.
  exp:   for j in [0, 63]:    2^(j/64) as a double-double; ln(2)/64 split so that
         n * hi is exact for |n| < 2^17; pi/2 as a double-double.

Run from the repository root:  python fastjet_gpu_amd/csrc/gen_math_tables.py
"""
import os
from decimal import Decimal as D, getcontext

getcontext().prec = 80
HERE = os.path.dirname(os.path.abspath(__file__))


def dd(x):
    hi = float(x)
    lo = float(x - D(hi))
    return hi, lo


def atan(x):
    x = D(x)
    n = 0
    while abs(x) > D("0.001"):
        x = x / (1 + (1 + x * x).sqrt())
        n += 1
    s, t, k, x2 = D(0), x, 0, x * x
    while True:
        term = t / (2 * k + 1)
        if abs(term) < D(10) ** -78:
            break
        s += term if k % 2 == 0 else -term
        t *= x2
        k += 1
    return s * (2 ** n)


def pi():
    """pi to the context precision (Machin's formula)."""
    def arctan_inv(k):
        x = D(1) / k
        s, t, n, x2 = D(0), x, 0, x * x
        while True:
            term = t / (2 * n + 1)
            if abs(term) < D(10) ** -79:
                break
            s += term if n % 2 == 0 else -term
            t *= x2
            n += 1
        return s
    return 4 * (4 * arctan_inv(5) - arctan_inv(239))


def main():
    ln2 = D(2).ln()
    import struct
    bits = struct.unpack("<Q", struct.pack("<d", float(ln2)))[0] & ~((1 << 11) - 1)
    ln2hi = struct.unpack("<d", struct.pack("<Q", bits))[0]
    ln2lo = float(ln2 - D(ln2hi))
    lines = ["/* Generated by gen_math_tables.py (80-digit decimal arithmetic); do not edit. */",
             "#ifndef FJMI_MATH_TABLES_H", "#define FJMI_MATH_TABLES_H", "",
             "#define FJMI_LOG_J0 90", "#define FJMI_LOG_J1 182", "",
             "/* ln 2 = FJMI_LN2_HI + FJMI_LN2_LO, FJMI_LN2_HI with 42 significant bits (e * FJMI_LN2_HI is exact) */",
             f"#define FJMI_LN2_HI {ln2hi.hex()}", f"#define FJMI_LN2_LO {ln2lo.hex()}", "",
             "/* {RN(128/j), hi and lo of -log(RN(128/j))} for j = FJMI_LOG_J0..FJMI_LOG_J1 */",
             "FJMI_TABLE const double fjmi_log_tab[][3] = {"]
    for j in range(90, 183):
        inv = float(D(128) / D(j))
        hi, lo = dd(-(D(inv).ln()))
        lines.append(f"    {{{inv.hex()}, {hi.hex()}, {lo.hex()}}}, /* j={j} */")
    lines += ["};", "", "/* {hi, lo} of atan(j/64) for j = 0..64 */", "FJMI_TABLE const double fjmi_atan_tab[][2] = {"]
    for j in range(65):
        hi, lo = dd(atan(D(j) / 64))
        lines.append(f"    {{{hi.hex()}, {lo.hex()}}}, /* j={j} */")
    lines += ["};", ""]
    l64 = ln2 / 64
    bits = struct.unpack("<Q", struct.pack("<d", float(l64)))[0] & ~((1 << 17) - 1)
    l64hi = struct.unpack("<d", struct.pack("<Q", bits))[0]
    l64lo = float(l64 - D(l64hi))
    pio2 = pi() / 2
    ph, pl = dd(pio2)
    lines += ["/* ln(2)/64 = FJMI_LN2_64_HI + FJMI_LN2_64_LO, the high part with 36 significant bits",
              " * (n * FJMI_LN2_64_HI is exact for |n| < 2^17); 64/ln(2) rounded; pi/2 as a double-double */",
              f"#define FJMI_LN2_64_HI {l64hi.hex()}", f"#define FJMI_LN2_64_LO {l64lo.hex()}",
              f"#define FJMI_INV_LN2_64 {float(D(64) / ln2).hex()}",
              f"#define FJMI_PIO2_HI {ph.hex()}", f"#define FJMI_PIO2_LO {pl.hex()}", "",
              "/* {hi, lo} of 2^(j/64) for j = 0..63 */", "FJMI_TABLE const double fjmi_exp2_tab[][2] = {"]
    for j in range(64):
        hi, lo = dd((D(j) / 64 * ln2).exp())
        lines.append(f"    {{{hi.hex()}, {lo.hex()}}}, /* j={j} */")
    lines += ["};", "", "#endif", ""]
    with open(os.path.join(HERE, "fjmi_math_tables.h"), "w") as fh:
        fh.write("\n".join(lines))


if __name__ == "__main__":
    main()
