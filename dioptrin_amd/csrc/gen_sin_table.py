"""Writes sin_table.inc: double-double values of sin(k/64) and cos(k/64),
k = 0..128, and of the reciprocal factorials used by the certified sine.

Values come from exact rational arithmetic (the Taylor series of sin and cos
at a rational point, summed with Python's Fraction until the terms fall below
2^-300), split into (hi, lo) with hi = RN(v) and lo = RN(v - hi). Run:
python gen_sin_table.py
"""
from fractions import Fraction
import math
import os

PREC = 1 << 400


def sin_cos_frac(x: Fraction):
    s, c = Fraction(0), Fraction(0)
    term, n = Fraction(1), 0       # term = x^n / n!
    while True:
        if abs(term) < Fraction(1, 1 << 300) and n > 2:
            break
        if n % 2 == 0:
            c += term if n % 4 == 0 else -term
        else:
            s += term if n % 4 == 1 else -term
        n += 1
        term = term * x / n
        term = Fraction(int(term * PREC), PREC)
    return s, c


def split(v: Fraction):
    hi = float(v)  # Fraction -> float rounds to nearest
    lo = float(v - Fraction(hi))
    return hi, lo


def main():
    # Cross-check the series: sin(1) and cos(1) against known expansions.
    s1, c1 = sin_cos_frac(Fraction(1))
    assert abs(s1 - Fraction("0.84147098480789650665250232163029899962256306079837")) < \
        Fraction(1, 10 ** 48), "sin series failed"
    assert abs(c1 - Fraction("0.54030230586813971740093660744297660373231042061792")) < \
        Fraction(1, 10 ** 48), "cos series failed"
    lines = ["// Generated by gen_sin_table.py: double-double constants.",
             "// sin(k/64), cos(k/64) for k = 0..128 as (hi, lo) pairs."]
    sins, coss = [], []
    for k in range(129):
        s, c = sin_cos_frac(Fraction(k, 64))
        sins.append(split(s))
        coss.append(split(c))
    for name, vals in (("kSinTable", sins), ("kCosTable", coss)):
        lines.append(f"static const double {name}[129][2] = {{")
        for hi, lo in vals:
            lines.append(f"  {{{hi.hex()}, {lo.hex()}}},")
        lines.append("};")
    lines.append("// 1/n! for n = 2..13 as (hi, lo) pairs (index n - 2).")
    lines.append("static const double kInvFact[12][2] = {")
    for n in range(2, 14):
        hi, lo = split(Fraction(1, math.factorial(n)))
        lines.append(f"  {{{hi.hex()}, {lo.hex()}}},")
    lines.append("};")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sin_table.inc")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
