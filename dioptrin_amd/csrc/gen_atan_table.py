"""Writes atan_table.inc: double-double values of atan(i/64), i = 0..64, and
the double-double constants used by the certified atan2 (pi, pi/2, 1/(2k+1)).

Values come from exact rational arithmetic (Taylor series of atan with a
rigorous tail bound, evaluated with Python's Fraction at 200 bits), split into
(hi, lo) with hi = RN(v) and lo = RN(v - hi). Run: python gen_atan_table.py
"""
from fractions import Fraction
import math, os

PREC = 1 << 400


def atan_frac(x: Fraction) -> Fraction:
    # atan(x) = 2 atan(x / (1 + sqrt(1 + x^2))) reduces |x| <= 1 below 0.42;
    # then the alternating Taylor series converges fast. Sqrt via integer isqrt.
    def fsqrt(v: Fraction) -> Fraction:
        num = v.numerator * PREC * PREC // v.denominator
        return Fraction(math.isqrt(num), PREC)
    halvings = 0
    while abs(x) > Fraction(1, 8):
        x = x / (1 + fsqrt(1 + x * x))
        halvings += 1
    s, term, k, x2 = Fraction(0), x, 0, x * x
    while True:
        t = term / (2 * k + 1)
        if abs(t) < Fraction(1, 1 << 300):
            break
        s += t if k % 2 == 0 else -t
        term *= x2
        k += 1
        # keep denominators bounded
        term = Fraction(int(term * PREC), PREC)
        s = Fraction(int(s * PREC), PREC)
    return s * (1 << halvings)


def split(v: Fraction):
    hi = float(v)  # Fraction -> float rounds to nearest
    lo = float(v - Fraction(hi))
    return hi, lo


def main():
    pi = 4 * atan_frac(Fraction(1))
    # Cross-check pi against a known 60-digit decimal expansion.
    known = Fraction("3.14159265358979323846264338327950288419716939937510582097494")
    assert abs(pi - known) < Fraction(1, 10 ** 58), "pi series failed"
    lines = ["// Generated by gen_atan_table.py: double-double constants.",
             "// atan(i/64) for i = 0..64 as (hi, lo) pairs."]
    lines.append("static const double kAtanTable[65][2] = {")
    for i in range(65):
        hi, lo = split(atan_frac(Fraction(i, 64)))
        lines.append(f"  {{{hi.hex()}, {lo.hex()}}},")
    lines.append("};")
    for name, v in [("kPi", pi), ("kHalfPi", pi / 2)]:
        hi, lo = split(v)
        lines.append(f"static const double {name}[2] = {{{hi.hex()}, {lo.hex()}}};")
    lines.append("// 1/3, 1/5, 1/7 as (hi, lo) pairs.")
    lines.append("static const double kInvOdd[3][2] = {")
    for d in (3, 5, 7):
        hi, lo = split(Fraction(1, d))
        lines.append(f"  {{{hi.hex()}, {lo.hex()}}},")
    lines.append("};")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "atan_table.inc")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
