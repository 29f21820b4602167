"""Prints the constants of gahhg_trig.h, derived exactly with rationals:
pi/2 in four parts (three of 33 significant bits, so that k * part is exact
for |k| < 2^20, and a 53-bit tail), 2/pi rounded to double, and the Taylor
coefficients (-1)^i / (2i+1)! and (-1)^i / (2i)! as double-double pairs
(hi = nearest double, lo = nearest double to the remainder).
pi comes from Machin's formula pi = 16 atan(1/5) - 4 atan(1/239)."""
from fractions import Fraction
import math


def atan_inv(n, terms):
    s = Fraction(0)
    for k in range(terms):
        s += Fraction((-1) ** k, (2 * k + 1) * n ** (2 * k + 1))
    return s


PI = 16 * atan_inv(5, 120) - 4 * atan_inv(239, 40)  # error < 10^-160


def top_bits(x, bits):
    """x rounded toward zero to `bits` significant bits (x > 0)."""
    e = math.floor(math.log2(x))
    scale = Fraction(2) ** (bits - 1 - e)
    return Fraction(math.floor(x * scale)) / scale


def dd(x):
    hi = float(x)
    lo = float(x - Fraction(hi))
    return hi, lo


def main():
    half = PI / 2
    parts = []
    rem = half
    for bits in (33, 33, 33):
        p = top_bits(rem, bits)
        parts.append(p)
        rem -= p
    parts.append(Fraction(float(rem)))
    for i, p in enumerate(parts):
        assert Fraction(float(p)) == p
        print(f"constexpr double kPio2_{i + 1} = {float(p).hex()};  // {float(p):.17g}")
    print(f"constexpr double kTwoOverPi = {float(2 / PI).hex()};")
    print("// sin: (-1)^i / (2i+1)!, i = 1..14; cos: (-1)^i / (2i)!, i = 1..15")
    for name, off in (("kSinC", 1), ("kCosC", 0)):
        n = 14 if off else 15
        rows = []
        for i in range(1, n + 1):
            hi, lo = dd(Fraction((-1) ** i, math.factorial(2 * i + off)))
            rows.append(f"    {{{hi.hex()}, {lo.hex()}}},")
        print(f"constexpr double {name}[{n}][2] = {{")
        print("\n".join(rows))
        print("};")


if __name__ == "__main__":
    main()
