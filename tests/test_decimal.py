"""The device loaders' decimal parser (gahhg_decimal.h), built for the host
from the same source and checked against Python's float(), which rounds
correctly as strtod does.

* Every plain decimal of up to 19 significant digits (trailing zeros aside)
  in the normal range is accepted and gives strtod's bits: random
  significands, exponents and point positions, shortest and 17-digit
  renderings of random doubles, exactly representable values written with
  16-19 digits, and the exact halfway cases between doubles (round half to
  even) that fit in 19 digits, including 2^53 + 1.
* Text it must not decide (a nonzero digit past the 19th, subnormal or
  overflowing results, hex, inf/nan, an exponent marker without digits, no
  digits) is refused, so the loader hands it to the host's sscanf.
"""
import os
import random
import shutil
import struct
import subprocess
from fractions import Fraction

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "gahhg_amd", "csrc")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("dec") / "dec")
    subprocess.check_call([cxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC,
                           os.path.join(HERE, "decimal_harness.cpp"), "-o", exe])

    def run(strings):
        out = subprocess.run([exe], input="\n".join(strings) + "\n", capture_output=True, text=True,
                             check=True).stdout.split("\n")
        res = []
        for s, o in zip(strings, out):
            if o.startswith("OK"):
                _, b, k = o.split()
                res.append((struct.unpack("<d", bytes.fromhex(b)[::-1])[0], int(k)))
            else:
                res.append(None)
        return res
    return run


def bits(x):
    return struct.pack("<d", x)


def plain_cases(rng):
    out = []
    for _ in range(40000):
        nd = rng.randint(1, 19)
        digits = str(rng.randint(10 ** (nd - 1), 10 ** nd - 1))
        pos = rng.randint(0, nd)
        s = digits[:pos] + "." + digits[pos:] if rng.random() < 0.8 else digits
        if rng.random() < 0.4:
            s += rng.choice("eE") + rng.choice(["", "+", "-"]) + str(rng.randint(0, 40))
        if rng.random() < 0.5:
            s = rng.choice("+-") + s
        out.append(s)
    for _ in range(20000):
        d = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(64)))[0]
        if d != d or d in (float("inf"), float("-inf")) or (d != 0 and abs(d) < 2.3e-308):
            continue
        out.append(repr(d))
        out.append("%.17g" % d)
        out.append("%.6f" % (d % 360.0 - 180.0))
    for _ in range(10000):
        out.append("%.*f" % (rng.randint(0, 12), rng.uniform(-400, 400)))
    # exactly representable values written with 16-19 digits (the products
    # whose truncated low bits are all ones)
    for _ in range(20000):
        k = rng.randint(0, 60)
        v = rng.randrange(1, 2 ** 53) * 2 ** rng.randint(0, 10)
        s = str(v)
        if len(s) < 19:
            s += "0" * rng.randint(0, 19 - len(s))
            e = len(s) - len(str(v))
            s = s + "e-%d" % e if e else s
        out.append(s)
        f = Fraction(rng.randrange(1, 2 ** 53), 2 ** rng.randint(1, 12))
        d = f.denominator.bit_length() - 1
        t = str(f.numerator * 5 ** d)
        t = t.rjust(d + 1, "0")
        out.append(t[:-d] + "." + t[-d:] + "0" * rng.randint(0, 3))
        del k
    return [s for s in out if "inf" not in s and "nan" not in s]


def halfway_cases(rng):
    out = ["9007199254740993", "9007199254740995", "18014398509481986", "18014398509481990"]
    while len(out) < 3000:
        m = rng.randrange(2 ** 53, 2 ** 54) | 1  # odd 54-bit: halfway between two doubles
        k = rng.randint(-4, 12)
        x = Fraction(m) * Fraction(2) ** k
        if x.denominator == 1:
            s = str(x.numerator)
        else:
            digits = -k
            s = str(x.numerator * 5 ** digits)
            s = s[:-digits] + "." + s[-digits:]
        if len(s.replace(".", "").lstrip("0")) <= 19:
            out.append(s)
    return out


def test_plain_decimals_match_strtod(harness):
    rng = random.Random(5)
    cases = plain_cases(rng) + halfway_cases(rng)
    res = harness(cases)
    refused = []
    for s, r in zip(cases, res):
        exp = float(s)
        sig = len(s.lstrip("+-").split("e")[0].split("E")[0].replace(".", "").strip("0"))
        plain = sig <= 19 and (exp == 0 or 2.2250738585072014e-308 <= abs(exp) < float("inf"))
        if r is None:
            if plain:
                refused.append(s)
            continue
        assert bits(r[0]) == bits(exp), (s, r[0], exp)
        assert r[1] == len(s), s
    assert not refused, refused[:10]


def test_halfway_rounds_to_even(harness):
    r = harness(["9007199254740993", "9007199254740995", "-9007199254740993", "0.5", "2.5e-1"])
    assert [x[0] for x in r] == [9007199254740992.0, 9007199254740996.0, -9007199254740992.0, 0.5, 0.25]


def test_signed_zero_and_forms(harness):
    r = harness(["-0", "0.000", "-0.0e5", ".5", "5.", "00012.3400", "1E+2", "1e-0"])
    assert [x[0] for x in r] == [-0.0, 0.0, -0.0, 0.5, 5.0, 12.34, 100.0, 1.0]
    assert bits(r[0][0]) == bits(-0.0) and bits(r[1][0]) == bits(0.0)


def test_refused_text(harness):
    cases = ["1234567890123456789012", "12345678901234567890.5", "4.9e-324", "1e-320", "1e309", "1.8e308", "0x1p3", "inf", "nan",
             "1e", "1e+", "-", "+.", ".", "e5", ""]
    assert harness(cases) == [None] * len(cases)


def test_stops_at_first_non_number_byte(harness):
    r = harness(["1.5,2.5", "-3e2 7", "12.5abc"])
    assert r == [(1.5, 3), (-300.0, 4), (12.5, 4)]
