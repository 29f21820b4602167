"""Prints gahhg_pow5.h: 128-bit approximations of 5^q, q = -342..308, for the
Eisel-Lemire decimal-to-double conversion (Lemire, "Number parsing at a
gigabyte per second", Software: Practice and Experience 51(8), 2021; exactness
without fallback for significands of up to 19 digits: Mushtak and Lemire,
"Fast number parsing without fallback", 2023).

Entry q holds m with 2^127 <= m < 2^128 and m ~ 5^q * 2^s for some integer s:
for q >= 0 the leading 128 bits of 5^q (truncated; exact for q <= 55); for
-27 <= q < 0, floor(2^(z+127) / 5^-q) + 1 with z = bit length of 5^-q (an
upper bound with exactly 128 bits, so that exact halfway decimals, which
need q >= -4, are recognised); for q < -27 the leading 128 bits of
floor(2^(2z+128) / 5^-q) + 1.  These are the conventions under which the
papers' bounds on the 64 x 128-bit product hold."""
Q_MIN, Q_MAX = -342, 308


def entry(q):
    if q >= 0:
        m = 5 ** q
        while m < (1 << 127):
            m <<= 1
        while m >= (1 << 128):
            m >>= 1
        return m
    p = 5 ** -q
    z = p.bit_length()
    if q >= -27:  # 2^(z+127) / p has exactly 128 bits: floor + 1 is an upper bound
        m = (1 << (z + 127)) // p + 1
    else:
        m = (1 << (2 * z + 128)) // p + 1
        while m >= (1 << 128):
            m >>= 1
    assert (1 << 127) <= m < (1 << 128)
    return m


def main():
    print("// gahhg_pow5.h -- generated by pow5_table.py; see that script.")
    print("#pragma once")
    print("#include <stdint.h>")
    print("namespace gahhg {")
    print(f"constexpr int kPow5Min = {Q_MIN}, kPow5Max = {Q_MAX};")
    print(f"// {{high 64 bits, low 64 bits}} for q = {Q_MIN} .. {Q_MAX}")
    print(f"GAHHG_POW5_STORAGE uint64_t kPow5[{Q_MAX - Q_MIN + 1}][2] = {{")
    for q in range(Q_MIN, Q_MAX + 1):
        m = entry(q)
        print(f"    {{0x{m >> 64:016x}ull, 0x{m & ((1 << 64) - 1):016x}ull}},")
    print("};")
    print("}  // namespace gahhg")


if __name__ == "__main__":
    main()
