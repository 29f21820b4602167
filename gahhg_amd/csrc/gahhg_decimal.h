// gahhg_decimal.h -- decimal text to double, correctly rounded, for host and
// device (the device loaders parse with it; a host build of the same source
// is fuzzed against strtod in the tests).
//
// Grammar: what strtod accepts for plain decimal input -- optional sign,
// digits with at most one '.', at least one digit, optional exponent
// ('e'/'E', optional sign, at least one digit).  parse_decimal returns false
// (the caller hands the text to the host's sscanf) for anything else and for
// what it does not convert exactly here: a nonzero digit past the 19th
// significant one,
// hex or inf/nan text, an 'e' without exponent digits, results that would be
// subnormal or overflow.
//
// Conversion of w * 10^q (w < 2^64 exact):
//   * w <= 2^53 and |q| <= 22: w and 10^|q| are exact doubles, so one IEEE
//     multiply or divide is the correctly rounded result (Clinger 1990);
//   * otherwise the Eisel-Lemire method (Lemire, "Number parsing at a
//     gigabyte per second", Softw. Pract. Exp. 51(8), 2021): w normalised to
//     64 bits times a 128-bit approximation of 5^q (gahhg_pow5.h), the
//     product's leading 54 bits rounded half-even, the binary exponent
//     floor(q log2 10) + 63 + (product's top bit) - (w's leading zeros).
//     For w < 2^64 the refined product decides every case, including the
//     one the 2021 paper sent to a fallback (all low product bits ones after
//     the refinement): Mushtak and Lemire, "Fast number parsing without
//     fallback", Softw. Pract. Exp. 53(6), 2023.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define GAHHG_DEC_FN __host__ __device__ inline
#else
#define GAHHG_DEC_FN inline
#endif

#define GAHHG_POW5_STORAGE constexpr
#include "gahhg_pow5.h"

namespace gahhg {
namespace dec {

GAHHG_DEC_FN void mul64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
#ifdef __HIP_DEVICE_COMPILE__
  *lo = a * b;
  *hi = __umul64hi(a, b);
#else
  const unsigned __int128 p = (unsigned __int128)a * b;
  *lo = (uint64_t)p;
  *hi = (uint64_t)(p >> 64);
#endif
}

GAHHG_DEC_FN int clz64(uint64_t w) {
#ifdef __HIP_DEVICE_COMPILE__
  return __clzll((long long)w);
#else
  return __builtin_clzll(w);
#endif
}

GAHHG_DEC_FN double bits_to_double(uint64_t b) {
#ifdef __HIP_DEVICE_COMPILE__
  return __longlong_as_double((long long)b);
#else
  double d;
  __builtin_memcpy(&d, &b, 8);
  return d;
#endif
}

// Exact powers of ten 10^0 .. 10^22 (5^22 < 2^53).
GAHHG_DEC_FN double pow10_exact(int k) {
  double p = 1.0;
  double b = 10.0;
  while (k) {  // exact: every partial product is an exact power of ten <= 1e22
    if (k & 1) p *= b;
    b *= b;
    k >>= 1;
  }
  return p;
}

// w * 10^q rounded to nearest-even, w != 0; false where not decided here.
GAHHG_DEC_FN bool eisel_lemire(uint64_t w, int q, double* out) {
  if (q < kPow5Min || q > kPow5Max) return false;
  const uint64_t* T = kPow5[q - kPow5Min];
  const int lz = clz64(w);
  const uint64_t W = w << lz;
  uint64_t hi, lo;
  mul64(W, T[0], &hi, &lo);
  if ((hi & 0x1ff) == 0x1ff) {  // a carry from the low half could reach the kept bits
    uint64_t h2, l2;
    mul64(W, T[1], &h2, &l2);
    const uint64_t s = lo + h2;
    hi += (s < lo) ? 1 : 0;
    lo = s;
  }
  const int upper = (int)(hi >> 63);
  const int shift = upper + 9;
  uint64_t m = hi >> shift;  // 54 bits: significand and the rounding bit
  // exactly halfway (possible only for -4 <= q <= 23): round to even
  if (q >= -4 && q <= 23 && lo <= 1 && (m << shift) == hi && (m & 3) == 1) m -= 1;
  m = (m + (m & 1)) >> 1;
  int be = ((q * 217706) >> 16) + 63 + upper - lz + 1023;
  if (m == (1ull << 53)) {
    m = 1ull << 52;
    be++;
  }
  if (be <= 0 || be >= 2047) return false;
  *out = bits_to_double((m & ((1ull << 52) - 1)) | ((uint64_t)be << 52));
  return true;
}

GAHHG_DEC_FN bool is_digit(uint8_t ch) { return ch >= '0' && ch <= '9'; }
GAHHG_DEC_FN bool is_space(uint8_t ch) { return ch == ' ' || (ch >= 9 && ch <= 13); }

// Parses the number at s[i] (i < end); on success i is past it.
GAHHG_DEC_FN bool parse_decimal(const uint8_t* s, int64_t& i, int64_t end, double& out) {
  bool neg = false;
  if (i < end && (s[i] == '+' || s[i] == '-')) {
    neg = s[i] == '-';
    i++;
  }
  // strtod reads "0x..." as hexadecimal: left to the host
  if (i + 1 < end && s[i] == '0' && (s[i + 1] == 'x' || s[i + 1] == 'X')) return false;
  uint64_t w = 0;
  int nd = 0, e10 = 0;
  bool any = false;
  // significant digits go into w (at most 19); zeros past the 19th only
  // scale it, any other digit there is left to the host
  while (i < end && is_digit(s[i])) {
    const int d = s[i] - '0';
    any = true;
    if (nd == 19) {
      if (d != 0) return false;
      e10++;
    } else if (w != 0 || d != 0) {
      w = w * 10 + (uint64_t)d;
      nd++;
    }
    i++;
  }
  if (i < end && s[i] == '.') {
    i++;
    while (i < end && is_digit(s[i])) {
      const int d = s[i] - '0';
      any = true;
      if (nd == 19) {
        if (d != 0) return false;
      } else {
        if (w != 0 || d != 0) {
          w = w * 10 + (uint64_t)d;
          nd++;
        }
        e10--;
      }
      i++;
    }
  }
  if (!any) return false;
  if (i < end && (s[i] == 'e' || s[i] == 'E')) {
    int64_t j = i + 1;
    bool eneg = false;
    if (j < end && (s[j] == '+' || s[j] == '-')) {
      eneg = s[j] == '-';
      j++;
    }
    if (j >= end || !is_digit(s[j])) return false;
    int e = 0;
    while (j < end && is_digit(s[j])) {
      e = e * 10 + (s[j] - '0');
      if (e > 100000) e = 100000;
      j++;
    }
    e10 += eneg ? -e : e;
    i = j;
  }
  if (w == 0) {
    out = neg ? -0.0 : 0.0;
    return true;
  }
  double v;
  if (w <= (1ull << 53) && e10 >= -22 && e10 <= 22) {
    v = (double)w;
    v = e10 >= 0 ? v * pow10_exact(e10) : v / pow10_exact(-e10);
  } else if (!eisel_lemire(w, e10, &v)) {
    return false;
  }
  out = neg ? -v : v;
  return true;
}

GAHHG_DEC_FN int64_t skip_space(const uint8_t* s, int64_t i, int64_t end) {
  while (i < end && is_space(s[i])) i++;
  return i;
}

}  // namespace dec
}  // namespace gahhg
