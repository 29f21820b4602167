// gahhg_trig.h -- correctly rounded sin and cos of a double, one source for
// host and device, so that a point converted on the GPU has the same bits as
// one converted on the host.
//
// The reference converts longitude/latitude with the platform's sin and cos
// (FakeLLVertex::FakeLL2XYZ, source/GPISPPolygon.cuh:82-92), whose last bit
// depends on the C library.  This build uses the correctly rounded values
// instead, a function of the input alone:
//   * argument reduction x = k pi/2 + r in double-double arithmetic, with
//     pi/2 in four parts (three of 33 bits, so k * part is exact for
//     |k| < 2^20, and a tail), r kept as an unevaluated sum hi + lo;
//   * Taylor series of sin r and cos r on |r| <= pi/4 (+ rounding slack),
//     the leading terms in double-double, the terms below 2^-52 of the
//     result in double (truncation < 2^-105 relative);
//   * one final rounding of the double-double sum.
// The double-double result carries an error of about 2^-100 relative, so it
// rounds to the correctly rounded double unless the exact value lies within
// that distance of a midpoint between doubles (probability about 2^-47 per
// call).  Correct rounding holds for |x| < 2^20 (about 6e7 degrees); beyond,
// results are deterministic but not correctly rounded.
//
// Exact products use Dekker's splitting (no fused multiply-add), and every
// expression is evaluated in the written order (the library compiles with
// -ffp-contract=off), so host and device give identical bits.
// The constants are printed by trig_constants.py from exact rationals.
#pragma once
#include <math.h>

#ifdef __HIPCC__
#define GAHHG_TRIG_FN __host__ __device__ inline
#define GAHHG_TRIG_UNROLL _Pragma("unroll")
#else
#define GAHHG_TRIG_FN inline
#define GAHHG_TRIG_UNROLL
#endif

namespace gahhg {
namespace trig {

struct DD {
  double hi, lo;
};

GAHHG_TRIG_FN DD two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}

GAHHG_TRIG_FN DD fast_two_sum(double a, double b) {  // |a| >= |b| or a == 0
  const double s = a + b;
  return {s, b - (s - a)};
}

GAHHG_TRIG_FN DD two_prod(double a, double b) {  // Dekker: a * b = p + e exactly
  const double p = a * b;
  const double ca = 134217729.0 * a, ah = ca - (ca - a), al = a - ah;
  const double cb = 134217729.0 * b, bh = cb - (cb - b), bl = b - bh;
  return {p, (((ah * bh - p) + ah * bl) + al * bh) + al * bl};
}

GAHHG_TRIG_FN DD dd_add(DD a, DD b) {
  DD s = two_sum(a.hi, b.hi);
  const DD t = two_sum(a.lo, b.lo);
  s = fast_two_sum(s.hi, s.lo + t.hi);
  return fast_two_sum(s.hi, s.lo + t.lo);
}

GAHHG_TRIG_FN DD dd_mul(DD a, DD b) {
  const DD p = two_prod(a.hi, b.hi);
  return fast_two_sum(p.hi, p.lo + (a.hi * b.lo + a.lo * b.hi));
}

constexpr double kPio2_1 = 0x1.921fb54400000p+0;  // 1.5707963267341256
constexpr double kPio2_2 = 0x1.0b4611a600000p-34;  // 6.077100506303966e-11
constexpr double kPio2_3 = 0x1.3198a2e000000p-69;  // 2.0222662487111665e-21
constexpr double kPio2_4 = 0x1.b839a252049c1p-104;  // 8.4784276603688996e-32
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
// sin: (-1)^i / (2i+1)!, i = 1..14; cos: (-1)^i / (2i)!, i = 1..15
constexpr double kSinC[14][2] = {
    {-0x1.5555555555555p-3, -0x1.5555555555555p-57},
    {0x1.1111111111111p-7, 0x1.1111111111111p-63},
    {-0x1.a01a01a01a01ap-13, -0x1.a01a01a01a01ap-73},
    {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73},
    {-0x1.ae64567f544e4p-26, 0x1.c062e06d1f209p-80},
    {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87},
    {-0x1.ae7f3e733b81fp-41, -0x1.1d8656b0ee8cbp-97},
    {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103},
    {-0x1.2f49b46814157p-57, -0x1.2650f61dbdcb4p-112},
    {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120},
    {-0x1.761b41316381ap-75, 0x1.3423c7d91404fp-130},
    {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139},
    {-0x1.d1ab1c2dccea3p-94, -0x1.054d0c78aea14p-149},
    {0x1.259f98b4358adp-103, 0x1.eaf8c39dd9bc5p-157},
};
constexpr double kCosC[15][2] = {
    {-0x1.0000000000000p-1, 0x0.0p+0},
    {0x1.5555555555555p-5, 0x1.5555555555555p-59},
    {-0x1.6c16c16c16c17p-10, 0x1.f49f49f49f49fp-65},
    {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76},
    {-0x1.27e4fb7789f5cp-22, -0x1.cbbc05b4fa99ap-76},
    {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83},
    {-0x1.93974a8c07c9dp-37, -0x1.05d6f8a2efd1fp-92},
    {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101},
    {-0x1.6827863b97d97p-53, -0x1.eec01221a8b0bp-107},
    {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120},
    {-0x1.0ce396db7f853p-70, 0x1.aebcdbd20331cp-124},
    {0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135},
    {-0x1.88e85fc6a4e5ap-89, 0x1.71c37ebd16540p-143},
    {0x1.0a18a2635085dp-98, 0x1.b9e2e28e1aa54p-153},
    {-0x1.3932c5047d60ep-108, -0x1.832b7b530a627p-162},
};

// sin x and cos x, correctly rounded (see above).
GAHHG_TRIG_FN void sincos_cr(double x, double* s, double* c) {
  if (x == 0.0) {  // keeps the sign of zero
    *s = x;
    *c = 1.0;
    return;
  }
  if (!(fabs(x) <= 1.7976931348623157e308)) {  // inf or nan
    *s = *c = x - x;
    return;
  }
  const double kd = rint(x * kTwoOverPi);
  DD r = {x, 0.0};
  if (kd != 0.0) {
    r = two_sum(x, -(kd * kPio2_1));
    r = dd_add(r, DD{-(kd * kPio2_2), 0.0});
    r = dd_add(r, DD{-(kd * kPio2_3), 0.0});
    const DD p4 = two_prod(kd, kPio2_4);
    r = dd_add(r, DD{-p4.hi, -p4.lo});
  }
  const DD z = dd_mul(r, r);
  // sin r = r + r * sum_{i>=1} kSinC[i-1] z^i; terms i >= 8 in double
  double a = kSinC[13][0];
GAHHG_TRIG_UNROLL
  for (int i = 12; i >= 7; i--) a = kSinC[i][0] + z.hi * a;
  DD A = {a, 0.0};
GAHHG_TRIG_UNROLL
  for (int i = 6; i >= 0; i--) A = dd_add(DD{kSinC[i][0], kSinC[i][1]}, dd_mul(z, A));
  const DD sr = dd_add(r, dd_mul(r, dd_mul(z, A)));
  // cos r = 1 + sum_{i>=1} kCosC[i-1] z^i; terms i >= 9 in double
  double b = kCosC[14][0];
GAHHG_TRIG_UNROLL
  for (int i = 13; i >= 8; i--) b = kCosC[i][0] + z.hi * b;
  DD B = {b, 0.0};
GAHHG_TRIG_UNROLL
  for (int i = 7; i >= 0; i--) B = dd_add(DD{kCosC[i][0], kCosC[i][1]}, dd_mul(z, B));
  const DD cr = dd_add(DD{1.0, 0.0}, dd_mul(z, B));
  // quadrant kd mod 4, exact for any integer-valued kd (kd * 0.25, floor
  // and the products are exact; beyond 2^54 kd is a multiple of 4)
  const int k = (int)(kd - 4.0 * floor(kd * 0.25));
  switch (k) {
    case 0: *s = sr.hi; *c = cr.hi; break;
    case 1: *s = cr.hi; *c = -sr.hi; break;
    case 2: *s = -sr.hi; *c = -cr.hi; break;
    default: *s = -cr.hi; *c = sr.hi; break;
  }
}

}  // namespace trig
}  // namespace gahhg
