// gahhg_geom.h -- FP64 vector arithmetic and predicates shared by the host
// code and the gfx950 kernels.
//
// Bit-exactness contract: every expression here is a plain sequence of IEEE
// double adds, multiplies, one correctly rounded division and one correctly
// rounded square root, with FMA contraction disabled (the translation units
// are compiled with -ffp-contract=off and the pragma below).  The operation
// order follows the reference's vector type (source/GPISPDS.cuh:30-106):
//   cross  = (y*z'-z*y', z*x'-x*z', x*y'-y*x')
//   dot    = (x*x' + y*y') + z*z'
//   v / k  = v * (1.0/k)
//   normalize(v) = v * (1.0/sqrt(x*x + y*y + z*z))
#pragma once
#include <stdint.h>

#pragma clang fp contract(off)

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GH_HD __host__ __device__ __forceinline__
#else
#define GH_HD inline
#endif

#include <math.h>

namespace gahhg {

// Tolerances and codes: source/GPISPParam.h:6-17.
constexpr double kEps = 1.0e-8;
constexpr double kTravelEps = 1.0e-5;
constexpr double kPi = 3.14159265;  // source/GPISPDS.cuh:7
constexpr int kUndefined = 0;
constexpr int kOut = 1;
constexpr int kIn = 2;
constexpr int kInterNo = 0;
constexpr int kInterX = 1;
constexpr int kInterT1 = 2;
constexpr int kInterT2 = 3;
constexpr int kInterOverlap = 4;
constexpr int kMarked = 999;          // crossed-cell marker, GPISPAHHG.cu:1212
constexpr int64_t kStepCap = 1 << 24; // walk length guard (reference: unbounded)

GH_HD bool eq0(double a) { return a < kEps && a > -kEps; }
GH_HD bool lt0(double a) { return a < -kEps; }
GH_HD bool gt0(double a) { return a > kEps; }
GH_HD bool le0(double a) { return a < kEps; }
GH_HD bool ge0(double a) { return a > -kEps; }
// "strictly on the same side" test used by every arc predicate
GH_HD bool same_side(double a, double b) { return (lt0(a) && lt0(b)) || (gt0(a) && gt0(b)); }

struct V3 {
  double x, y, z;
};

GH_HD V3 v3(double x, double y, double z) {
  V3 r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
GH_HD V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
GH_HD V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
GH_HD V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
GH_HD V3 mul(V3 a, double k) { return v3(a.x * k, a.y * k, a.z * k); }
GH_HD V3 divk(V3 a, double k) {
  double s = 1.0 / k;
  return v3(a.x * s, a.y * s, a.z * s);
}
GH_HD double sqlen(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
GH_HD V3 unit(V3 a) {
  double t = 1.0 / sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return v3(a.x * t, a.y * t, a.z * t);
}
GH_HD V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
GH_HD double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

GH_HD V3 load3(const double* p) { return v3(p[0], p[1], p[2]); }
GH_HD void store3(double* p, V3 a) {
  p[0] = a.x;
  p[1] = a.y;
  p[2] = a.z;
}

// Plane normal of the great circle through a then b:
// normalize(a x (b - a)), source/GPISPAHHG.cu:18-25 (also the first half of
// every arc predicate).
GH_HD V3 arc_normal(V3 a, V3 b) { return unit(cross(a, sub(b, a))); }

// Border-plane vs edge classification given the precomputed border normal nb
// (= arc_normal(v0, v1)) and edge normal ne (= arc_normal(p0, p1)); the same
// values device_isEdgeBorderInter (source/GPISPAHHG.cu:226-284) computes.
GH_HD int edge_border_class_r(double r0, double r1, double r2, double r3, int* reserved);
GH_HD int edge_border_class(V3 nb, V3 ne, V3 p0, V3 p1, V3 v0, V3 v1, int* reserved) {
  double r0 = dot(nb, p0);
  double r1 = dot(nb, p1);
  double r2 = dot(ne, v0);
  double r3 = dot(ne, v1);
  return edge_border_class_r(r0, r1, r2, r3, reserved);
}

// The classification from the four plane values r0 = nb.p0, r1 = nb.p1,
// r2 = ne.v0, r3 = ne.v1.
GH_HD int edge_border_class_r(double r0, double r1, double r2, double r3, int* reserved) {
  if (eq0(r0) && eq0(r1)) return kInterOverlap;
  bool s23 = same_side(r2, r3);
  bool s01 = same_side(r0, r1);
  if (!s23 && eq0(r0)) {
    *reserved = 0;
    return kInterT1;
  }
  if (!s23 && eq0(r1)) {
    *reserved = 1;
    return kInterT1;
  }
  if (eq0(r2) && eq0(r3)) return kInterOverlap;
  if (!s01 && eq0(r2)) {
    *reserved = 0;
    return kInterT2;
  }
  if (!s01 && eq0(r3)) {
    *reserved = 1;
    return kInterT2;
  }
  if (s01 || s23) return kInterNo;
  return kInterX;
}

}  // namespace gahhg
