// gahhg_query.hip -- point-in-spherical-polygon queries on gfx950.
//
// kernel_PISP (source/GPISPAHHG.cu:1422-1504) semantics: locate the first
// level-0 cell (index order) whose border planes all have n.p >= -eps,
// descend trying the centre child first and then the border children in
// border order, and at the leaf return its label, flipped when the arc from
// the leaf centre to p crosses an odd number of the leaf's polygon edges
// (device_isEEInter, :351-378).  A point that reaches no leaf gets 0 (the
// reference leaves that output unwritten).
#include <float.h>
#include <string.h>

#include <algorithm>


#include "gahhg_internal.h"

#pragma clang fp contract(off)

namespace gahhg {

constexpr int kQBlock = 256;

// The reference's inside test (every border plane n.p >= -eps, in border
// order, stopping at the first failing plane: the early exit keeps the
// number of gathered normals, which bounds the query's memory requests,
// as small as possible).
__device__ __forceinline__ int q_inside(const GridView& g, V3 p, int cid) {
  const int bs = g.c_bstart[cid];
  const int bn = g.c_bnum[cid];
  for (int i = 0; i < bn; i++) {
    const double* n = g.b_n + 3 * (int64_t)(bs + i);
    double r = n[0] * p.x + n[1] * p.y + n[2] * p.z;
    if (r < -kEps) return -1;
  }
  return cid;
}

// q_inside that also counts the border planes it evaluates.
__device__ __forceinline__ int q_inside_n(const GridView& g, V3 p, int cid, int& tests) {
  const int bs = g.c_bstart[cid];
  const int bn = g.c_bnum[cid];
  for (int i = 0; i < bn; i++) {
    const double* n = g.b_n + 3 * (int64_t)(bs + i);
    double r = n[0] * p.x + n[1] * p.y + n[2] * p.z;
    if (r < -kEps) {
      tests += i + 1;
      return -1;
    }
  }
  tests += bn;
  return cid;
}

// First-match descent (centre child, then border children in border order)
// from cell cur to the leaf holding p, -1 when no child holds it.
template <bool TRACE>
__device__ __forceinline__ int descend(const GridView& g, V3 p, int cur, int& tests) {
  int sc;
  while ((sc = g.c_subcell[cur]) != -1) {
    int r = TRACE ? q_inside_n(g, p, sc, tests) : q_inside(g, p, sc);
    if (r != -1) {
      cur = r;
      continue;
    }
    const int bs = g.c_bstart[cur], bn = g.c_bnum[cur];
    bool found = false;
    for (int i = 0; i < bn; i++) {
      const int tc = g.b_subcell[bs + i];
      if (tc == -1) continue;
      r = TRACE ? q_inside_n(g, p, tc, tests) : q_inside(g, p, tc);
      if (r != -1) {
        cur = r;
        found = true;
        break;
      }
    }
    if (!found) return -1;
  }
  return cur;
}

// The reference's inside test of cell x on its node record: every border
// plane n.p >= -eps, in border order (the same normals and operation order
// as q_inside).  On a pass, sub / meta receive x's centre child and meta.
template <bool TRACE>
__device__ __forceinline__ bool node_inside(const int32_t* __restrict__ nodes, int x, V3 p, int& sub, int& meta,
                                            int& tests) {
  // all six plane slots are loaded at once (one round trip instead of one
  // per plane); the slots past bn hold zero normals, which pass (0 >= -eps)
  const int32_t* r = nodes + (int64_t)x * kNodeInts;
  const int2 h = *reinterpret_cast<const int2*>(r);
  const double* n = reinterpret_cast<const double*>(r + 10);
  double w[18];
#pragma unroll
  for (int i = 0; i < 18; i++) w[i] = n[i];
  int fail = -1;
#pragma unroll
  for (int i = 5; i >= 0; i--) {
    const double v = w[3 * i] * p.x + w[3 * i + 1] * p.y + w[3 * i + 2] * p.z;
    if (v < -kEps) fail = i;
  }
  if (fail >= 0) {
    if (TRACE) tests += fail + 1;
    return false;
  }
  if (TRACE) tests += h.y & 0xff;
  sub = h.x;
  meta = h.y;
  return true;
}

// descend() on the node records: one record per level holds the chosen
// child's planes, its own centre child and border children, type and refs
// flag.  Returns the leaf (-1 = none) and its meta word.
template <bool TRACE>
__device__ __forceinline__ int descend_nodes(const int32_t* __restrict__ nodes, V3 p, int cur, int& meta,
                                             int& tests) {
  const int2 h = *reinterpret_cast<const int2*>(nodes + (int64_t)cur * kNodeInts);
  int sub = h.x;
  meta = h.y;
  while (sub != -1) {
    int nsub = 0, nmeta = 0;
    if (node_inside<TRACE>(nodes, sub, p, nsub, nmeta, tests)) {
      cur = sub;
      sub = nsub;
      meta = nmeta;
      continue;
    }
    const int32_t* r = nodes + (int64_t)cur * kNodeInts;
    const int nch = (meta >> 8) & 0xff;
    bool found = false;
    for (int i = 0; i < nch; i++) {
      const int tc = r[2 + i];
      if (node_inside<TRACE>(nodes, tc, p, nsub, nmeta, tests)) {
        cur = tc;
        sub = nsub;
        meta = nmeta;
        found = true;
        break;
      }
    }
    if (!found) return -1;
  }
  return cur;
}

// Node record of every cell (layout at QueryLayout::nodes); a cell with
// more than 6 borders (none in a built or validated loaded hierarchy) sets
// st->bad_nodes, and the layout is refused.
__global__ void k_pack_nodes(GridView g, int64_t C, int32_t* __restrict__ nodes, DevState* st) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  int32_t* r = nodes + c * kNodeInts;
  const int bs = g.c_bstart[c], bn = g.c_bnum[c];
  if (bn > 6) st->bad_nodes = 1;  // any writer: same value
  const int b6 = min(bn, 6);
  const int sc = g.c_subcell[c];
  int nch = 0;
  for (int i = 0; i < 8; i++) r[i] = -1;
  if (sc != -1)
    for (int i = 0; i < b6; i++) {
      const int tc = g.b_subcell[bs + i];
      if (tc != -1) r[2 + nch++] = tc;
    }
  r[0] = sc;
  r[1] = b6 | (nch << 8) | ((g.c_type[c] & 0xff) << 16) | ((g.c_rnum[c] != 0 ? 1 : 0) << 24);
  r[8] = r[9] = 0;
  double* n = reinterpret_cast<double*>(r + 10);
  for (int i = 0; i < 6; i++) {
    const V3 v = i < b6 ? load3(g.b_n + 3 * (int64_t)(bs + i)) : v3(0.0, 0.0, 0.0);
    n[3 * i] = v.x, n[3 * i + 1] = v.y, n[3 * i + 2] = v.z;
  }
  r[46] = r[47] = 0;
}

// Plain first-match descent from level 0 to the leaf holding p (-1 = none).
__device__ int locate_leaf(const GridView& g, V3 p, int cell0) {
  int cur = -1;
  for (int i = 0; i < kGrid0Cells; i++) {
    cur = q_inside(g, p, cell0 + i);
    if (cur != -1) break;
  }
  if (cur == -1) return -1;
  int sc;
  while ((sc = g.c_subcell[cur]) != -1) {
    int r = q_inside(g, p, sc);
    if (r != -1) {
      cur = r;
      continue;
    }
    const int bs = g.c_bstart[cur], bn = g.c_bnum[cur];
    bool found = false;
    for (int i = 0; i < bn; i++) {
      const int tc = g.b_subcell[bs + i];
      if (tc == -1) continue;
      r = q_inside(g, p, tc);
      if (r != -1) {
        cur = r;
        found = true;
        break;
      }
    }
    if (!found) return -1;
  }
  return cur;
}

__device__ __forceinline__ int level_of(const LevelTab& lt, int cid) {
  int l = 0;
  while (l + 1 < lt.n && cid >= lt.cstart[l + 1]) l++;
  return l;
}

// Crossing test of the arc c -> p against polygon edge (a, b) with plane
// normals nq = arc_normal(c, p) and ne = arc_normal(a, b): the four dot
// products device_isEEInter evaluates, in its order.
__device__ __forceinline__ bool ee_cross(V3 nq, V3 ne, V3 a, V3 b, V3 c, V3 p) {
  if (same_side(dot(nq, a), dot(nq, b))) return false;
  if (same_side(dot(ne, c), dot(ne, p))) return false;
  return true;
}

// Straightforward query over the build layout.
__global__ void __launch_bounds__(kQBlock) k_query_soa(GridView g, const PolyDev* __restrict__ pd,
                                                       const double* __restrict__ edge,
                                                       const double* __restrict__ xyz,
                                                       const int32_t* __restrict__ qpoly, int64_t n,
                                                       int32_t* __restrict__ labels) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const V3 p = load3(xyz + 3 * q);
  const int cell0 = pd[qpoly ? qpoly[q] : 0].cell0;
  int cur = -1;
  for (int i = 0; i < kGrid0Cells; i++) {
    cur = q_inside(g, p, cell0 + i);
    if (cur != -1) break;
  }
  if (cur == -1) {
    labels[q] = kUndefined;
    return;
  }
  while (g.c_subcell[cur] != -1) {
    int r = q_inside(g, p, g.c_subcell[cur]);
    if (r != -1) {
      cur = r;
      continue;
    }
    const int bs = g.c_bstart[cur], bn = g.c_bnum[cur];
    bool found = false;
    for (int i = 0; i < bn; i++) {
      const int tc = g.b_subcell[bs + i];
      if (tc == -1) continue;
      r = q_inside(g, p, tc);
      if (r != -1) {
        cur = r;
        found = true;
        break;
      }
    }
    if (!found) {
      labels[q] = kUndefined;
      return;
    }
  }
  const int rn = g.c_rnum[cur];
  const int type = g.c_type[cur];
  if (rn == 0) {
    labels[q] = type;
    return;
  }
  const int bs = g.c_bstart[cur], bn = g.c_bnum[cur];
  V3 c = v3(0.0, 0.0, 0.0);
  for (int i = 0; i < bn; i++) c = add(c, load3(g.v + 3 * (int64_t)g.b_v0[bs + i]));
  c = unit(divk(c, bn));
  const V3 nq = arc_normal(c, p);
  const int rs = g.c_rstart[cur];
  int cnt = 0;
  for (int i = 0; i < rn; i++) {
    const double* er = edge + 9 * (int64_t)g.refs[rs + i];
    if (ee_cross(nq, load3(er + 6), load3(er), load3(er + 3), c, p)) cnt++;
  }
  labels[q] = (cnt & 1) ? 3 - type : type;
}

__global__ void __launch_bounds__(256) k_label_counts(const int32_t* __restrict__ labels, int64_t n,
                                                      unsigned long long* out2) {
  __shared__ unsigned long long part[2][4];
  unsigned long long in = 0, un = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int l = labels[i];
    in += l == kIn;
    un += l == kUndefined;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    in += __shfl_xor(in, o, 64);
    un += __shfl_xor(un, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    part[0][threadIdx.x >> 6] = in;
    part[1][threadIdx.x >> 6] = un;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long a = part[0][0] + part[0][1] + part[0][2] + part[0][3];
    const unsigned long long b = part[1][0] + part[1][1] + part[1][2] + part[1][3];
    if (a) atomicAdd(&out2[0], a);
    if (b) atomicAdd(&out2[1], b);
  }
}

// ---------------------------------------------------------------------------
// Start-cell table.  Bin (f, i, j) of a cube map: face f in {+x,-x,+y,-y,
// +z,-z}, (u, v) in [-1,1]^2 split into bins x bins squares; the bin of p is
// found from the dominant axis with one division (any rounding of that
// choice is covered by the cap margin below).
// ---------------------------------------------------------------------------
__device__ __forceinline__ V3 face_point(int f, double u, double v) {
  switch (f) {
    case 0: return v3(1.0, u, v);
    case 1: return v3(-1.0, u, v);
    case 2: return v3(u, 1.0, v);
    case 3: return v3(u, -1.0, v);
    case 4: return v3(u, v, 1.0);
    default: return v3(u, v, -1.0);
  }
}

// Bin of p at B bins per face side, and p's position (fx, fy) inside that
// bin in [0, 1) for the sub-table lookup.
__device__ __forceinline__ int bin_of(V3 p, int B, double& fx, double& fy) {
  const double ax = fabs(p.x), ay = fabs(p.y), az = fabs(p.z);
  int f;
  double m, a, b;
  if (ax >= ay && ax >= az) {
    f = p.x >= 0.0 ? 0 : 1;
    m = ax;
    a = p.y;
    b = p.z;
  } else if (ay >= az) {
    f = p.y >= 0.0 ? 2 : 3;
    m = ay;
    a = p.x;
    b = p.z;
  } else {
    f = p.z >= 0.0 ? 4 : 5;
    m = az;
    a = p.x;
    b = p.y;
  }
  if (!(m > 0.0)) return -1;
  const double inv = 1.0 / m;
  const double x = (a * inv + 1.0) * 0.5 * B;
  const double y = (b * inv + 1.0) * 0.5 * B;
  int i = (int)floor(x);
  int j = (int)floor(y);
  i = min(max(i, 0), B - 1);
  j = min(max(j, 0), B - 1);
  fx = x - i;
  fy = y - j;
  return (f * B + i) * B + j;
}

// Conservative state of the reference's inside test (n.p >= -eps for every
// border, in border order) over a cap (centre u, angular radius rho, given
// as cr = cos rho, sr = sin rho): 0 = fails for every point, 1 = passes for
// every point, 2 = undecided.  With theta the angle between n and u, n.p
// over the cap spans [|n| cos(min(theta + rho, pi)), |n| cos(max(theta -
// rho, 0))]; the cosines come from cos theta and sin theta by the addition
// formulas (no inverse or direct trigonometric call per border).  Near the
// decision (n.p near 0, theta near pi/2 -+ rho) every term is well
// conditioned, so the bounds are good to ~1e-15; they carry a 1e-10 margin.
__device__ int cell_state_cap(const GridView& g, int cid, V3 u, double cr, double sr) {
  const int bs = g.c_bstart[cid], bn = g.c_bnum[cid];
  bool all_in = true;
  for (int i = 0; i < bn; i++) {
    const V3 n = load3(g.b_n + 3 * (int64_t)(bs + i));
    const double nn = sqrt(dot(n, n));
    const double ct = fmin(1.0, fmax(-1.0, dot(n, u) / nn));
    const double st = sqrt(fmax(0.0, (1.0 - ct) * (1.0 + ct)));
    // theta + rho <= pi  <=>  cos theta >= -cos rho;  theta >= rho  <=>  cos theta <= cos rho
    const double lo = (ct >= -cr ? nn * (ct * cr - st * sr) : -nn) - 1e-10;
    const double hi = (ct <= cr ? nn * (ct * cr + st * sr) : nn) + 1e-10;
    if (hi < -kEps) return 0;  // this border rejects every point of the cap
    if (!(lo >= -kEps)) all_in = false;
  }
  return all_in ? 1 : 2;
}

// Cap (centre, angular radius) around sub-bin (si, sj) of an S x S split of
// bin (i, j) of face f at B bins per side; S = 1 is the bin itself.  The
// radius carries a 1e-9 margin over the exact bin, far above the rounding
// of the bin coordinates a query computes.
__device__ void bin_cap(int f, int B, int i, int j, int S, int si, int sj, V3& cen, double& rho) {
  const double u0 = -1.0 + 2.0 * (i + (double)si / S) / B, u1 = -1.0 + 2.0 * (i + (double)(si + 1) / S) / B;
  const double w0 = -1.0 + 2.0 * (j + (double)sj / S) / B, w1 = -1.0 + 2.0 * (j + (double)(sj + 1) / S) / B;
  cen = unit(face_point(f, 0.5 * (u0 + u1), 0.5 * (w0 + w1)));
  rho = 0.0;
  const double us[2] = {u0, u1}, ws[2] = {w0, w1};
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      const V3 cr = unit(face_point(f, us[a], ws[b]));
      rho = fmax(rho, acos(fmin(1.0, dot(cen, cr))));
    }
  rho += 1e-9;
}

// Table entry for the points of a cap: the deepest cell that the
// first-match descent certainly passes through for every point of the cap,
// starting from cell `from` (-1: the level-0 scan of the polygon's cells
// cell0 .. cell0 + 41).  Codes: >= 0 start cell, -1 start with the level-0
// scan, -(2 + label) resolved leaf without refs.
// cell_state_cap on a node record: the same normals, in the same order,
// from one fetch.
__device__ int node_state_cap(const int32_t* __restrict__ nodes, int cid, V3 u, double cr, double sr) {
  const int32_t* r = nodes + (int64_t)cid * kNodeInts;
  const int bn = r[1] & 0xff;
  const double* nr = reinterpret_cast<const double*>(r + 10);
  bool all_in = true;
  for (int i = 0; i < bn; i++) {
    const V3 n = v3(nr[3 * i], nr[3 * i + 1], nr[3 * i + 2]);
    const double nn = sqrt(dot(n, n));
    const double ct = fmin(1.0, fmax(-1.0, dot(n, u) / nn));
    const double st = sqrt(fmax(0.0, (1.0 - ct) * (1.0 + ct)));
    const double lo = (ct >= -cr ? nn * (ct * cr - st * sr) : -nn) - 1e-10;
    const double hi = (ct <= cr ? nn * (ct * cr + st * sr) : nn) + 1e-10;
    if (hi < -kEps) return 0;
    if (!(lo >= -kEps)) all_in = false;
  }
  return all_in ? 1 : 2;
}

// Level-0 part of cap_entry: the level-0 cell (0..41, relative to cell0)
// that the scan certainly matches first for every point of the cap, or -1.
// With node records (nodes != nullptr) every cell test is one record fetch.
__device__ int cap_level0(const GridView& g, const int32_t* __restrict__ nodes, int cell0, V3 cen, double cr,
                          double sr) {
  for (int c = 0; c < kGrid0Cells; c++) {
    const int s = nodes ? node_state_cap(nodes, cell0 + c, cen, cr, sr) : cell_state_cap(g, cell0 + c, cen, cr, sr);
    if (s == 0) continue;
    return s == 1 ? c : -1;  // certain first match, or undecided: stop either way
  }
  return -1;
}

__device__ int32_t cap_descend(const GridView& g, const int32_t* __restrict__ nodes, int cur, V3 cen, double cr,
                               double sr);

__device__ int32_t cap_entry(const GridView& g, const int32_t* __restrict__ nodes, int cell0, int from, V3 cen,
                             double rho) {
  const double cr = cos(rho), sr = sin(rho);
  int cur = from;
  if (cur < 0) {
    const int c = cap_level0(g, nodes, cell0, cen, cr, sr);
    if (c < 0) return -1;
    cur = cell0 + c;
  }
  return cap_descend(g, nodes, cur, cen, cr, sr);
}

// From cell cur: descend while the first-match child is certain.  With node
// records a level reads the record of cur (centre child, border children)
// and one record per candidate child.
__device__ int32_t cap_descend(const GridView& g, const int32_t* __restrict__ nodes, int cur, V3 cen, double cr,
                               double sr) {
  if (nodes) {
    for (;;) {
      const int32_t* r = nodes + (int64_t)cur * kNodeInts;
      const int sc = r[0];
      if (sc == -1) break;
      const int nch = (r[1] >> 8) & 0xff;
      int next = -1;
      bool stop = false;
      int s = node_state_cap(nodes, sc, cen, cr, sr);
      if (s == 1) {
        next = sc;
      } else if (s == 2) {
        stop = true;
      } else {
        for (int k = 0; k < nch; k++) {
          const int tc = r[2 + k];
          s = node_state_cap(nodes, tc, cen, cr, sr);
          if (s == 0) continue;
          if (s == 1) next = tc;
          else stop = true;
          break;
        }
      }
      if (stop || next < 0) break;
      cur = next;
    }
    const int2 h = *reinterpret_cast<const int2*>(nodes + (int64_t)cur * kNodeInts);
    if (h.x == -1 && (h.y >> 24) == 0) return -(2 + ((h.y >> 16) & 0xff));
    return cur;
  }
  for (;;) {
    const int sc = g.c_subcell[cur];
    if (sc == -1) break;
    int next = -1;
    bool stop = false;
    int s = cell_state_cap(g, sc, cen, cr, sr);
    if (s == 1) {
      next = sc;
    } else if (s == 2) {
      stop = true;
    } else {
      const int bs = g.c_bstart[cur], bn = g.c_bnum[cur];
      for (int k = 0; k < bn; k++) {
        const int tc = g.b_subcell[bs + k];
        if (tc == -1) continue;
        s = cell_state_cap(g, tc, cen, cr, sr);
        if (s == 0) continue;
        if (s == 1) next = tc;
        else stop = true;
        break;
      }
    }
    if (stop || next < 0) break;
    cur = next;
  }
  if (g.c_subcell[cur] == -1 && g.c_rnum[cur] == 0) return -(2 + g.c_type[cur]);
  return cur;
}

// Bins per cube-face side of the shared level-0 lookup of a batch's
// table-less polygons (24576 entries, 96 KB).
constexpr int kL0Bins = 64;

// First entry of the level-0 table of B bins per side (B = 16 * 2^k): the
// tables of 16, 32, ... bins follow one another.
__host__ __device__ __forceinline__ int64_t l0_offset(int B) {
  int64_t off = 0;
  for (int b = 16; b < B; b *= 2) off += 6LL * b * b;
  return off;
}

// Level-0 part of every bin's entry for Bmin, 2 Bmin, ..., Bmax bins per
// side (powers of two >= 16), written from out[0] on: a relative level-0
// cell (0..41) that every point of the bin reaches first, or -1.  Level-0
// geometry is the same for every polygon and build (polygon 0's cells are
// used), so it is computed once per resolution and context.
__global__ void k_level0_table(GridView g, const int32_t* __restrict__ nodes, int cell0, int Bmin, int Bmax,
                               int32_t* __restrict__ out) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t t = t0 + l0_offset(Bmin);
  if (t >= l0_offset(2 * Bmax)) return;
  int B = Bmin;
  while (t >= l0_offset(2 * B)) B *= 2;
  const int64_t u = t - l0_offset(B);
  const int f = (int)(u / ((int64_t)B * B));
  const int i = (int)((u / B) % B);
  const int j = (int)(u % B);
  V3 cen;
  double rho;
  bin_cap(f, B, i, j, 1, 0, 0, cen, rho);
  out[t0] = cap_level0(g, nodes, cell0, cen, cos(rho), sin(rho));
}

// Entry tt of all polygons' tables: the polygon owning it is found by a
// binary search over the tables' first entries (tstart, np + 1 values).
// With l0 (power-of-two bins per side), the level-0 part comes from
// k_level0_table.
__global__ void k_build_table(GridView g, const int32_t* __restrict__ nodes, const PolyDev* __restrict__ pd,
                              int npoly, const int32_t* __restrict__ tstart, const int32_t* __restrict__ l0,
                              int32_t* __restrict__ table_all) {
  const int64_t tt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tt >= tstart[npoly]) return;
  int lo = 0, hi = npoly - 1;  // last polygon with tstart <= tt
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tstart[mid] <= tt) lo = mid;
    else hi = mid - 1;
  }
  const int poly = lo;
  const int B = pd[poly].tbins;
  const int64_t t = tt - tstart[poly];
  const int f = (int)(t / ((int64_t)B * B));
  const int i = (int)((t / B) % B);
  const int j = (int)(t % B);
  V3 cen;
  double rho;
  bin_cap(f, B, i, j, 1, 0, 0, cen, rho);
  if (l0) {
    const int c = l0[l0_offset(B) + t];
    table_all[tt] = c < 0 ? -1 : cap_descend(g, nodes, pd[poly].cell0 + c, cen, cos(rho), sin(rho));
  } else {
    table_all[tt] = cap_entry(g, nodes, pd[poly].cell0, -1, cen, rho);
  }
}

// A bin gets a sub-table when its entry is not final: the level-0 scan or a
// cell that still has children.
__global__ void k_sub_flag(GridView g, const int32_t* __restrict__ table, int64_t n, int32_t* __restrict__ flag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int e = table[t];
  flag[t] = (e == -1 || (e >= 0 && g.c_subcell[e] != -1)) ? 1 : 0;
}

// Flagged bins per polygon from the exclusive scan of the flags.
__global__ void k_sub_counts(int npoly, int64_t per, const int32_t* __restrict__ flag,
                             const int32_t* __restrict__ scan, int32_t* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npoly) return;
  const int64_t a = per * p, z = per * (p + 1) - 1;
  out[p] = scan[z] + flag[z] - scan[a];
}

// Sub-table size of every bin (S^2 when flagged and the polygon has S > 1)
// and whether it gets one.
__global__ void k_sub_size(const PolyDev* __restrict__ pd, int64_t per, int64_t n, const int32_t* __restrict__ flag,
                           int32_t* __restrict__ size, int32_t* __restrict__ has) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int S = pd[t / per].tsub;
  const int h = flag[t] && S > 1;
  size[t] = h ? S * S : 0;
  has[t] = h;
}

__global__ void k_sub_list(int64_t n, const int32_t* __restrict__ has, const int32_t* __restrict__ hscan,
                           int32_t* __restrict__ list) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  if (has[t]) list[hscan[t]] = (int32_t)t;
}

// One 64-lane block per bin with a sub-table: its S^2 entries, each
// descending from the bin's own entry.
__global__ void __launch_bounds__(64, 6) k_sub_build(GridView g, const int32_t* __restrict__ nodes,
                                                  const PolyDev* __restrict__ pd, int B, int64_t per,
                                                  const int32_t* __restrict__ list,
                                                  const int32_t* __restrict__ table,
                                                  const int32_t* __restrict__ off, int32_t* __restrict__ sub) {
  const int64_t tt = list[blockIdx.x];
  const int poly = (int)(tt / per);
  const int64_t t = tt - per * poly;
  const int f = (int)(t / ((int64_t)B * B));
  const int i = (int)((t / B) % B);
  const int j = (int)(t % B);
  const PolyDev& P = pd[poly];
  const int S = P.tsub;
  const int from = table[tt];
  int32_t* o = sub + off[tt];
  for (int s = threadIdx.x; s < S * S; s += 64) {
    V3 cen;
    double rho;
    bin_cap(f, B, i, j, S, s / S, s % S, cen, rho);
    o[s] = cap_entry(g, nodes, P.cell0, from, cen, rho);
  }
}

__global__ void k_sub_link(int64_t n, const int32_t* __restrict__ has, const int32_t* __restrict__ off,
                           int32_t* __restrict__ table) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  if (has[t]) table[t] = -(kSubCode + off[t]);
}

__device__ __forceinline__ bool chained(const int32_t* pe, int e, int e2) { return pe[2 * e + 1] == pe[2 * e2]; }

// Records of a bottom cell's edge stream: its refs plus one closing record
// per chain.
__device__ __forceinline__ bool is_bottom(const GridView& g, const PolyDev* pd, int cid) {
  const PolyDev& P = pd[g.c_poly[cid]];
  return cid >= P.bcs && cid < P.bcs + P.bcn;
}

// G lanes per cell (G a power of two up to 64): lane i of the group looks
// at ref i (chunks of G); a ref closes a chain when the next ref does not
// continue it (or it is the cell's last ref).  Conditions are uniform within
// a group.
__device__ __forceinline__ bool chain_break(const int32_t* pe, const int32_t* r, int rn, int i) {
  return i + 1 == rn || !chained(pe, r[i], r[i + 1]);
}

__device__ __forceinline__ unsigned long long group_mask(int lane, int G) {
  return G == 64 ? ~0ull : (((1ull << G) - 1ull) << (lane & ~(G - 1)));
}

__global__ void __launch_bounds__(256) k_leaf_count(GridView g, const PolyDev* __restrict__ pd,
                                                    const int32_t* __restrict__ pe, int cnum,
                                                    int32_t* __restrict__ cnt, int G) {
  const int t = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G);
  const int lane = threadIdx.x & 63, gl = lane & (G - 1);
  const unsigned long long gm = group_mask(lane, G);
  int rn = 0, chains = 0;
  if (t < cnum) {
    const int cid = t;
    rn = is_bottom(g, pd, cid) ? g.c_rnum[cid] : 0;
    if (rn > 0) {
      const int32_t* r = g.refs + g.c_rstart[cid];
      for (int base = 0; base < rn; base += G) {
        const int i = base + gl;
        chains += __popcll(__ballot(i < rn && chain_break(pe, r, rn, i)) & gm);
      }
    }
    if (gl == 0) cnt[t] = rn > 0 ? rn + chains : 0;
  }
}

// Per polygon sums over its bottom cells, kSumChunk cells per block (grid:
// chunks x polygons), one atomic per block and sum into zeroed totals.
constexpr int kSumChunk = 8192;

// pcnt[2p] += edge-stream records, pcnt[2p+1] += leaves with refs.
__global__ void __launch_bounds__(256) k_leaf_poly_sums(const PolyDev* __restrict__ pd,
                                                        const int32_t* __restrict__ cnt,
                                                        int32_t* __restrict__ pcnt) {
  __shared__ int part[2][4];
  const PolyDev& P = pd[blockIdx.y];
  const int i0 = blockIdx.x * kSumChunk, i1 = min(P.bcn, i0 + kSumChunk);
  int rec = 0, leaves = 0;
  for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
    const int k = cnt[P.bcs + i];
    rec += k;
    leaves += k > 0 ? 1 : 0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    rec += __shfl_xor(rec, o, 64);
    leaves += __shfl_xor(leaves, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    part[0][threadIdx.x >> 6] = rec;
    part[1][threadIdx.x >> 6] = leaves;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int r = part[0][0] + part[0][1] + part[0][2] + part[0][3];
    const int l = part[1][0] + part[1][1] + part[1][2] + part[1][3];
    if (r) atomicAdd(&pcnt[2 * blockIdx.y], r);
    if (l) atomicAdd(&pcnt[2 * blockIdx.y + 1], l);
  }
}

// Centre and edge stream of every bottom cell with refs (G lanes per cell):
// ref i goes to record k0 + i + (chain breaks before i), and a closing
// record (the chain's last vertex) follows every break.
__global__ void __launch_bounds__(256) k_leaf_write(GridView g, const PolyDev* __restrict__ pd,
                                                    const int32_t* __restrict__ pe,
                                                    const double* __restrict__ edge, int cnum,
                                                    const int32_t* __restrict__ roff, double* __restrict__ qc,
                                                    int32_t* __restrict__ hdr, double* __restrict__ sv,
                                                    double* __restrict__ sn, uint8_t* __restrict__ sf, int G) {
  const int t = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G);
  const int lane = threadIdx.x & 63, gl = lane & (G - 1);
  const unsigned long long gm = group_mask(lane, G);
  if (t >= cnum) return;
  const int cid = t;
  const int rn = is_bottom(g, pd, cid) ? g.c_rnum[cid] : 0;
  const int k0 = roff[t];
  if (rn <= 0) {
    if (gl == 0) {
      hdr[2 * t] = k0;
      hdr[2 * t + 1] = 0;
    }
    return;
  }
  const int bs = g.c_bstart[cid], bn = g.c_bnum[cid];
  V3 c = v3(0.0, 0.0, 0.0);
  for (int i = 0; i < bn; i++) c = add(c, load3(g.v + 3 * (int64_t)g.b_v0[bs + i]));
  c = unit(divk(c, bn));
  if (gl == 0) store3(qc + 3 * (int64_t)t, c);
  const int32_t* r = g.refs + g.c_rstart[cid];
  int carry = 0;  // chain breaks in earlier chunks
  for (int base = 0; base < rn; base += G) {
    const int i = base + gl;
    const bool brk = i < rn && chain_break(pe, r, rn, i);
    const unsigned long long bal = __ballot(brk) & gm;
    if (i < rn) {
      const int before = carry + __popcll(bal & ((1ull << lane) - 1ull));
      const int k = k0 + i + before;
      // the edge record holds the edge's vertices (copies of pv at pe[2e],
      // pe[2e + 1]) beside its normal: one fetch instead of pe then pv
      const int e = r[i];
      const double* er = edge + 9 * (int64_t)e;
      const V3 ne = load3(er + 6);
      const double r2 = dot(ne, c);
      store3(sv + 3 * (int64_t)k, load3(er));
      store3(sn + 3 * (int64_t)k, ne);
      sf[k] = (uint8_t)(4 | (lt0(r2) ? 1 : 0) | (gt0(r2) ? 2 : 0));
      if (brk) {
        store3(sv + 3 * (int64_t)(k + 1), load3(er + 3));
        store3(sn + 3 * (int64_t)(k + 1), v3(0.0, 0.0, 0.0));
        sf[k + 1] = 0;
      }
    }
    carry += __popcll(bal);
  }
  if (gl == 0) {
    hdr[2 * t] = k0;
    hdr[2 * t + 1] = rn + carry;
  }
}

// ---- angular index of a leaf's edge stream --------------------------------
// The crossing test of record k (edge a -> b of the leaf's stream) against
// the arc from the leaf centre c to a point p fails at its first condition
// when a and b lie strictly on one side of the arc's plane: dot(nq, a) and
// dot(nq, b) both below -eps or both above eps, nq = arc_normal(c, p).  nq
// is perpendicular to c, so it is fixed by one angle psi in the tangent
// plane of c: nq = cos(psi) u + sin(psi) v, and dot(nq, x) =
// R_x cos(psi - phi_x) for x = R_x (cos phi_x u + sin phi_x v) + (c.x) c.
// The signs of dot(nq, a) and dot(nq, b) can only differ, or either come
// within eps of 0, for psi (mod pi) on the arc from phi_a + pi/2 to
// phi_b + pi/2, widened by asin(2 eps / R) + 1e-7 rad at each end (this
// absorbs the rounding of nq, of the dot products and of the angles).  A
// leaf with kIdxMin or more records gets kIdxB buckets over psi in [0, pi),
// each listing the records whose widened arc meets it; a point then tests
// only the records of its psi's bucket.  The records left out fail the
// first condition for every psi of the bucket, so the crossing count, and
// the label, are those of the full stream.
constexpr int kIdxB = 16;
constexpr int kIdxMin = 16;
constexpr double kPiD = 3.141592653589793;  // pi rounded to double (angles only)

// Tangent basis (u, v) at c, the same in the build and the query.
__device__ __forceinline__ void leaf_basis(V3 c, V3& u, V3& v) {
  const V3 axis = fabs(c.y) < 0.9 ? v3(0.0, 1.0, 0.0) : v3(1.0, 0.0, 0.0);
  u = unit(cross(c, axis));
  v = cross(c, u);
}

// Bucket of psi = angle of (nq.u, nq.v) mod pi, by a binary search over
// the bucket boundaries j pi / kIdxB (no transcendental in the query; a psi
// within rounding of a boundary may land on either side, which the 1e-7 rad
// widening of the records' arcs covers).
__device__ __forceinline__ int psi_bucket(V3 nq, V3 u, V3 v) {
  double x = dot(nq, u), y = dot(nq, v);
  if (y < 0.0 || (y == 0.0 && x < 0.0)) {
    x = -x;
    y = -y;
  }
  static_assert(kIdxB == 16, "boundary tables are for 16 buckets");
  // boundary directions (cos, sin) of j pi / 16, j = 0..15
  constexpr double bc[16] = {0x1.0000000000000p+0, 0x1.f6297cff75cb0p-1, 0x1.d906bcf328d46p-1, 0x1.a9b66290ea1a3p-1, 0x1.6a09e667f3bcdp-1, 0x1.1c73b39ae68c9p-1, 0x1.87de2a6aea964p-2, 0x1.8f8b83c69a60dp-3, 0x1.1a62633145c07p-54, -0x1.8f8b83c69a608p-3, -0x1.87de2a6aea962p-2, -0x1.1c73b39ae68c6p-1, -0x1.6a09e667f3bccp-1, -0x1.a9b66290ea1a4p-1, -0x1.d906bcf328d46p-1, -0x1.f6297cff75cb0p-1};
  constexpr double bs[16] = {0x0.0p+0, 0x1.8f8b83c69a60ap-3, 0x1.87de2a6aea963p-2, 0x1.1c73b39ae68c8p-1, 0x1.6a09e667f3bccp-1, 0x1.a9b66290ea1a3p-1, 0x1.d906bcf328d46p-1, 0x1.f6297cff75cb0p-1, 0x1.0000000000000p+0, 0x1.f6297cff75cb0p-1, 0x1.d906bcf328d46p-1, 0x1.a9b66290ea1a5p-1, 0x1.6a09e667f3bcdp-1, 0x1.1c73b39ae68c8p-1, 0x1.87de2a6aea965p-2, 0x1.8f8b83c69a617p-3};
  int b = 0;
#pragma unroll
  for (int step = 8; step > 0; step >>= 1) {
    const int j = b + step;
    if (bc[j] * y - bs[j] * x >= 0.0) b = j;  // psi at or past boundary j
  }
  return b;
}

// Buckets met by the widened arc of edge a -> b: first bucket and count
// (count kIdxB = all).
__device__ __forceinline__ void edge_buckets(V3 a, V3 b, V3 u, V3 v, int& b0, int& nb) {
  const double ua = dot(u, a), va = dot(v, a), ub = dot(u, b), vb = dot(v, b);
  const double Ra = sqrt(ua * ua + va * va), Rb = sqrt(ub * ub + vb * vb);
  const double Rm = fmin(Ra, Rb);
  if (!(Rm > 1e-6)) {
    b0 = 0;
    nb = kIdxB;
    return;
  }
  const double pa = atan2(va, ua), pb = atan2(vb, ub);
  double d = pb - pa;  // to (-pi, pi]
  if (d > kPiD) d -= 2.0 * kPiD;
  if (d <= -kPiD) d += 2.0 * kPiD;
  const double w = asin(fmin(1.0, 2.0 * kEps / Rm)) + 1e-7;
  const double len = fabs(d) + 2.0 * w;
  if (len >= kPiD - 1e-6) {
    b0 = 0;
    nb = kIdxB;
    return;
  }
  double start = (d >= 0.0 ? pa : pb) + 0.5 * kPiD - w;
  start = fmod(start, kPiD);
  if (start < 0.0) start += kPiD;
  const double scale = kIdxB / kPiD;
  const int s = min(max((int)(start * scale), 0), kIdxB - 1);
  const int e = (int)((start + len) * scale);  // may pass kIdxB: wraps
  b0 = s;
  nb = min(e - s + 1, kIdxB);
}

// Index sizes: one wave per cell; isize[cell] = kIdxB + 1 offsets plus the
// candidate entries of a leaf with >= kIdxMin records, 0 otherwise.
// Leaves to index (>= kIdxMin records): flag per cell; every cell starts
// with no index (ib = -1, isize = 0).
__global__ void k_index_flag(const int32_t* __restrict__ hdr, int cnum, int32_t* __restrict__ flag,
                             int32_t* __restrict__ isize, int32_t* __restrict__ ib) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= cnum) return;
  flag[cell] = hdr[2 * cell + 1] >= kIdxMin ? 1 : 0;
  isize[cell] = 0;
  ib[cell] = -1;
}

__global__ void k_index_list(const int32_t* __restrict__ flag, const int32_t* __restrict__ pos, int cnum,
                             int32_t* __restrict__ list) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell < cnum && flag[cell]) list[pos[cell]] = cell;
}

// One wave per listed leaf.
__global__ void __launch_bounds__(256, 6) k_index_size(const double* __restrict__ qc, const int32_t* __restrict__ hdr,
                                                    const double* __restrict__ sv, const uint8_t* __restrict__ sf,
                                                    const int32_t* __restrict__ list, int nlist,
                                                    int32_t* __restrict__ isize, uint16_t* __restrict__ code) {
  const int j = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= nlist) return;
  const int cell = list[j];
  const int r0 = hdr[2 * cell], m = hdr[2 * cell + 1];
  V3 u, v;
  leaf_basis(load3(qc + 3 * (int64_t)cell), u, v);
  int tot = 0;
  for (int base = 0; base < m - 1; base += 64) {
    const int k = base + lane;
    int nb = 0;
    if (k < m - 1) {
      int b0 = 0;
      if (sf[r0 + k] & 4)
        edge_buckets(load3(sv + 3 * (int64_t)(r0 + k)), load3(sv + 3 * (int64_t)(r0 + k + 1)), u, v, b0, nb);
      code[r0 + k] = (uint16_t)(b0 | (nb << 8));  // the buckets again for k_index_fill
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nb += __shfl_xor(nb, o, 64);
    tot += nb;
  }
  if (lane == 0) isize[cell] = kIdxB + 1 + tot;
}

// Per polygon over its bottom cells (grid: chunks x polygons, as
// k_leaf_poly_sums): pstat[2p] += records of its indexed leaves,
// pstat[2p+1] += their index entries (to size the query's lane groups).
__global__ void __launch_bounds__(256) k_index_poly_sums(const PolyDev* __restrict__ pd,
                                                         const int32_t* __restrict__ hdr,
                                                         const int32_t* __restrict__ isize,
                                                         unsigned long long* __restrict__ pstat) {
  __shared__ unsigned long long part[2][4];
  const PolyDev& P = pd[blockIdx.y];
  const int i0 = blockIdx.x * kSumChunk, i1 = min(P.bcn, i0 + kSumChunk);
  unsigned long long sm = 0, st = 0;
  for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
    const int cell = P.bcs + i;
    const int sz = isize[cell];
    if (sz > 0) {
      sm += (unsigned long long)hdr[2 * cell + 1];
      st += (unsigned long long)(sz - (kIdxB + 1));
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sm += __shfl_xor(sm, o, 64);
    st += __shfl_xor(st, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    part[0][threadIdx.x >> 6] = sm;
    part[1][threadIdx.x >> 6] = st;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long a = part[0][0] + part[0][1] + part[0][2] + part[0][3];
    const unsigned long long b = part[1][0] + part[1][1] + part[1][2] + part[1][3];
    if (a) atomicAdd(&pstat[2 * blockIdx.y], a);
    if (b) atomicAdd(&pstat[2 * blockIdx.y + 1], b);
  }
}

// Index contents at ioff[cell] (leaves sized by k_index_size): kIdxB + 1
// offsets (relative to the entries), then per bucket its records in
// ascending order; ib[cell] = ioff[cell].  The buckets of each record are
// the codes k_index_size left (first bucket | count << 8).
__global__ void __launch_bounds__(256) k_index_fill(const int32_t* __restrict__ hdr,
                                                    const uint16_t* __restrict__ code,
                                                    const int32_t* __restrict__ list, int nlist,
                                                    const int32_t* __restrict__ ioff, int32_t* __restrict__ ib,
                                                    int32_t* __restrict__ idx) {
  const int j = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= nlist) return;
  const int cell = list[j];
  const int r0 = hdr[2 * cell], m = hdr[2 * cell + 1];
  const int base = ioff[cell];
  int32_t* off = idx + base;
  int32_t* ent = off + kIdxB + 1;
  const unsigned long long below = (1ull << lane) - 1ull;
  // pass 1: entries per bucket
  int cnt[kIdxB];
#pragma unroll
  for (int b = 0; b < kIdxB; b++) cnt[b] = 0;
  for (int c0 = 0; c0 < m - 1; c0 += 64) {
    const int k = c0 + lane;
    const int cb = k < m - 1 ? code[r0 + k] : 0;
    const int b0 = cb & 0xff, nb = cb >> 8;
#pragma unroll
    for (int b = 0; b < kIdxB; b++) {
      const bool in = nb > 0 && ((b - b0 + kIdxB) % kIdxB) < nb;
      cnt[b] += __popcll(__ballot(in));
    }
  }
  int start[kIdxB];
  int acc = 0;
#pragma unroll
  for (int b = 0; b < kIdxB; b++) {
    start[b] = acc;
    acc += cnt[b];
  }
  if (lane == 0) {
    for (int b = 0; b < kIdxB; b++) off[b] = start[b];
    off[kIdxB] = acc;
    ib[cell] = base;
  }
  // pass 2: entries, in record order within each bucket
  for (int c0 = 0; c0 < m - 1; c0 += 64) {
    const int k = c0 + lane;
    const int cb = k < m - 1 ? code[r0 + k] : 0;
    const int b0 = cb & 0xff, nb = cb >> 8;
#pragma unroll
    for (int b = 0; b < kIdxB; b++) {
      const bool in = nb > 0 && ((b - b0 + kIdxB) % kIdxB) < nb;
      const unsigned long long bal = __ballot(in);
      if (in) ent[start[b] + __popcll(bal & below)] = k;
      start[b] += __popcll(bal);
    }
  }
}

// Edge tests of a wave's boundary points, by groups of G lanes: each group
// resolves the points of its own lanes one at a time, its G lanes splitting
// the leaf's edge stream (coalesced loads) or, for an indexed leaf, the
// records of the point's psi bucket, and the crossing count is the popcount
// of the group's part of a ballot.  Every condition below is uniform within
// a group, so a group's lanes stay active together.
__device__ __forceinline__ bool record_cross(int rk, V3 nq, V3 pp, const double* __restrict__ sv,
                                             const double* __restrict__ sn, const uint8_t* __restrict__ sf) {
  const int fl = sf[rk];
  if (!(fl & 4)) return false;
  const double ra = dot(nq, load3(sv + 3 * (int64_t)rk));
  const double rb = dot(nq, load3(sv + 3 * (int64_t)(rk + 1)));
  if (same_side(ra, rb)) return false;
  const double r3 = dot(load3(sn + 3 * (int64_t)rk), pp);
  return !(((fl & 1) && lt0(r3)) || ((fl & 2) && gt0(r3)));
}

template <int G>
__device__ __forceinline__ void leaf_tests(int lane, int leaf, V3 p, int type, int64_t q,
                                           const double* __restrict__ qc, const int32_t* __restrict__ hdr,
                                           const int32_t* __restrict__ ib, const int32_t* __restrict__ idx,
                                           const double* __restrict__ sv, const double* __restrict__ sn,
                                           const uint8_t* __restrict__ sf, int32_t* __restrict__ labels) {
  const int gbase = lane & ~(G - 1);
  const unsigned long long gmask = G == 64 ? ~0ull : (((1ull << G) - 1ull) << gbase);
  // each lane first gathers its own leaf's centre, stream range and index
  // bucket (all lanes at once), so that the group's loop below only fetches
  // index entries and records for each leaf in turn
  V3 nq_own = v3(0.0, 0.0, 0.0);
  int r0_own = 0, m_own = 0, e0_own = 0, e1_own = 0, base_own = -1;
  if (leaf >= 0) {
    const V3 c = load3(qc + 3 * (int64_t)leaf);
    nq_own = arc_normal(c, p);
    r0_own = hdr[2 * leaf];
    m_own = hdr[2 * leaf + 1];
    base_own = ib ? ib[leaf] : -1;
    if (base_own >= 0) {
      V3 u, v;
      leaf_basis(c, u, v);
      const int b = psi_bucket(nq_own, u, v);
      e0_own = idx[base_own + b];
      e1_own = idx[base_own + b + 1];
    }
  }
  unsigned long long mine = __ballot(leaf >= 0) & gmask;
  while (mine) {
    const int src = __ffsll((long long)mine) - 1;
    mine &= mine - 1;
    const V3 pp = v3(__shfl(p.x, src, 64), __shfl(p.y, src, 64), __shfl(p.z, src, 64));
    const V3 nq = v3(__shfl(nq_own.x, src, 64), __shfl(nq_own.y, src, 64), __shfl(nq_own.z, src, 64));
    const int r0 = __shfl(r0_own, src, 64), m = __shfl(m_own, src, 64);
    const int base = __shfl(base_own, src, 64);
    int cnt = 0;
    if (base < 0) {
      for (int b0 = 0; b0 < m - 1; b0 += G) {
        const int k = b0 + (lane - gbase);
        const bool cross = k < m - 1 && record_cross(r0 + k, nq, pp, sv, sn, sf);
        cnt += __popcll(__ballot(cross) & gmask);
      }
    } else {
      const int e0 = __shfl(e0_own, src, 64), e1 = __shfl(e1_own, src, 64);
      const int* ent = idx + base + kIdxB + 1;
      for (int i0 = e0; i0 < e1; i0 += G) {
        const int i = i0 + (lane - gbase);
        const bool cross = i < e1 && record_cross(r0 + ent[i], nq, pp, sv, sn, sf);
        cnt += __popcll(__ballot(cross) & gmask);
      }
    }
    if (lane == src) labels[q] = (cnt & 1) ? 3 - type : type;
  }
}

// Table-accelerated query; labels identical to k_query_soa.
// Each lane locates its own point (table start cell, then the reference's
// first-match descent).  Points that end in a leaf with refs are then
// resolved cooperatively, one at a time per wave: the 64 lanes split that
// leaf's edge stream (coalesced loads), each lane tests one edge, and the
// crossing count is the popcount of a ballot.
// TRACE adds one record of 6 int32 per point: {leaf cell (-1 = none),
// its top-level cell, its level, level of the table start cell (-1 = level-0
// scan, -2 = resolved by the table), border planes evaluated, leaf edge
// records scanned}; labels are unchanged.
// The points of one wave: lane l takes position t (the wave's 64
// consecutive positions), point order[t] (points grouped by polygon and
// start bin, so that a wave's gathers share cache lines) or point t.
template <bool TRACE>
__device__ __forceinline__ void query_lanes(int64_t t, int64_t n, const GridView& g, const PolyDev* __restrict__ pd,
                                            const int32_t* __restrict__ table, const int32_t* __restrict__ sub,
                                            int B, const double* __restrict__ qc, const int32_t* __restrict__ hdr,
                                            const int32_t* __restrict__ ib, const int32_t* __restrict__ idx,
                                            const double* __restrict__ sv, const double* __restrict__ sn,
                                            const uint8_t* __restrict__ sf, const double* __restrict__ xyz,
                                            const int32_t* __restrict__ qpoly, const int32_t* __restrict__ order,
                                            int32_t* __restrict__ labels, const LevelTab& lt,
                                            int32_t* __restrict__ trace, int G_override,
                                            const int32_t* __restrict__ nodes) {
  const int64_t q = (order && t < n) ? (int64_t)order[t] : t;
  const int lane = threadIdx.x & 63;
  V3 p = v3(0.0, 0.0, 0.0);
  int leaf = -1;  // bottom-cell index still to resolve, or -1
  int type = kUndefined;
  int tests = 0, start_level = -1;
  if (t < n) {
    p = load3(xyz + 3 * q);
    const PolyDev& P = pd[qpoly ? qpoly[q] : 0];
    int cur = -1;
    bool done = false;
    if (B > 0 && P.tbins > 0) {
      double fx = 0.0, fy = 0.0;
      const int bin = bin_of(p, P.tbins, fx, fy);
      int e = bin >= 0 ? table[(int64_t)P.tbl + bin] : -1;
      if (e <= -kSubCode) {
        const int S = P.tsub;
        const int si = min(max((int)(fx * S), 0), S - 1), sj = min(max((int)(fy * S), 0), S - 1);
        e = sub[(int64_t)(-e - kSubCode) + si * S + sj];
      }
      if (e <= -2) {
        labels[q] = -e - 2;
        done = true;
      }
      if (P.trel && e >= 0) e += P.cell0;  // the shared level-0 lookup
      cur = e;
      if (TRACE) start_level = e <= -2 ? -2 : (e >= 0 ? level_of(lt, e) : -1);
    }
    if (!done && cur < 0) {
      for (int i = 0; i < kGrid0Cells; i++) {
        if (nodes) {
          int sub0 = 0, meta0 = 0;
          cur = node_inside<TRACE>(nodes, P.cell0 + i, p, sub0, meta0, tests) ? P.cell0 + i : -1;
        } else {
          cur = TRACE ? q_inside_n(g, p, P.cell0 + i, tests) : q_inside(g, p, P.cell0 + i);
        }
        if (cur != -1) break;
      }
      if (cur == -1) {
        labels[q] = kUndefined;
        done = true;
      }
    }
    if (!done) {
      int meta = 0;
      bool refs = false;
      if (nodes) {
        cur = descend_nodes<TRACE>(nodes, p, cur, meta, tests);
        type = (meta >> 16) & 0xff;
        refs = (meta >> 24) != 0;
      } else {
        cur = descend<TRACE>(g, p, cur, tests);
        if (cur != -1) {
          type = g.c_type[cur];
          refs = g.c_rnum[cur] != 0;
        }
      }
      if (cur == -1) {
        type = kUndefined;
        labels[q] = kUndefined;
      } else if (!refs) {
        labels[q] = type;
      } else {
        leaf = cur;
      }
    }
    if (TRACE) {
      const int lc = locate_leaf(g, p, P.cell0);
      int32_t* tr = trace + 6 * q;
      tr[0] = lc;
      tr[1] = lc >= 0 ? g.c_root[lc] : -1;
      tr[2] = lc >= 0 ? level_of(lt, lc) : -1;
      tr[3] = start_level;
      tr[4] = tests;
      tr[5] = leaf >= 0 ? hdr[2 * leaf + 1] : 0;
    }
  }
  // lanes per edge-test group: the smallest of the wave's polygons' widths
  // (points arrive grouped by polygon, so a wave rarely spans two); the
  // width decides only how the work is split, never a label
  int gw = 64;
  if (t < n) gw = G_override > 0 ? G_override : pd[qpoly ? qpoly[q] : 0].qg;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) gw = min(gw, __shfl_xor(gw, o, 64));
  switch (gw) {
    case 1: leaf_tests<1>(lane, leaf, p, type, q, qc, hdr, ib, idx, sv, sn, sf, labels); break;
    case 2: leaf_tests<2>(lane, leaf, p, type, q, qc, hdr, ib, idx, sv, sn, sf, labels); break;
    case 4: leaf_tests<4>(lane, leaf, p, type, q, qc, hdr, ib, idx, sv, sn, sf, labels); break;
    case 8: leaf_tests<8>(lane, leaf, p, type, q, qc, hdr, ib, idx, sv, sn, sf, labels); break;
    case 16: leaf_tests<16>(lane, leaf, p, type, q, qc, hdr, ib, idx, sv, sn, sf, labels); break;
    case 32: leaf_tests<32>(lane, leaf, p, type, q, qc, hdr, ib, idx, sv, sn, sf, labels); break;
    default: leaf_tests<64>(lane, leaf, p, type, q, qc, hdr, ib, idx, sv, sn, sf, labels); break;
  }
}

// Queue counters of the persistent launch: one per XCD, 256 bytes apart.
constexpr int kQueueStride = 32;  // uint64 entries
constexpr int kQueueWords = 8 * kQueueStride;

__global__ void k_zero_queue(unsigned long long* queue) {
  if (threadIdx.x < kQueueWords) queue[threadIdx.x] = 0ull;
}

// Thread t of the launch takes position t; with xcd > 0 each XCD takes
// runs of xcd consecutive blocks (xcd_block).  With d_n the count is *d_n
// (at most n, the launch bound): blocks past it leave at once.
// 7 waves per SIMD (72 VGPRs): the walk is bound by the latency of its
// dependent gathers; the registers that holding a node record's six planes
// at once needs are worth more than an eighth wave (8 waves: 2 VGPRs
// spilled, fractal_1m's pass 8.52-8.61 ms against 8.36-8.45 ms).
template <bool TRACE>
__global__ void __launch_bounds__(kQBlock, 7) k_query(GridView g, const PolyDev* __restrict__ pd,
                                                   const int32_t* __restrict__ table,
                                                   const int32_t* __restrict__ sub, int B,
                                                   const double* __restrict__ qc, const int32_t* __restrict__ hdr,
                                                   const int32_t* __restrict__ ib, const int32_t* __restrict__ idx,
                                                   const double* __restrict__ sv, const double* __restrict__ sn,
                                                   const uint8_t* __restrict__ sf, const double* __restrict__ xyz,
                                                   const int32_t* __restrict__ qpoly,
                                                   const int32_t* __restrict__ order, int64_t n,
                                                   const int32_t* __restrict__ d_n,
                                                   int32_t* __restrict__ labels, LevelTab lt,
                                                   int32_t* __restrict__ trace, int G_override, int xcd,
                                                   const int32_t* __restrict__ nodes) {
  if (d_n) n = *d_n;
  const int64_t blk = xcd_block(xcd);
  if (blk * blockDim.x >= n) return;
  query_lanes<TRACE>(blk * blockDim.x + threadIdx.x, n, g, pd, table, sub, B, qc, hdr, ib, idx, sv, sn, sf, xyz,
                     qpoly, order, labels, lt, trace, G_override, nodes);
}

// The persistent launch (queue counters cleared before it): a grid of
// resident blocks whose waves take tiles of 64 positions from queues.  The
// positions are cut into the runs of the launch above (xcd blocks' worth of
// tiles, run r belonging to XCD r mod 8); queue x hands out XCD x's runs
// tile by tile in order; a wave takes from its own XCD's queue (blockIdx
// mod 8: workgroups reach the XCDs round-robin) and, once that is empty,
// from the others in turn.  Every tile is taken exactly once, so labels do
// not depend on the order.
__global__ void __launch_bounds__(kQBlock) __attribute__((amdgpu_waves_per_eu(7, 8)))
k_query_persist(GridView g, const PolyDev* __restrict__ pd, const int32_t* __restrict__ table,
                const int32_t* __restrict__ sub, int B, const double* __restrict__ qc,
                const int32_t* __restrict__ hdr, const int32_t* __restrict__ ib, const int32_t* __restrict__ idx,
                const double* __restrict__ sv, const double* __restrict__ sn, const uint8_t* __restrict__ sf,
                const double* __restrict__ xyz, const int32_t* __restrict__ qpoly, const int32_t* __restrict__ order,
                int64_t n, const int32_t* __restrict__ d_n, int32_t* __restrict__ labels, LevelTab lt, int G_override,
                int xcd, const int32_t* __restrict__ nodes, unsigned long long* __restrict__ queue) {
  if (d_n) n = *d_n;
  const int lane = threadIdx.x & 63;
  const int64_t tiles = (n + 63) / 64;
  const int64_t K = xcd > 0 ? (int64_t)xcd * (blockDim.x / 64) : 1;  // tiles per run
  const int home = blockIdx.x & 7;
  for (int v = 0; v < 8; v++) {
    const int x = (home + v) & 7;
    for (;;) {
      unsigned long long k = 0;
      if (lane == 0) k = atomicAdd(queue + (int64_t)x * kQueueStride, 1ull);
      k = ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(k >> 32)) << 32) |
          (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)k);
      const int64_t run = (int64_t)k / K, j = (int64_t)k - run * K;
      const int64_t tile = (run * 8 + x) * K + j;
      if (tile >= tiles) break;  // tiles of queue x increase with k: x is empty
      query_lanes<false>(tile * 64 + lane, n, g, pd, table, sub, B, qc, hdr, ib, idx, sv, sn, sf, xyz, qpoly,
                         order, labels, lt, nullptr, G_override, nodes);
    }
  }
}

// Sub-tables for the bins whose level-1 entry is not final.  Polygon p gets
// S_p x S_p sub-bins per such bin: the smallest power of two that makes a
// sub-bin no wider than a cell of p's bottom level (level-0 cells span about
// 0.55 rad, halving per level), capped by sub_cap (query_sub, or the side of
// the polygon's start-table candidate) and by a budget of sub_budget entries
// per edge of p (S_p = 1: no sub-tables).
static int build_sub_tables(gahhg_ctx* c, int64_t per, int64_t nb, int sub_cap, int sub_budget) {
  QueryLayout& q = c->q;
  const int np = (int)c->polys.size();
  const int B = q.bins;
  const int64_t need = 5 * (nb + 1) + np;
  if (use_buf(c, q.b_tmp, need * 4, (void**)&q.tmp)) return GAHHG_E_NOMEM;
  q.tmpcap = need;
  int32_t* flag = q.tmp;
  int32_t* scan = flag + (nb + 1);
  int32_t* size = scan + (nb + 1);
  int32_t* off = size + (nb + 1);
  int32_t* list = off + (nb + 1);
  int32_t* cnt = list + (nb + 1);
  GridView g = view(c);
  const unsigned nblk = (unsigned)((nb + 255) / 256);
  hipLaunchKernelGGL(k_sub_flag, dim3(nblk), dim3(256), 0, c->stream, g, (const int32_t*)q.table, nb, flag);
  int rc = hip_check(c, hipGetLastError(), "sub flags");
  if (!rc) rc = scan_i32(c, flag, scan, nb, nullptr, nullptr, 0);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sub_counts, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, np, per,
                     (const int32_t*)flag, (const int32_t*)scan, cnt);
  std::vector<int32_t> K((size_t)np);
  rc = hip_check(c, hipGetLastError(), "sub counts");
  if (!rc) rc = hip_check(c, hipMemcpyAsync(K.data(), cnt, np * 4, hipMemcpyDeviceToHost, c->stream), "sub D2H");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sub sync");
  if (rc) return rc;
  const double bin_angle = 1.5707963267948966 / B;
  int64_t total = 0;
  for (int p = 0; p < np; p++) {
    PolyDev& P = c->pd[p];
    const double cell = 0.55 / (double)(1LL << std::min(P.L, 40));
    const double budget = (double)sub_budget * P.N + 4096.0;
    int S = 1;
    while (S < sub_cap && bin_angle / S > cell) S *= 2;
    while (S > 1 && (double)K[p] * S * S > budget) S /= 2;
    P.tsub = S;
    if (S > 1) total += (int64_t)K[p] * S * S;
  }
  if (total + kSubCode > INT32_MAX) return set_err(c, GAHHG_E_LIMIT, "sub-tables exceed int32 indexing");
  rc = hip_check(c, hipMemcpyAsync(c->d_pd, c->pd.data(), c->pd.size() * sizeof(PolyDev), hipMemcpyHostToDevice,
                                   c->stream),
                 "polygon upload");
  if (rc || total == 0) return rc;
  hipLaunchKernelGGL(k_sub_size, dim3(nblk), dim3(256), 0, c->stream, (const PolyDev*)c->d_pd, per, nb,
                     (const int32_t*)flag, size, flag);
  rc = hip_check(c, hipGetLastError(), "sub sizes");
  if (!rc) rc = scan_i32(c, size, off, nb, nullptr, nullptr, 0);
  if (!rc) rc = scan_i32(c, flag, scan, nb, nullptr, nullptr, 0);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sub_list, dim3(nblk), dim3(256), 0, c->stream, nb, (const int32_t*)flag,
                     (const int32_t*)scan, list);
  rc = hip_check(c, hipGetLastError(), "sub list");
  if (rc) return rc;
  int64_t nbin = 0;
  for (int p = 0; p < np; p++) nbin += c->pd[p].tsub > 1 ? K[p] : 0;
  if (use_buf(c, q.b_sub, total * 4, (void**)&q.sub)) return GAHHG_E_NOMEM;
  q.nsub = total;
  q.bytes += total * 4;
  hipLaunchKernelGGL(k_sub_build, dim3((unsigned)nbin), dim3(64), 0, c->stream, g, (const int32_t*)q.nodes,
                     (const PolyDev*)c->d_pd, B, per,
                     (const int32_t*)list, (const int32_t*)q.table, (const int32_t*)off, q.sub);
  rc = hip_check(c, hipGetLastError(), "sub build");
  if (rc) return rc;
  hipLaunchKernelGGL(k_sub_link, dim3(nblk), dim3(256), 0, c->stream, nb, (const int32_t*)flag, (const int32_t*)off,
                     q.table);
  return hip_check(c, hipGetLastError(), "sub link");
}

// Grid of the per-polygon sum kernels: chunks of kSumChunk bottom cells x
// polygons.
static dim3 sum_grid(const gahhg_ctx* c) {
  int chunks = 1;
  for (const PolyDev& P : c->pd) chunks = std::max(chunks, (P.bcn + kSumChunk - 1) / kSumChunk);
  return dim3((unsigned)chunks, (unsigned)c->pd.size());
}

// Angular indexes of the leaves with kIdxMin or more records (see above).
static int build_leaf_index(gahhg_ctx* c) {
  QueryLayout& q = c->q;
  const int64_t nc = c->tot_c;
  if (nc <= 0 || c->prm.query_index == 0) return GAHHG_OK;
  const int np = (int)c->polys.size();
  int32_t* scratch = nullptr;
  // sizes, offsets, total (+pad), per-polygon sums, flags, positions, count,
  // list, bucket codes (one uint16 per record)
  const int64_t words = 2 * nc + 2 + 4 * (int64_t)np + 3 * nc + 2 + ((int64_t)q.nrec + 1) / 2 + 1;
  if (use_buf(c, q.b_ib, nc * 4, (void**)&q.ib) || use_buf(c, q.b_iscr, words * 4, (void**)&scratch))
    return GAHHG_E_NOMEM;
  int32_t* isize = scratch;
  int32_t* ioff = scratch + nc;
  int32_t* total = scratch + 2 * nc;
  unsigned long long* pstat = (unsigned long long*)(scratch + 2 * nc + 2);
  int32_t* flag = scratch + 2 * nc + 2 + 4 * np;
  int32_t* pos = flag + nc;
  int32_t* nlist_d = pos + nc;
  int32_t* list = nlist_d + 2;
  uint16_t* code = (uint16_t*)(list + nc);
  const unsigned cb = (unsigned)((nc + 255) / 256);
  hipLaunchKernelGGL(k_index_flag, dim3(cb), dim3(256), 0, c->stream, (const int32_t*)q.hdr, (int)nc, flag, isize,
                     q.ib);
  int rc = hip_check(c, hipGetLastError(), "index flags");
  if (!rc) rc = scan_i32(c, flag, pos, nc, nullptr, nlist_d, 0);
  int32_t nlist = 0;
  if (!rc) {
    hipLaunchKernelGGL(k_index_list, dim3(cb), dim3(256), 0, c->stream, (const int32_t*)flag, (const int32_t*)pos,
                       (int)nc, list);
    rc = hip_check(c, hipMemcpyAsync(&nlist, nlist_d, 4, hipMemcpyDeviceToHost, c->stream), "index list");
  }
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "index list");
  const unsigned lb = (unsigned)(((int64_t)std::max(nlist, 1) * 64 + 255) / 256);
  if (!rc && nlist > 0) {
    hipLaunchKernelGGL(k_index_size, dim3(lb), dim3(256), 0, c->stream, (const double*)q.qc, (const int32_t*)q.hdr,
                       (const double*)q.sv, (const uint8_t*)q.sf, (const int32_t*)list, nlist, isize, code);
    rc = hip_check(c, hipGetLastError(), "index size");
  }
  if (!rc) rc = fill_i32(c, (int32_t*)pstat, 4 * (int64_t)np, 0);
  if (!rc) {
    hipLaunchKernelGGL(k_index_poly_sums, sum_grid(c), dim3(256), 0, c->stream, (const PolyDev*)c->d_pd,
                       (const int32_t*)q.hdr, (const int32_t*)isize, pstat);
    rc = hip_check(c, hipGetLastError(), "index sums");
  }
  if (!rc) rc = scan_i32(c, isize, ioff, nc, nullptr, total, 0);
  int32_t h = 0;
  std::vector<unsigned long long> ps(2 * (size_t)np);
  if (!rc) rc = hip_check(c, hipMemcpyAsync(&h, total, 4, hipMemcpyDeviceToHost, c->stream), "index size");
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(ps.data(), pstat, 16 * (size_t)np, hipMemcpyDeviceToHost, c->stream),
                   "index size");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "index size");
  if (!rc) {
    // lanes per edge-test group, from the records a point tests on average
    // (the whole stream of an unindexed leaf, one bucket of an indexed one):
    // about half of them without an index (measured), an eighth with one,
    // where more groups resolving points side by side pay more than fewer
    // chunks per point (fractal_1m, 100M points: G = 4, 8, 16, 32 give
    // 9.0, 9.5, 10.7, 13.7 ms)
    std::vector<int32_t> pc((size_t)np * 2);
    rc = hip_check(c, hipMemcpy(pc.data(), q.pcnt, np * 8, hipMemcpyDeviceToHost), "leaf counts");
    for (int p = 0; !rc && p < np; p++) {
      const double leaves = pc[2 * p + 1];
      const double tested = (double)pc[2 * p] - (double)ps[2 * p] + (double)ps[2 * p + 1] / kIdxB;
      const double mean = leaves > 0 ? tested / leaves : 1.0;
      const double share = pc[2 * p] > 0 ? (double)ps[2 * p] / pc[2 * p] : 0.0;  // records in indexed leaves
      const double per_lane = 2.0 + 6.0 * share;
      int Gp = 4;
      while (Gp < 64 && Gp < mean / per_lane) Gp *= 2;
      c->pd[p].qg = Gp;
      if (c->prm.verbose && np == 1)
        printf("Leaf index: %d of %d leaves indexed, %lld records in indexed leaves, %lld entries; %.1f records "
               "tested per point on average, %d lanes per group.\n",
               nlist, pc[2 * p + 1], (long long)ps[2 * p], (long long)ps[2 * p + 1], mean, Gp);
    }
    if (!rc)
      rc = hip_check(c, hipMemcpyAsync(c->d_pd, c->pd.data(), c->pd.size() * sizeof(PolyDev), hipMemcpyHostToDevice,
                                       c->stream),
                     "polygon upload");
  }
  if (!rc && use_buf(c, q.b_idx, std::max<int64_t>(h, 1) * 4, (void**)&q.idx)) rc = GAHHG_E_NOMEM;
  if (rc) return rc;
  q.nidx = h;
  q.bytes += nc * 4 + (int64_t)h * 4;
  if (nlist > 0)
    hipLaunchKernelGGL(k_index_fill, dim3(lb), dim3(256), 0, c->stream, (const int32_t*)q.hdr,
                       (const uint16_t*)code, (const int32_t*)list, nlist, (const int32_t*)ioff, q.ib, q.idx);
  return hip_check(c, hipGetLastError(), "index fill");
}

int pack_query_layout(gahhg_ctx* c) {
  QueryLayout& q = c->q;
  c->qepoch++;
  const int np = (int)c->polys.size();
  // start tables: the requested resolution for one polygon; batches keep
  // each polygon's table to at most 128 bins per face side (393 KB) and
  // build no sub-tables
  if (c->prm.query_layout == 1) return GAHHG_OK;  // the plain path queries the build arrays only
  q.bins = np == 1 ? c->prm.query_bins : std::min(c->prm.query_bins, 128);
  const int64_t nc = std::max<int64_t>(c->tot_c, 1);
  GridView g = view(c);
  // record counts -> offsets (flag scratch holds >= tot_c entries)
  // lanes per cell from the mean refs of the bottom cells
  int64_t refs = 0, bottom = 0;
  for (const PolyDev& P : c->pd) {
    refs += P.R;
    bottom += P.bcn;
  }
  int G = 1;
  while (G < 64 && G < (bottom > 0 ? 0.25 * refs / bottom : 1.0)) G *= 2;
  const unsigned wb = (unsigned)((nc * G + 255) / 256);
  if (use_buf(c, q.b_pcnt, (int64_t)np * 8, (void**)&q.pcnt)) return GAHHG_E_NOMEM;
  hipLaunchKernelGGL(k_leaf_count, dim3(wb), dim3(256), 0, c->stream, g, (const PolyDev*)c->d_pd,
                     (const int32_t*)c->d_pe, (int)c->tot_c, c->d_f0, G);
  int rc = fill_i32(c, q.pcnt, 2 * (int64_t)np, 0);
  if (rc) return rc;
  hipLaunchKernelGGL(k_leaf_poly_sums, sum_grid(c), dim3(256), 0, c->stream, (const PolyDev*)c->d_pd,
                     (const int32_t*)c->d_f0, q.pcnt);
  rc = hip_check(c, hipGetLastError(), "leaf count");
  if (rc) return rc;
  rc = scan_i32(c, c->d_f0, c->d_o0, c->tot_c, nullptr, &c->d_st->nrec, 0);
  if (rc) return rc;
  // node records for the descent (one 192-byte record per cell), checked
  // with the readback below
  if (use_buf(c, q.b_nodes, nc * kNodeInts * 4, (void**)&q.nodes)) return GAHHG_E_NOMEM;
  q.bytes_nodes = nc * kNodeInts * 4;
  rc = fill_i32(c, &c->d_st->bad_nodes, 1, 0);
  if (rc) return rc;
  hipLaunchKernelGGL(k_pack_nodes, dim3((unsigned)((c->tot_c + 255) / 256)), dim3(256), 0, c->stream, g,
                     (int64_t)c->tot_c, q.nodes, c->d_st);
  rc = hip_check(c, hipGetLastError(), "node records");
  if (rc) return rc;
  rc = hip_check(c, hipMemcpyAsync(c->h_st, c->d_st, sizeof(DevState), hipMemcpyDeviceToHost, c->stream), "readback");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  if (c->h_st->bad_nodes) return set_err(c, GAHHG_E_FORMAT, "query layout: a cell has more than 6 borders");
  q.nrec = std::max<int32_t>(c->h_st->nrec, 1);
  // lanes per edge-test group of each polygon's queries: the smallest power
  // of two >= half its mean edge-stream records per leaf, in [4, 64]
  {
    std::vector<int32_t> pc((size_t)np * 2);
    rc = hip_check(c, hipMemcpy(pc.data(), q.pcnt, np * 8, hipMemcpyDeviceToHost), "leaf counts");
    if (rc) return rc;
    q.leaves = 0;
    for (int p = 0; p < np; p++) q.leaves += pc[2 * p + 1];
    for (int p = 0; p < np; p++) {
      const double mean = pc[2 * p + 1] > 0 ? (double)pc[2 * p] / pc[2 * p + 1] : 1.0;
      int Gp = 4;
      while (Gp < 64 && Gp < 0.5 * mean) Gp *= 2;
      c->pd[p].qg = Gp;
    }
    rc = hip_check(c, hipMemcpyAsync(c->d_pd, c->pd.data(), c->pd.size() * sizeof(PolyDev), hipMemcpyHostToDevice,
                                     c->stream),
                   "polygon upload");
    if (rc) return rc;
  }
  if (use_buf(c, q.b_qc, nc * 24, (void**)&q.qc) || use_buf(c, q.b_hdr, nc * 8, (void**)&q.hdr) ||
      use_buf(c, q.b_sv, q.nrec * 24, (void**)&q.sv) || use_buf(c, q.b_sn, q.nrec * 24, (void**)&q.sn) ||
      use_buf(c, q.b_sf, q.nrec, (void**)&q.sf))
    return GAHHG_E_NOMEM;
  q.bytes = nc * 32 + q.nrec * 49;
  hipLaunchKernelGGL(k_leaf_write, dim3(wb), dim3(256), 0, c->stream, g, (const PolyDev*)c->d_pd,
                     (const int32_t*)c->d_pe, (const double*)c->d_edge, (int)c->tot_c,
                     (const int32_t*)c->d_o0, q.qc, q.hdr, q.sv, q.sn, q.sf, G);
  rc = hip_check(c, hipGetLastError(), "leaf write");
  if (rc) return rc;
  q.bytes += q.bytes_nodes;
  rc = build_leaf_index(c);
  if (rc) return rc;
  // start-table candidate (bins per face side, sub-table side) per polygon:
  // fixed by the caller, else chosen by the learned table model when one is
  // set, else (-1) the size rule below.  Every candidate gives the same
  // labels: a table entry is only a cell that every point of its bin reaches.
  std::vector<int32_t> cand((size_t)np, -1);
  if ((int)c->table_fixed.size() == np) cand = c->table_fixed;
  bool open = false;
  for (int p = 0; p < np; p++) open |= cand[p] < 0;
  if (open && c->learn.table_on && c->prm.query_bins > 0) {
    std::vector<int32_t> chosen((size_t)np, 0);
    rc = table_choose(c, chosen.data());
    if (rc) return rc;
    for (int p = 0; p < np; p++)
      if (cand[p] < 0) cand[p] = chosen[p];
  }
  int sub_cap = c->prm.query_sub, sub_budget = c->prm.query_sub_budget;
  if (np == 1 && cand[0] >= 0) {
    q.bins = kTableB[cand[0]];
    sub_cap = kTableS[cand[0]];
    sub_budget = std::max(1, sub_cap / 2);
  }
  c->table_used.assign((size_t)np * 2, 0);
  if (q.bins > 0) {
    // bins per face side: the requested ones for one polygon; in a batch
    // each polygon gets at most query_table_budget entries per edge (at
    // least 16 and at most q.bins bins per side), so that many small
    // polygons do not each pay for a whole-sphere table (624 polygons of
    // ~500 edges: 30 ms of table build at 128 bins)
    const int64_t tb = c->prm.query_table_budget;
    std::vector<int32_t> ts((size_t)np + 1, 0);
    int nrel = 0;
    for (int p = 0; p < np; p++) {
      int Bp = q.bins;
      if (np > 1 && cand[p] >= 0) {
        Bp = std::min(kTableB[cand[p]], q.bins);
      } else if (np > 1 && tb > 0) {
        Bp = 16;
        while (2 * Bp <= q.bins && 6LL * (2 * Bp) * (2 * Bp) <= tb * c->pd[p].N) Bp *= 2;
        Bp = std::min(Bp, q.bins);
      }
      const int64_t next = (int64_t)ts[p] + 6LL * Bp * Bp;
      if (next > INT32_MAX - 6LL * kL0Bins * kL0Bins)
        return set_err(c, GAHHG_E_LIMIT, "start tables exceed int32 indexing");
      c->pd[p].tbl = ts[p];
      c->pd[p].tbins = Bp;
      c->pd[p].tsub = 1;
      c->pd[p].trel = 0;
      nrel += Bp == 0;
      ts[p + 1] = (int32_t)next;
    }
    const int64_t nb = ts[np];
    // polygons of a batch without a table of their own share one level-0
    // lookup appended to the tables (kL0Bins per face side, entries relative
    // to the polygon's first level-0 cell), in place of scanning up to 42
    // level-0 cells per point
    const int64_t nl0r = nb > 0 && nrel > 0 ? 6LL * kL0Bins * kL0Bins : 0;
    for (int p = 0; p < np && nl0r > 0; p++)
      if (c->pd[p].tbins == 0) c->pd[p].tbl = (int32_t)nb, c->pd[p].tbins = kL0Bins, c->pd[p].trel = 1;
    if (nb == 0) {  // no polygon gets a table: the query starts at level 0
      q.bins = 0;
      for (int p = 0; p < np; p++) c->pd[p].tbl = -1, c->pd[p].tbins = 0, c->pd[p].tsub = 1, c->pd[p].trel = 0;
      rc = hip_check(c, hipMemcpyAsync(c->d_pd, c->pd.data(), c->pd.size() * sizeof(PolyDev), hipMemcpyHostToDevice,
                                       c->stream),
                     "polygon upload");
      if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "polygon upload");
      if (!rc)
        for (int p = 0; p < np; p++) c->table_used[2 * p] = 0, c->table_used[2 * p + 1] = 1;
      return rc;
    }
    if (use_buf(c, q.b_table, (nb + nl0r) * 4, (void**)&q.table) ||
        use_buf(c, q.b_tstart, ((int64_t)np + 1) * 4, (void**)&q.tstart))
      return GAHHG_E_NOMEM;
    q.bytes += (nb + nl0r) * 4;
    rc = hip_check(c, hipMemcpyAsync(q.tstart, ts.data(), ((size_t)np + 1) * 4, hipMemcpyHostToDevice, c->stream),
                   "table starts");
    if (!rc)
      rc = hip_check(c, hipMemcpyAsync(c->d_pd, c->pd.data(), c->pd.size() * sizeof(PolyDev), hipMemcpyHostToDevice,
                                       c->stream),
                     "polygon upload");
    if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "table starts");  // ts leaves scope
    if (rc) return rc;
    const int64_t per = 6LL * q.bins * q.bins;
    // tables of 16, 32, ... bins per side: the level-0 part of their entries
    // comes from b_l0, computed once per resolution and shared by the
    // polygons and by later builds (level-0 geometry never changes)
    int32_t* l0 = nullptr;
    int Bmax = 0;
    bool pow2 = true;
    uint32_t need = 0;
    for (int p = 0; p < np && pow2; p++) {
      const int B = c->pd[p].tbins;
      if (B == 0 || c->pd[p].trel) continue;  // no table of its own
      pow2 = B >= 16 && B <= 4096 && (B & (B - 1)) == 0;
      if (pow2) need |= 1u << (31 - __builtin_clz((unsigned)(B / 16)));
      Bmax = std::max(Bmax, B);
    }
    pow2 = pow2 && Bmax > 0;
    if (pow2) {
      const int64_t nl0 = l0_offset(2 * Bmax);
      if (q.b_l0.p == nullptr || std::max<int64_t>(nl0 * 4, 256) > q.b_l0.cap) q.l0_have = 0;  // reallocated
      if (use_buf(c, q.b_l0, nl0 * 4, (void**)&l0)) return GAHHG_E_NOMEM;
      for (int k = 0; (16 << k) <= Bmax; k++) {
        if (!((need >> k) & 1u) || ((q.l0_have >> k) & 1u)) continue;
        const int B = 16 << k;
        const int64_t n = 6LL * B * B;
        hipLaunchKernelGGL(k_level0_table, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, g,
                           (const int32_t*)q.nodes, c->pd[0].cell0, B, B, l0 + l0_offset(B));
        rc = hip_check(c, hipGetLastError(), "level-0 table");
        if (rc) return rc;
        q.l0_have |= 1u << k;
      }
    }
    hipLaunchKernelGGL(k_build_table, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, c->stream, g,
                       (const int32_t*)q.nodes,
                       (const PolyDev*)c->d_pd, np, (const int32_t*)q.tstart, (const int32_t*)l0, q.table);
    rc = hip_check(c, hipGetLastError(), "table build");
    if (!rc && nl0r > 0) {
      hipLaunchKernelGGL(k_level0_table, dim3((unsigned)((nl0r + 255) / 256)), dim3(256), 0, c->stream, g,
                         (const int32_t*)q.nodes, c->pd[0].cell0, kL0Bins, kL0Bins, q.table + nb);
      rc = hip_check(c, hipGetLastError(), "level-0 lookup");
    }
    // sub-tables for one polygon only: on the 256-polygon batch they cost
    // 11.5 ms of build for 0.9 ms less per 64M points
    if (!rc && sub_cap > 1 && np == 1 && nb > 0) rc = build_sub_tables(c, per, nb, sub_cap, sub_budget);
  }
  if (!rc && q.bins == 0)
    for (int p = 0; p < np; p++) c->pd[p].tbins = 0, c->pd[p].tsub = 1, c->pd[p].trel = 0;  // no table this time
  if (!rc)
    for (int p = 0; p < np; p++)
      c->table_used[2 * p] = c->pd[p].trel ? 0 : c->pd[p].tbins, c->table_used[2 * p + 1] = c->pd[p].tsub;
  return rc;
}
// Spreads the low 16 bits of v to the even bit positions.
__device__ __forceinline__ uint32_t spread_bits(uint32_t v) {
  v &= 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// Sort key of a point: its polygon, then its cube-map face, then the
// Z-order (Morton) code of its face coordinates quantised to r bits per
// axis, so that points close on the sphere get close keys and a wave's
// lanes walk the same cells and leaves; value = index.
__global__ void k_sort_key(const double* __restrict__ xyz, const int32_t* __restrict__ qpoly, int64_t n, int r,
                           uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const V3 p = load3(xyz + 3 * t);
  double fx = 0.0, fy = 0.0;
  const int R = 1 << r;
  const int cell = bin_of(p, R, fx, fy);  // (face * R + i) * R + j
  uint32_t code = 0;
  if (cell >= 0) {
    const uint32_t face = (uint32_t)(cell >> (2 * r));
    const uint32_t i = (uint32_t)((cell >> r) & (R - 1)), j = (uint32_t)(cell & (R - 1));
    code = (face << (2 * r)) | (spread_bits(i) << 1) | spread_bits(j);
  }
  const uint32_t poly = qpoly ? (uint32_t)qpoly[t] : 0u;
  key[t] = (poly << (3 + 2 * r)) | code;
  val[t] = (uint32_t)t;
}

static int ceil_log2(int64_t v) {
  int b = 0;
  while ((1LL << b) < v) b++;
  return b;
}

// The grouping key of k_sort_key for one point.
__device__ __forceinline__ uint32_t group_key(V3 p, uint32_t poly, int r) {
  double fx = 0.0, fy = 0.0;
  const int R = 1 << r;
  const int cell = bin_of(p, R, fx, fy);
  uint32_t code = 0;
  if (cell >= 0) {
    const uint32_t face = (uint32_t)(cell >> (2 * r));
    const uint32_t i = (uint32_t)((cell >> r) & (R - 1)), j = (uint32_t)(cell & (R - 1));
    code = (face << (2 * r)) | (spread_bits(i) << 1) | spread_bits(j);
  }
  return (poly << (3 + 2 * r)) | code;
}

// First pass of a grouped query: the start table answers the points it
// resolves (label written here); every other point gets its grouping key,
// the rest kNoKey (no real key has it: the face field never exceeds 5).
// counts[block] = the block's points left for the walk.
constexpr uint32_t kNoKey = 0xffffffffu;

// Each block takes kTableSub consecutive 256-point tiles (counts per tile,
// as the compaction reads them), with every tile's loads in flight
// together: a tile's chain (point, table entry, sub-table entry) is a few
// dependent loads, and four chains per thread hide each other's latency.
constexpr int kTableSub = 4;
__global__ void __launch_bounds__(256) k_query_table(const PolyDev* __restrict__ pd, const int32_t* __restrict__ table,
                                                     const int32_t* __restrict__ sub, int B,
                                                     const double* __restrict__ xyz,
                                                     const int32_t* __restrict__ qpoly, int64_t n, int r,
                                                     int32_t* __restrict__ labels, uint32_t* __restrict__ key,
                                                     int32_t* __restrict__ counts, int64_t nblk) {
  __shared__ int part[kTableSub][4];
  V3 p[kTableSub];
  uint32_t poly[kTableSub];
  int64_t t[kTableSub];
  bool ok[kTableSub];
#pragma unroll
  for (int j = 0; j < kTableSub; j++) {
    t[j] = ((int64_t)blockIdx.x * kTableSub + j) * 256 + threadIdx.x;
    ok[j] = t[j] < n;
    p[j] = ok[j] ? load3(xyz + 3 * t[j]) : v3(1.0, 0.0, 0.0);
    poly[j] = (ok[j] && qpoly) ? (uint32_t)qpoly[t[j]] : 0u;
  }
  int e[kTableSub];
  double fx[kTableSub], fy[kTableSub];
#pragma unroll
  for (int j = 0; j < kTableSub; j++) {
    const PolyDev& P = pd[poly[j]];
    fx[j] = fy[j] = 0.0;
    // the shared level-0 lookup never answers a point: left to the walk
    const int bin = P.tbins > 0 && !P.trel ? bin_of(p[j], P.tbins, fx[j], fy[j]) : -1;
    e[j] = (ok[j] && bin >= 0) ? table[(int64_t)P.tbl + bin] : -1;
  }
#pragma unroll
  for (int j = 0; j < kTableSub; j++) {
    if (e[j] <= -kSubCode) {
      const int S = pd[poly[j]].tsub;
      const int si = min(max((int)(fx[j] * S), 0), S - 1), sj = min(max((int)(fy[j] * S), 0), S - 1);
      e[j] = sub[(int64_t)(-e[j] - kSubCode) + si * S + sj];
    }
  }
#pragma unroll
  for (int j = 0; j < kTableSub; j++) {
    bool keep = false;
    if (ok[j]) {
      // every point's label slot is written, so that the stores fill whole
      // lines; the walk overwrites the placeholder of each kept point
      keep = e[j] > -2;
      labels[t[j]] = keep ? kUndefined : -e[j] - 2;
      key[t[j]] = keep ? group_key(p[j], poly[j], r) : kNoKey;
    }
    const int c = __popcll(__ballot(keep));
    if ((threadIdx.x & 63) == 0) part[j][threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x < kTableSub) {
    const int64_t sb = (int64_t)blockIdx.x * kTableSub + threadIdx.x;
    if (sb < nblk) counts[sb] = part[threadIdx.x][0] + part[threadIdx.x][1] + part[threadIdx.x][2] + part[threadIdx.x][3];
  }
}

// Second pass: the kept points of table-pass block sb, in point order, to
// [offs[sb], ...).  Each block takes 4 table-pass blocks at once (their
// loads in flight together; 4x fewer blocks).
constexpr int kCompactSub = 4;
__global__ void __launch_bounds__(256) k_query_compact(const uint32_t* __restrict__ key, int64_t n, int64_t nblk,
                                                       const int32_t* __restrict__ offs, uint32_t* __restrict__ kout,
                                                       uint32_t* __restrict__ vout) {
  __shared__ int part[kCompactSub][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t k[kCompactSub];
  unsigned long long m[kCompactSub];
#pragma unroll
  for (int j = 0; j < kCompactSub; j++) {
    const int64_t sb = (int64_t)blockIdx.x * kCompactSub + j;
    const int64_t t = sb * 256 + threadIdx.x;
    k[j] = (sb < nblk && t < n) ? key[t] : kNoKey;
  }
#pragma unroll
  for (int j = 0; j < kCompactSub; j++) {
    m[j] = __ballot(k[j] != kNoKey);
    if (lane == 0) part[j][wid] = __popcll(m[j]);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kCompactSub; j++) {
    const int64_t sb = (int64_t)blockIdx.x * kCompactSub + j;
    if (sb >= nblk || k[j] == kNoKey) continue;
    int before = offs[sb];
    for (int w = 0; w < wid; w++) before += part[j][w];
    const int pos = before + __popcll(m[j] & ((1ull << lane) - 1ull));
    kout[pos] = k[j];
    vout[pos] = (uint32_t)(sb * 256 + threadIdx.x);
  }
}

// Bytes of the block counts, offsets and total of a grouping of cap points.
static int64_t count_bytes(int64_t cap) { return ((8 * ((cap + 255) / 256) + 4) + 255) / 256 * 256; }

// Grows the grouping buffer to n points (contents not kept).
static int sort_buffers(gahhg_ctx* c, int64_t n) {
  QuerySort& s = c->qsort;
  if (s.cap >= n) return GAHHG_OK;
  (void)hipStreamSynchronize(c->stream);
  dfree(s.buf);
  s.buf = nullptr;
  s.cap = 0;
  c->qepoch++;
  if (dalloc(c, (void**)&s.buf, count_bytes(n) + 16 * n + 4 * radix_hist_entries(n))) return GAHHG_E_NOMEM;
  s.cap = n;
  return GAHHG_OK;
}

// Groups the points by (polygon, position on the sphere) with the device
// radix sort of gahhg_sort.hip over the key bits in use.  With a start
// table, the points it resolves are answered first (labels written) and
// only the rest are compacted (block counts, a scan, a stable scatter) and
// sorted; their number stays on the device (the sort reads it there).  With
// host_count it is also copied to pinned memory as soon as the compaction
// has written it, and read after the sort is queued, so the GPU does not
// idle while the host waits.  Buffers are kept by the context and grow.
int64_t sort_buffer_bytes(int64_t cap) { return count_bytes(cap) + 16 * cap + 4 * radix_hist_entries(cap); }

int sort_queries(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels,
                 const int32_t** order, int64_t* m, const int32_t** d_count, bool host_count,
                 const SortScratch* ext) {
  if (n > INT32_MAX) return set_err(c, GAHHG_E_LIMIT, "sort: more than 2^31 points");
  // face coordinates at r bits per axis, r as large as the 32-bit key
  // leaves room for after the polygon and face bits (at most query_sort_bits)
  const int poly_bits = ceil_log2((int64_t)c->polys.size());
  const int r = std::max(1, std::min(c->prm.query_sort_bits, (32 - 3 - poly_bits) / 2));
  const int end_bit = std::min(32, 3 + 2 * r + poly_bits);
  const int64_t nb = (n + 255) / 256;  // blocks of the table passes
  int rc = GAHHG_OK;
  SortScratch s;
  if (ext) {
    if (ext->cap < n) return set_err(c, GAHHG_E_LIMIT, "sort: %lld points exceed the scratch's %lld", (long long)n,
                                     (long long)ext->cap);
    s = *ext;
  } else {
    rc = sort_buffers(c, n);
    if (rc) return rc;
    s.buf = c->qsort.buf;
    s.cap = c->qsort.cap;
  }
  // buffer: [block counts | block offsets | total] [k0 | k1 | v0 | v1] [radix histograms]
  const int64_t cnt_cap = count_bytes(s.cap);
  int32_t* bcount = (int32_t*)s.buf;
  int32_t* boffs = bcount + nb;
  int32_t* btotal = boffs + nb;
  uint32_t* k0 = (uint32_t*)((char*)s.buf + cnt_cap);
  uint32_t* k1 = k0 + s.cap;
  uint32_t* v0 = k1 + s.cap;
  uint32_t* v1 = v0 + s.cap;
  int32_t* hist = (int32_t*)(v1 + s.cap);
  const unsigned blocks = (unsigned)nb;
  if (c->q.table) {
    const QueryLayout& L = c->q;
    hipLaunchKernelGGL(k_query_table, dim3((unsigned)((nb + kTableSub - 1) / kTableSub)), dim3(256), 0, c->stream,
                       (const PolyDev*)c->d_pd, (const int32_t*)L.table, (const int32_t*)L.sub, L.bins, d_xyz, d_poly,
                       n, r, d_labels, k1, bcount, nb);
    rc = hip_check(c, hipGetLastError(), "table pass");
    if (rc) return rc;
    rc = scan_i32_with(c, bcount, boffs, nb, nullptr, btotal, 0, s.scan);
    if (rc) return rc;
    hipLaunchKernelGGL(k_query_compact, dim3((unsigned)((nb + kCompactSub - 1) / kCompactSub)), dim3(256), 0,
                       c->stream, (const uint32_t*)k1, n, nb, (const int32_t*)boffs, k0, v0);
    rc = hip_check(c, hipGetLastError(), "compaction");
    if (rc) return rc;
  } else {
    hipLaunchKernelGGL(k_sort_key, dim3(blocks), dim3(256), 0, c->stream, d_xyz, d_poly, n, r, k0, v0);
    rc = hip_check(c, hipGetLastError(), "sort keys");
    if (!rc) rc = fill_i32(c, btotal, 1, (int32_t)n);
    if (rc) return rc;
  }
  if (host_count) {
    rc = hip_check(c, hipMemcpyAsync(c->h_cnt, btotal, 4, hipMemcpyDeviceToHost, c->stream), "point count");
    if (!rc) rc = hip_check(c, hipEventRecord(c->ev_cnt, c->stream), "point count");
    if (rc) return rc;
  }
  uint32_t *kres = nullptr, *vres = nullptr;
  rc = radix_sort_pairs(c, k0, v0, k1, v1, btotal, n, end_bit, hist, &kres, &vres, s.scan);
  if (rc) return rc;
  *order = (const int32_t*)vres;
  *d_count = btotal;
  *m = -1;
  if (host_count) {
    rc = hip_check(c, hipEventSynchronize(c->ev_cnt), "point count");
    if (rc) return rc;
    *m = *c->h_cnt;
  }
  return GAHHG_OK;
}

// Scratch a grouped pass over n points needs, allocated now (a captured
// pass must neither allocate nor synchronise).
static int ensure_queue(gahhg_ctx* c);

static int reserve_query(gahhg_ctx* c, int64_t n) {
  const int64_t nb = (n + 255) / 256;
  if (c->prm.query_persistent) {
    const int qrc = ensure_queue(c);
    if (qrc) return qrc;
  }
  const int rc = scan_reserve(c, std::max<int64_t>(nb, radix_hist_entries(n)));
  return rc ? rc : sort_buffers(c, n);
}

// The persistent launch's queue counters and its grid (the blocks of
// k_query<false> resident at once on the device), set up once per context.
static int ensure_queue(gahhg_ctx* c) {
  if (c->q_queue) return GAHHG_OK;
  int per_cu = 0, cus = 0;
  int rc = hip_check(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_query_persist, c->prm.block_query, 0),
                     "query occupancy");
  if (!rc) rc = hip_check(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device), "CUs");
  if (rc) return rc;
  if (dalloc(c, (void**)&c->q_queue, kQueueWords * 8)) return GAHHG_E_NOMEM;
  c->q_resident = std::max(1, per_cu) * std::max(1, cus);
  c->qepoch++;
  return GAHHG_OK;
}

// allow_persistent: the persistent launch may be used (its queue exists);
// the allocation-free entry point passes false and keeps the static one.
template <bool TRACE>
static void launch_query(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, const int32_t* d_order,
                         int64_t n, const int32_t* d_n, int32_t* d_labels, int32_t* d_trace,
                         bool allow_persistent = false) {
  unsigned blocks = (unsigned)((n + c->prm.block_query - 1) / c->prm.block_query);
  const QueryLayout& L = c->q;
  unsigned long long* queue = nullptr;
  if (!TRACE && allow_persistent && c->prm.query_persistent && c->q_queue) {
    queue = c->q_queue;
    hipLaunchKernelGGL(k_zero_queue, dim3(1), dim3(kQueueWords), 0, c->stream, queue);
    blocks = std::min<unsigned>(blocks, (unsigned)c->q_resident);
  }
  if (queue)
    hipLaunchKernelGGL(k_query_persist, dim3(blocks), dim3(c->prm.block_query), 0, c->stream, view(c),
                       (const PolyDev*)c->d_pd, (const int32_t*)L.table, (const int32_t*)L.sub,
                       L.table ? L.bins : 0, (const double*)L.qc, (const int32_t*)L.hdr, (const int32_t*)L.ib,
                       (const int32_t*)L.idx, (const double*)L.sv, (const double*)L.sn, (const uint8_t*)L.sf, d_xyz,
                       d_poly, d_order, n, d_n, d_labels, level_tab(c), c->prm.query_group, c->prm.query_xcd,
                       (const int32_t*)L.nodes, queue);
  else
    hipLaunchKernelGGL((k_query<TRACE>), dim3(blocks), dim3(c->prm.block_query), 0, c->stream, view(c),
                       (const PolyDev*)c->d_pd, (const int32_t*)L.table, (const int32_t*)L.sub,
                       L.table ? L.bins : 0, (const double*)L.qc, (const int32_t*)L.hdr, (const int32_t*)L.ib,
                       (const int32_t*)L.idx, (const double*)L.sv,
                       (const double*)L.sn, (const uint8_t*)L.sf, d_xyz, d_poly, d_order, n, d_n, d_labels,
                       level_tab(c), d_trace, c->prm.query_group, c->prm.query_xcd, (const int32_t*)L.nodes);
}

// The work of one accelerated query pass, enqueued with no host wait when
// host_count is false (then it can be recorded into a graph).
static int enqueue_query(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels,
                         bool host_count, const SortScratch* ext = nullptr) {
  const int32_t* order = nullptr;
  const int32_t* d_m = nullptr;
  int64_t m = n;
  if (c->prm.query_sort == 2 || (c->prm.query_sort == 1 && n >= kSortMin)) {
    const int rc = sort_queries(c, d_xyz, d_poly, n, d_labels, &order, &m, &d_m, host_count, ext);
    if (rc) return rc;
  }
  if (host_count) (void)hipEventRecord(c->ev[2], c->stream);
  const bool pers = ext == nullptr;  // the allocation-free path keeps the static launch
  if (m < 0)
    launch_query<false>(c, d_xyz, d_poly, order, n, d_m, d_labels, nullptr, pers);
  else if (m > 0)
    launch_query<false>(c, d_xyz, d_poly, order, m, nullptr, d_labels, nullptr, pers);
  if (host_count) (void)hipEventRecord(c->ev[3], c->stream);
  return hip_check(c, hipGetLastError(), "query kernel");
}

void drop_query_graph(gahhg_ctx* c) {
  QueryGraph& G = c->qgraph;
  if (G.exec) (void)hipGraphExecDestroy(G.exec);
  if (G.graph) (void)hipGraphDestroy(G.graph);
  G = QueryGraph();
}

// query_graph = 1: the pass is recorded once per (points, polygon ids,
// labels, count) and buffer epoch, then replayed with one launch.
static int graph_query(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels) {
  QueryGraph& G = c->qgraph;
  if (!(G.exec && G.xyz == d_xyz && G.poly == d_poly && G.labels == d_labels && G.n == n && G.epoch == c->qepoch)) {
    drop_query_graph(c);
    int rc = reserve_query(c, n);
    if (rc) return rc;
    rc = hip_check(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "graph capture");
    if (rc) return rc;
    const int qrc = enqueue_query(c, d_xyz, d_poly, n, d_labels, false);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (qrc) {
      if (graph) (void)hipGraphDestroy(graph);
      return qrc;
    }
    rc = hip_check(c, e, "graph capture");
    if (rc) return rc;
    G.graph = graph;
    rc = hip_check(c, hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0), "graph instantiate");
    if (rc) {
      drop_query_graph(c);
      return rc;
    }
    G.xyz = d_xyz, G.poly = d_poly, G.labels = d_labels, G.n = n, G.epoch = c->qepoch;
  }
  (void)hipEventRecord(c->ev[2], c->stream);
  const int rc = hip_check(c, hipGraphLaunch(G.exec, c->stream), "graph launch");
  (void)hipEventRecord(c->ev[3], c->stream);
  return rc;
}

int run_query_with(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels,
                   const SortScratch* ext) {
  if (!ext) return run_query(c, d_xyz, d_poly, n, d_labels);
  if (n <= 0) return GAHHG_OK;
  if (c->prm.query_layout == 1) return run_query(c, d_xyz, d_poly, n, d_labels);  // uses no scratch
  return enqueue_query(c, d_xyz, d_poly, n, d_labels, true, ext);
}

int run_query(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels) {
  if (n <= 0) return GAHHG_OK;
  // events: ev[2] before the query kernel, ev[3] after it (the sort, when
  // any, runs between the caller's ev[0] and ev[2])
  if (c->prm.query_layout == 1) {
    const unsigned blocks = (unsigned)((n + c->prm.block_query - 1) / c->prm.block_query);
    (void)hipEventRecord(c->ev[2], c->stream);
    hipLaunchKernelGGL(k_query_soa, dim3(blocks), dim3(c->prm.block_query), 0, c->stream, view(c),
                       (const PolyDev*)c->d_pd, (const double*)c->d_edge, d_xyz, d_poly, n, d_labels);
    (void)hipEventRecord(c->ev[3], c->stream);
  } else if (c->prm.query_graph) {
    return graph_query(c, d_xyz, d_poly, n, d_labels);
  } else {
    if (c->prm.query_persistent) {
      const int qrc = ensure_queue(c);
      if (qrc) return qrc;
    }
    return enqueue_query(c, d_xyz, d_poly, n, d_labels, true);
  }
  return hip_check(c, hipGetLastError(), "query kernel");
}

LevelTab level_tab(const gahhg_ctx* c) {
  LevelTab lt;
  memset(&lt, 0, sizeof(lt));
  lt.n = (int32_t)std::min<size_t>(c->glv.size(), kMaxLevels);
  for (int l = 0; l < lt.n; l++) lt.cstart[l] = c->glv[l].cstart;
  lt.cstart[lt.n] = (int32_t)c->tot_c;
  return lt;
}

int run_query_trace(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels,
                    int32_t* d_trace) {
  if (n <= 0) return GAHHG_OK;
  if (!c->q.hdr) return set_err(c, GAHHG_E_STATE, "trace: the accelerated query layout is not built");
  launch_query<true>(c, d_xyz, d_poly, nullptr, n, nullptr, d_labels, d_trace);
  return hip_check(c, hipGetLastError(), "traced query kernel");
}

__global__ void k_check_ids(const int32_t* __restrict__ poly, int64_t n, int npoly, DevState* st) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if ((uint32_t)poly[i] >= (uint32_t)npoly) st->bad_poly = 1;  // any writer: same value
}

__global__ void k_clear_bad(DevState* st) { st->bad_poly = 0; }

int check_poly_ids(gahhg_ctx* c, const int32_t* d_poly, int64_t n) {
  if (!d_poly || n <= 0) return GAHHG_OK;
  hipLaunchKernelGGL(k_clear_bad, dim3(1), dim3(1), 0, c->stream, c->d_st);
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_check_ids, dim3((unsigned)blocks), dim3(256), 0, c->stream, d_poly, n,
                     (int)c->polys.size(), c->d_st);
  int32_t bad = 0;
  int rc = hip_check(c, hipMemcpyAsync(&c->h_st->bad_poly, &c->d_st->bad_poly, 4, hipMemcpyDeviceToHost, c->stream),
                     "polygon id check");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "polygon id check");
  if (rc) return rc;
  bad = c->h_st->bad_poly;
  return bad ? set_err(c, GAHHG_E_ARG, "query: a polygon id is outside 0..%d", (int)c->polys.size() - 1) : GAHHG_OK;
}

__global__ void k_expand_pairs(const double* __restrict__ xyz, int64_t n, int64_t total, double* __restrict__ out,
                               int32_t* __restrict__ poly) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = t / n, i = t - p * n;
    out[3 * t] = xyz[3 * i];
    out[3 * t + 1] = xyz[3 * i + 1];
    out[3 * t + 2] = xyz[3 * i + 2];
    poly[t] = (int32_t)p;
  }
}

int expand_pairs(gahhg_ctx* c, const double* d_xyz, int64_t n, double* d_xyz_out, int32_t* d_poly_out) {
  const int64_t total = n * (int64_t)c->polys.size();
  if (total <= 0) return GAHHG_OK;
  const int64_t blocks = std::min<int64_t>((total + 255) / 256, 16384);
  hipLaunchKernelGGL(k_expand_pairs, dim3((unsigned)blocks), dim3(256), 0, c->stream, d_xyz, n, total, d_xyz_out,
                     d_poly_out);
  return hip_check(c, hipGetLastError(), "pair expansion");
}

__global__ void k_clear_census(DevState* st) {
  st->q_in = 0;
  st->q_undef = 0;
}

int count_labels(gahhg_ctx* c, const int32_t* d_labels, int64_t n, unsigned long long* d_out2) {
  hipLaunchKernelGGL(k_clear_census, dim3(1), dim3(1), 0, c->stream, c->d_st);
  if (n <= 0) return hip_check(c, hipGetLastError(), "census");
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(k_label_counts, dim3((unsigned)blocks), dim3(256), 0, c->stream, d_labels, n, d_out2);
  return hip_check(c, hipGetLastError(), "label counts");
}

// Known-answer kernel: 12 doubles in per case (a, b, c, d), 3 out.
__global__ void k_kat(int op, const double* __restrict__ in, int64_t n, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* x = in + 12 * i;
  const V3 a = load3(x), b = load3(x + 3), cc = load3(x + 6), d = load3(x + 9);
  double* o = out + 3 * i;
  if (op == 0) {
    store3(o, unit(a));
  } else if (op == 1) {
    store3(o, cross(a, b));
  } else if (op == 2) {
    // isEEInter(p0=a, p1=b, v0=c, v1=d)
    const V3 n0 = arc_normal(cc, d);
    const V3 n1 = arc_normal(a, b);
    const bool r = !(same_side(dot(n0, a), dot(n0, b)) || same_side(dot(n1, cc), dot(n1, d)));
    o[0] = r ? 1.0 : 0.0;
    o[1] = 0.0;
    o[2] = 0.0;
  } else {
    // isEdgeBorderInter(p0=a, p1=b, v0=c, v1=d)
    int res = -1;
    const int k = edge_border_class(arc_normal(cc, d), arc_normal(a, b), a, b, cc, d, &res);
    o[0] = (double)k;
    o[1] = (double)res;
    o[2] = 0.0;
  }
}

int run_kat(gahhg_ctx* c, int op, const double* d_in, int64_t n, double* d_out) {
  if (n <= 0) return GAHHG_OK;
  hipLaunchKernelGGL(k_kat, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, op, d_in, n, d_out);
  return hip_check(c, hipGetLastError(), "kat kernel");
}

// Loads this file's code object onto the current device (context set-up).
void load_code_query() {
  hipFuncAttributes a;
  (void)hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_pack_nodes));
}

}  // namespace gahhg
