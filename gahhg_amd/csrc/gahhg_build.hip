// gahhg_build.hip -- hierarchy construction on gfx950.
//
// Follows GAHHG::divGrid / dispatchSELast / GPISP::createAHHG
// (source/GPISPAHHG.cu:1558-1784,1947-2015).  Differences in mechanism, not
// in results:
//  - one host synchronisation per level (the reference makes seven blocking
//    scalar reads per level, :1598,1618,1624,1638,1643,1682,1686); every other
//    count stays in device memory (DevState) and kernels read it there;
//  - every prefix sum is the wave64/LDS scan of gahhg_scan.hip;
//  - the per-edge normal arc_normal(p0, p1) and both endpoints are packed
//    once per build (9 doubles per edge) instead of being re-derived inside
//    every border test; the value is the same double the reference computes
//    inline, so every predicate outcome is unchanged;
//  - border normals stored at creation are reused in the walk (the reference
//    recomputes normalize(v0 x (v1 - v0)) from the same vertices);
//  - refs inside a bottom cell are sorted by edge id (the reference's atomic
//    scatter leaves their order unspecified, :1011-1027).
// Undefined reference behaviour is made explicit and counted: a walk stops at
// a border with no partner (reference reads dev_b[-1]); an edge whose start
// point is not located is dropped from deeper levels; walks stop after
// kStepCap cells.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "gahhg_internal.h"

#pragma clang fp contract(off)

namespace gahhg {

constexpr int kBlock = 256;
// A level's walks use walk records when the steps they are expected to take
// beyond the first cell of each edge number at least this many per cell of
// the level: building a record costs about 0.2 ns per cell and a step read
// from it saves 10 to 80 ps (measured on MI355X: the 1M-vertex fractal ring's
// levels 7-10 have 20 to 80 extra steps per cell and gain 0.6 ms in all; the
// 256-polygon batch has at most 12 and loses 0.15 ms with records).
constexpr double kWalkRecStepsPerCell = 16.0;
// The mark walk keeps the first kSlotK cells of each edge's walk (walk
// slots, device-driven build): the dispatch's write walk of a polygon that
// stopped after marking its bottom level repeats that walk, so the edges
// whose whole walk fits are copied from the slots and only the others are
// walked again.  32 slots (4 B each, per edge) copy all but the longest
// walks: the 256-polygon batch builds 4% faster than with 8, and 64 gain
// nothing more.
constexpr int kSlotK = 32;
constexpr int32_t kSlotDirty = INT32_MIN;  // slot 0 of a walk that raised a diagnostic

static inline unsigned blocks_for(int64_t n, int bs) {
  int64_t b = (n + bs - 1) / bs;
  return (unsigned)std::max<int64_t>(b, 1);
}

// ---------------------------------------------------------------------------
// Small device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ V3 vtx(const GridView& g, int i) { return load3(g.v + 3 * (int64_t)i); }
__device__ __forceinline__ V3 bnorm(const GridView& g, int i) { return load3(g.b_n + 3 * (int64_t)i); }

// device_isPointInsideCell, source/GPISPAHHG.cu:208-224
__device__ __forceinline__ int inside_cell(const GridView& g, V3 p, int cid) {
  const int bs = g.c_bstart[cid];
  const int bn = g.c_bnum[cid];
  for (int i = 0; i < bn; i++) {
    const double* n = g.b_n + 3 * (int64_t)(bs + i);
    double r = n[0] * p.x + n[1] * p.y + n[2] * p.z;
    if (r < -kEps) return -1;
  }
  return cid;
}

// device_getSegExitBorder, source/GPISPAHHG.cu:286-348
__device__ int seg_exit_border(const GridView& g, V3 ne, V3 p0, V3 p1, int cid, int in,
                               int* intertype, int* info) {
  const int bs = g.c_bstart[cid];
  const int be = bs + g.c_bnum[cid];
  for (int i = bs; i < be; i++) {
    if (i == in) continue;
    int temp = 0;
    const int iv0 = g.b_v0[i], iv1 = g.b_v1[i];
    int inter = edge_border_class(bnorm(g, i), ne, p0, p1, vtx(g, iv0), vtx(g, iv1), &temp);
    if (inter == kInterX) {
      *intertype = kInterX;
      return i;
    }
    if (inter == kInterT1) {
      *intertype = kInterT1;
      *info = temp;
      return temp == 0 ? i : -1;
    }
    if (inter == kInterT2) {
      if (in != -1) {
        int vt = temp == 0 ? iv0 : iv1;
        if (vt == g.b_v0[in] || vt == g.b_v1[in]) continue;
      }
      *intertype = kInterT2;
      *info = temp;
      return i;
    }
    if (inter == kInterOverlap) {
      *intertype = kInterOverlap;
      return i;
    }
  }
  return -1;
}

// device_isPointInsideCell on a walk record (the same planes; all of them
// are fetched before any is tested, the result is the same).
__device__ __forceinline__ int inside_rec(const WalkRec* __restrict__ rec, V3 p, int cid) {
  const WalkRec& R = rec[cid];
  const int bn = R.bnum & 0xff;
  bool in = true;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (k < bn) {
      const double* n = R.b[k].n;
      const double r = n[0] * p.x + n[1] * p.y + n[2] * p.z;
      in = in && !(r < -kEps);
    }
  }
  return in ? cid : -1;
}

// seg_exit_border on the record of cell cid (the record's address is known,
// so a border costs one fetch, not the index-then-vertex chain of the grid
// arrays).  Returns the exit border's slot k or -1; *bs_out, *bn_out: the cell's
// border range.
__device__ __forceinline__ int seg_exit_lazy(const GridView& g, const WalkRec* __restrict__ rec, int cid, V3 ne,
                                             V3 p0, V3 p1, int in, int* intertype, int* info, int* bs_out,
                                             int* bn_out) {
  const WalkRec& R = rec[cid];
  const int bs = R.bstart, bn = R.bnum & 0xff;
  *bs_out = bs;
  *bn_out = bn;
  // every plane value of the record first: the loads of all borders are
  // independent and go out together (testing border by border would wait
  // for each border's fetch in turn)
  double r0[6], r1[6], sv[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (k < bn) {
      const WalkBorder& B = R.b[k];
      const V3 nb = V3{B.n[0], B.n[1], B.n[2]};
      r0[k] = dot(nb, p0);
      r1[k] = dot(nb, p1);
      sv[k] = dot(ne, V3{B.v0[0], B.v0[1], B.v0[2]});
    }
  }
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (k >= bn) break;
    const int kn = k + 1 < bn ? k + 1 : 0;
    const double s0 = sv[k], s1 = sv[kn];
    const int i = bs + k;
    if (i == in) continue;
    int temp = 0;
    const int inter = edge_border_class_r(r0[k], r1[k], s0, s1, &temp);
    if (inter == kInterX) {
      *intertype = kInterX;
      return k;
    }
    if (inter == kInterT1) {
      *intertype = kInterT1;
      *info = temp;
      return temp == 0 ? k : -1;
    }
    if (inter == kInterT2) {
      if (in != -1) {  // vertex indices from the grid arrays (this branch is rare)
        const int vt = temp == 0 ? g.b_v0[i] : g.b_v1[i];
        if (vt == g.b_v0[in] || vt == g.b_v1[in]) continue;
      }
      *intertype = kInterT2;
      *info = temp;
      return k;
    }
    if (inter == kInterOverlap) {
      *intertype = kInterOverlap;
      return k;
    }
  }
  return -1;
}

// Walk records of the cells [lo + first, lo + first + kRecBlock) of the
// range [lo, lo + n): each thread fills one record in LDS, then the block
// writes its records out as contiguous 16-byte chunks.  Records are indexed
// by global cell id.
constexpr int kRecBlock = 128;
__device__ __forceinline__ void walk_records_tile(const GridView& g, int lo, int n, int first,
                                                  WalkRec* __restrict__ rec, WalkRec* srec) {
  const int t = first + threadIdx.x;
  if (t < n) {
    const int cid = lo + t;
    WalkRec& R = srec[threadIdx.x];
    const int bs = g.c_bstart[cid], bn = min(g.c_bnum[cid], 6);
    bool open = false;
    for (int k = 0; k < 6; k++) {
      WalkBorder& B = R.b[k];
      if (k < bn) {
        const int i = bs + k;
        const V3 nn = load3(g.b_n + 3 * (int64_t)i);
        const V3 a = vtx(g, g.b_v0[i]);
        const int pt = g.b_partner[i];
        open |= g.b_v1[i] != g.b_v0[bs + (k + 1) % bn];
        B.n[0] = nn.x, B.n[1] = nn.y, B.n[2] = nn.z;
        B.v0[0] = a.x, B.v0[1] = a.y, B.v0[2] = a.z;
        B.partner = pt, B.pparent = pt >= 0 ? g.b_parent[pt] : -1;
      } else {
        B.n[0] = B.n[1] = B.n[2] = 0.0;
        B.v0[0] = B.v0[1] = B.v0[2] = 0.0;
        B.partner = B.pparent = -1;
      }
    }
    R.bstart = bs;
    R.bnum = bn | (open ? kWalkRecOpen : 0);
    R.pad[0] = R.pad[1] = 0;
  }
  __syncthreads();
  const int cnt = min(kRecBlock, n - first);
  const int chunks = cnt * (int)(sizeof(WalkRec) / 16);
  const uint4* src = (const uint4*)srec;
  uint4* dst = (uint4*)(rec + lo + first);
  for (int c = threadIdx.x; c < chunks; c += kRecBlock) dst[c] = src[c];
  __syncthreads();  // srec is refilled by the next tile
}

__global__ void __launch_bounds__(kRecBlock) k_walk_records(GridView g, int lo, int n, WalkRec* __restrict__ rec) {
  __shared__ WalkRec srec[kRecBlock];
  walk_records_tile(g, lo, n, blockIdx.x * kRecBlock, rec, srec);
}

// The same for level lv of the device-driven build, when its walks use
// records (bit lv of st->rec_levels).
__global__ void __launch_bounds__(kRecBlock) k_walk_records_dv(GridView g, const DevState* __restrict__ st, int lv,
                                                               WalkRec* __restrict__ rec) {
  __shared__ WalkRec srec[kRecBlock];
  if (!((st->rec_levels >> lv) & 1u)) return;
  const LevelDesc L = st->lv[lv];
  for (int first = blockIdx.x * kRecBlock; first < L.cnum; first += gridDim.x * kRecBlock)
    walk_records_tile(g, L.cstart, L.cnum, first, rec, srec);
}

struct WalkDiag {
  unsigned long long steps, partner_missing, cap_hits, unlocated, bad_index;
};

enum { kMark = 0, kCount = 1, kWrite = 2 };

// Maps a border index derived by markNECell's OVERLAP branch (computed in
// the polygon's local numbering, :1279-1285) back to a global index, or -1
// when the local index is outside the borders that exist at this level.
__device__ __forceinline__ int overlap_in_index(const GridView& g, const PolyDev& P, int nbs_global, int kk,
                                                WalkDiag& d) {
  const int local = nbs_global - P.bs + P.lbs + kk;
  if (local < 0 || local >= P.lbs + P.bn) {
    d.bad_index++;
    return -1;
  }
  int glob;
  if (local >= P.lbs)
    glob = P.bs + (local - P.lbs);
  else if (local >= P.plbs)
    glob = P.pbs + (local - P.plbs);
  else {
    d.bad_index++;
    return -1;
  }
  return g.b_partner[glob];
}

// One polygon edge across the cells of one level: markNECell (:1153-1289,
// MODE kMark), calcRefNum (:685-818, kCount), writeRef (:820-955, kWrite).
template <int MODE, bool REC>
__device__ void walk_edge(const GridView& g, const WalkRec* __restrict__ rec, const double* __restrict__ edge, int e,
                          int level, const PolyDev& P, int32_t* buf, int32_t* out_cells, WalkDiag& d,
                          int32_t* __restrict__ slot = nullptr, int64_t slot_n = 0) {
  // REC false: this level has no walk records; the steps read the grid
  // arrays (the same values) and the step state stays small
  const double* er = edge + 9 * (int64_t)e;
  const V3 p0 = load3(er), p1 = load3(er + 3), ne = load3(er + 6);
  int cur = -1;
  if (level == 0) {
    for (int i = 0; i < kGrid0Cells; i++) {
      cur = REC ? inside_rec(rec, p0, P.cell0 + i) : inside_cell(g, p0, P.cell0 + i);
      if (cur != -1) break;
    }
    if (cur == -1) {
      d.unlocated++;
      if (MODE == kMark) buf[e] = -1;
      return;
    }
    if (MODE == kMark) buf[e] = cur;
  } else {
    const int par = buf[e];
    if (par < 0) {
      d.unlocated++;
      return;
    }
    const int sc = g.c_subcell[par];
    cur = sc >= 0 ? (REC ? inside_rec(rec, p0, sc) : inside_cell(g, p0, sc)) : -1;
    if (cur == -1) {
      const int pbs = g.c_bstart[par], pbn = g.c_bnum[par];
      for (int i = 0; i < pbn; i++) {
        cur = g.b_subcell[pbs + i];
        if (cur == -1) continue;
        cur = REC ? inside_rec(rec, p0, cur) : inside_cell(g, p0, cur);
        if (cur != -1) break;
      }
    }
    if (cur == -1) {
      d.unlocated++;
      if (MODE == kMark) buf[e] = -1;
      return;
    }
    if (MODE != kCount) buf[e] = cur;
  }

  int in = -1;
  int64_t steps = 0;
  for (;;) {
    if (steps >= kStepCap) {
      d.cap_hits++;
      break;
    }
    if (MODE == kMark) {
      g.c_rnum[cur] = kMarked;
      if (slot && steps < kSlotK) slot[steps * slot_n + e] = cur;
    }
    if (MODE == kWrite) out_cells[steps] = cur;
    steps++;
    int intertype = 0, info = 0;
    int out, kpt = 0, kpp = 0, u;
    if (REC && !(rec[cur].bnum & kWalkRecOpen)) {
      int bs = 0;
      const int k = seg_exit_lazy(g, rec, cur, ne, p0, p1, in, &intertype, &info, &bs, &u);
      if (k == -1) break;
      out = bs + k;
      kpt = rec[cur].b[k].partner;
      kpp = rec[cur].b[k].pparent;
    } else {
      out = seg_exit_border(g, ne, p0, p1, cur, in, &intertype, &info);
      if (out == -1) break;
      kpt = g.b_partner[out];
      kpp = kpt >= 0 ? g.b_parent[kpt] : -1;
      u = g.c_bnum[cur];
    }
    if (intertype == kInterX || intertype == kInterT1) {
      if (kpt == -1) {
        d.partner_missing++;
        break;
      }
      in = kpt;
      cur = kpp;
    } else if (intertype == kInterT2) {
      // step a travel-epsilon past the grid vertex and test the two cells
      // that share it (:753-792)
      V3 cp = vtx(g, info == 0 ? g.b_v0[out] : g.b_v1[out]);
      V3 ed = unit(sub(p1, p0));
      cp = add(cp, mul(ed, kTravelEps));
      const int pt = kpt;
      if (pt != -1) {
        int r = REC ? inside_rec(rec, cp, kpp) : inside_cell(g, cp, kpp);
        if (r != -1) {
          cur = r;
          in = pt;
          continue;
        }
      }
      const int bsc = g.c_bstart[cur];
      const int other = info == 0 ? bsc + (g.b_id[out] + u - 1) % u : bsc + (g.b_id[out] + 1) % u;
      const int po = g.b_partner[other];
      if (po != -1) {
        int r = REC ? inside_rec(rec, cp, g.b_parent[po]) : inside_cell(g, cp, g.b_parent[po]);
        if (r != -1) {
          cur = r;
          in = po;
          continue;
        }
      }
      break;
    } else {  // overlap: continue through the next/previous border (:794-815)
      V3 e0 = unit(sub(p1, p0));
      V3 e1 = unit(sub(vtx(g, g.b_v1[out]), vtx(g, g.b_v0[out])));
      const double t = dot(e0, e1);
      const int u = g.c_bnum[cur];
      const int bs = g.c_bstart[cur];
      const int k = t > 0 ? (out - bs + 1) % u : (out - bs + u - 1) % u;
      const int pt = g.b_partner[bs + k];
      if (pt == -1) {
        d.partner_missing++;
        break;
      }
      const int nc = g.b_parent[pt];
      if (MODE == kMark) {
        // markNECell derives `in` from the new cell's bstart, after curcell
        // has changed (:1279-1285); the index is kept as the reference
        // computes it and treated as -1 outside the array.
        const int nbs = g.c_bstart[nc];
        const int kk = t > 0 ? (out - nbs + 1) % u : (out - nbs + u - 1) % u;
        in = overlap_in_index(g, P, nbs, kk, d);
      } else {
        in = pt;
      }
      cur = nc;
    }
  }
  d.steps += (unsigned long long)steps;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// MODE kMark walks the edges of refining polygons at `level` (and keeps
// each edge's cell count in cnt_out); kCount and kWrite walk every edge at
// its polygon's bottom level, starting from the level above it (buf_prev)
// when refinement stopped by the density rule or the model after marking
// that level (kCount then takes the count the mark walk kept).
// dv != 0 (the device-driven build): the record levels come from
// st->rec_levels, and the launch does nothing once the level loop has
// stopped (mark walk) or the build stops short (dispatch walks).
// Waves per SIMD: the walks wait on memory most of their time, so the
// grid-array variants are held to 6 (80 VGPRs, a few spilled; the 256-polygon
// batch's mark walks 8.25 -> 7.91 ms against 5); the record variants keep 4.
template <int MODE, bool REC>
__global__ void __launch_bounds__(kBlock, REC ? 4 : 6) k_walk(GridView g, const WalkRec* __restrict__ rec, uint32_t rec_levels,
                                                 const double* __restrict__ edge,
                                                 const int32_t* __restrict__ epoly, const PolyDev* __restrict__ pd,
                                                 int Ntot, int level, int32_t* buf, int32_t* buf_prev,
                                                 const int32_t* off, int32_t* cnt_out, int32_t* rcell,
                                                 DevState* st, int dv, int32_t* __restrict__ slot,
                                                 const int32_t* __restrict__ elist) {
  if (dv) {
    if (MODE == kMark ? st->run == 0 : st->abort != 0) return;  // uniform over the launch
    rec_levels = st->rec_levels;
  }
  // a mark walk walks one level: the launch whose record mode is not that
  // level's has nothing to do
  if (MODE == kMark && ((((rec_levels >> level) & 1u) != 0) != REC)) return;
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  WalkDiag d = {0, 0, 0, 0, 0};
  bool mine = e < Ntot;
  if (elist) {  // the edges listed in elist (st->nwalk of them)
    mine = e < st->nwalk;
    if (mine) e = elist[e];
  }
  PolyDev P;
  if (mine) {
    P = pd[epoly[e]];
    // edges walking a level whose records exist go to the REC launch, the
    // others to the plain one
    const int lvl = MODE == kMark ? level : P.L;
    mine = (((rec_levels >> lvl) & 1u) != 0) == REC;
  }
  if (mine) {
    if (MODE == kMark) {
      if (P.refining) {
        buf_prev[e] = buf[e];
        walk_edge<MODE, REC>(g, rec, edge, e, level, P, buf, nullptr, d, slot, Ntot);
        // the count walk of a polygon that stops at this level repeats this
        // walk exactly (same start, same cells): keep its count, unless
        // the walk raised a diagnostic (then the count walk runs and
        // counts it as the reference's dispatch does, and the write walk
        // walks it again)
        const bool clean = d.partner_missing == 0 && d.cap_hits == 0 && d.unlocated == 0 && d.bad_index == 0;
        cnt_out[e] = clean ? (int32_t)d.steps : -1;
        if (slot && !clean) slot[e] = kSlotDirty;
      }
    } else {
      int32_t* b = P.use_prev ? buf_prev : buf;
      if (MODE == kCount && P.use_prev && cnt_out[e] >= 0) {
        d.steps = cnt_out[e];  // walked by the mark walk of the bottom level
      } else {
        int32_t* oc = (MODE == kWrite) ? rcell + off[e] : nullptr;
        walk_edge<MODE, REC>(g, rec, edge, e, P.L, P, b, oc, d);
        if (MODE == kCount) cnt_out[e] = (int32_t)d.steps;
      }
    }
  }
  // counters: summed per wave, then per block, one atomic per block and
  // counter (per-wave atomics on one address serialise across the launch)
  __shared__ unsigned long long part[kBlock / 64][5];
  unsigned long long v[5] = {wave_sum_u64(d.steps), wave_sum_u64(d.partner_missing), wave_sum_u64(d.cap_hits),
                             wave_sum_u64(d.unlocated), wave_sum_u64(d.bad_index)};
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 5; k++) part[threadIdx.x >> 6][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++)
      for (int k = 0; k < 5; k++) v[k] += part[w][k];
    if (MODE == kMark && v[0]) atomicAdd(&st->steps_mark, v[0]);
    if (MODE == kCount && v[0]) atomicAdd(&st->steps_dispatch, v[0]);
    if (MODE != kWrite) {  // the write walk repeats the count walk exactly
      if (v[1]) atomicAdd(&st->partner_missing, v[1]);
      if (v[2]) atomicAdd(&st->cap_hits, v[2]);
      if (v[3]) atomicAdd(&st->unlocated, v[3]);
      if (v[4]) atomicAdd(&st->bad_index, v[4]);
    }
  }
}

// Per-edge record: p0, p1 and arc_normal(p0, p1).
__global__ void k_edges(const double* __restrict__ pv, const int32_t* __restrict__ pe, int N,
                        double* __restrict__ edge) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N) return;
  V3 a = load3(pv + 3 * (int64_t)pe[2 * e]);
  V3 b = load3(pv + 3 * (int64_t)pe[2 * e + 1]);
  double* r = edge + 9 * (int64_t)e;
  store3(r, a);
  store3(r + 3, b);
  store3(r + 6, arc_normal(a, b));
}

// kernel_DivGridFast_zeroRefCount, :1127-1150, over one level's cells of
// every polygon (level 0 also resets the fields the reference resets there).
__device__ __forceinline__ void zero_cell(const GridView& g, int cid, int level0) {
  g.c_rnum[cid] = 0;
  if (level0) {
    g.c_rstart[cid] = -1;
    g.c_subcell[cid] = -1;
    g.c_upcell[cid] = -1;
    g.c_type[cid] = kUndefined;
    const int bs = g.c_bstart[cid], bn = g.c_bnum[cid];
    for (int i = 0; i < bn; i++) g.b_subcell[bs + i] = -1;
  }
}

__global__ void k_zero(GridView g, int lo, int n, int level0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) zero_cell(g, lo + t, level0);
}

// Grid-stride loop over the cells of level lv of the device-driven build.
#define FOR_LEVEL_CELLS(L, t) \
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < (L).cnum; t += gridDim.x * blockDim.x)

__global__ void k_zero_dv(GridView g, const DevState* __restrict__ st, int lv) {
  if (!st->run) return;
  const LevelDesc L = st->lv[lv];
  FOR_LEVEL_CELLS(L, t) zero_cell(g, L.cstart + t, lv == 0 ? 1 : 0);
}

// Kind of the border child produced for border i of refined cell P
// (kernel_DivGrid_calcBorderSCOffsets, :473-509): 6 = hexagon shared with a
// refined neighbour (created by the lower id), 4 = quad against an
// unrefined or missing neighbour, 0 = none.  Refined means marked (rnum != 0)
// at this level, which is exactly the cells createCenter gives a subcell.
__device__ __forceinline__ int child_kind(const GridView& g, int P, int bnum, int i) {
  if (bnum == 4 && i == 0) return 0;
  const int pt = g.b_partner[g.c_bstart[P] + i];
  if (pt == -1) return 4;
  const int Q = g.b_parent[pt];
  const bool qref = g.c_rnum[Q] != 0;
  if (!qref) return 4;
  return P < Q ? 6 : 0;
}

// Plan of one level (kernel_DivGrid_calcCenterOffsets :381-404 and the
// border-child counts): per cell, new corners (= bnum when refined), the
// centre child, border children and their borders.
__device__ __forceinline__ void plan_cell(const GridView& g, int lo, int t, int32_t* fa, int32_t* fb, int32_t* fc,
                                          int32_t* fd) {
  const int cid = lo + t;
  const bool r = g.c_rnum[cid] != 0;
  const int bn = g.c_bnum[cid];
  int nc = 0, nb = 0;
  if (r)
    for (int i = 0; i < bn; i++) {
      const int k = child_kind(g, cid, bn, i);
      nc += k ? 1 : 0;
      nb += k;
    }
  fa[t] = r ? bn : 0;
  fb[t] = r ? 1 : 0;
  fc[t] = nc;
  fd[t] = nb;
}

__global__ void k_plan(GridView g, int lo, int n, int32_t* fa, int32_t* fb, int32_t* fc, int32_t* fd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) plan_cell(g, lo, t, fa, fb, fc, fd);
}

__global__ void k_plan_dv(GridView g, const DevState* __restrict__ st, int lv, int32_t* fa, int32_t* fb, int32_t* fc,
                          int32_t* fd) {
  if (!st->run) return;
  const LevelDesc L = st->lv[lv];
  FOR_LEVEL_CELLS(L, t) plan_cell(g, L.cstart, t, fa, fb, fc, fd);
}

// Per-polygon totals of the plan from the exclusive scans.
// (st non-null: the device-driven build; lo is then its level's first cell)
__global__ void k_plan_totals(PolyDev* pd, int npoly, int lo, const int32_t* fa, const int32_t* fb,
                              const int32_t* fc, const int32_t* fd, const int32_t* oa, const int32_t* ob,
                              const int32_t* oc, const int32_t* od, const DevState* st, int lv) {
  if (st) {
    if (!st->run) return;
    lo = st->lv[lv].cstart;
  }
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npoly) return;
  PolyDev& P = pd[p];
  if (!P.refining || P.cn <= 0) return;
  const int f = P.cs - lo, l = f + P.cn - 1;
  P.nv = oa[l] + fa[l] - oa[f];
  P.U = ob[l] + fb[l] - ob[f];
  P.nbc = oc[l] + fc[l] - oc[f];
  P.nbb = od[l] + fd[l] - od[f];
}

// kernel_DivGrid_createCenter, :406-470
// Lanes per refined cell of the centre-child and border-children kernels
// (a cell has at most 6 borders).
constexpr int kBcLanes = 8;

// Lane i of a refined cell's group (kBcLanes lanes per cell) creates the
// centre child's corner i and border i; it computes the next corner itself
// (the same operations, so the same bits as the stored vertex) for the
// border's normal.  Lane 0 writes the child's cell record.
__device__ __forceinline__ void create_center_lane(const GridView& g, const PolyDev* __restrict__ pd, int lo, int t,
                                                   int i, const int32_t* oa, const int32_t* ob) {
  const int pid = lo + t;
  if (g.c_rnum[pid] == 0) return;
  const int bn = g.c_bnum[pid];
  if (i >= bn) return;
  const int p = g.c_poly[pid];
  const PolyDev& P = pd[p];
  if (!P.refining) return;
  const int f = P.cs - lo;
  const int bs = g.c_bstart[pid];
  V3 center;
  if (bn == 4) {
    center = mul(add(vtx(g, g.b_v0[bs]), vtx(g, g.b_v1[bs])), 0.5);
  } else {
    center = v3(0.0, 0.0, 0.0);
    for (int k = 0; k < bn; k++) center = add(center, vtx(g, g.b_v0[bs + k]));
    center = divk(center, bn);
  }
  center = unit(center);
  const int vo = P.nvb + (oa[t] - oa[f]);
  const int bo = P.nbb0 + (oa[t] - oa[f]);
  const int child = P.ncb + (ob[t] - ob[f]);
  const int i1 = (i + 1) % bn;
  const V3 a = unit(mul(add(vtx(g, g.b_v0[bs + i]), center), 0.5));
  const V3 z = unit(mul(add(vtx(g, g.b_v0[bs + i1]), center), 0.5));
  store3(const_cast<double*>(g.v) + 3 * (int64_t)(vo + i), a);
  const int b = bo + i;
  g.b_id[b] = i;
  g.b_parent[b] = child;
  g.b_partner[b] = -1;
  g.b_subcell[b] = -1;
  g.b_v0[b] = vo + i;
  g.b_v1[b] = vo + i1;
  store3(g.b_n + 3 * (int64_t)b, arc_normal(a, z));
  if (i == 0) {
    g.c_bstart[child] = bo;
    g.c_bnum[child] = bn;
    g.c_rstart[child] = -1;
    g.c_rnum[child] = -1;
    g.c_upcell[child] = pid;
    g.c_subcell[child] = -1;
    g.c_type[child] = kUndefined;
    g.c_poly[child] = p;
    g.c_root[child] = g.c_root[pid];
    g.c_subcell[pid] = child;
  }
}

__global__ void k_create_center(GridView g, const PolyDev* __restrict__ pd, int lo, int n, const int32_t* oa,
                                const int32_t* ob) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int t = (int)(u / kBcLanes);
  if (t < n) create_center_lane(g, pd, lo, t, (int)(u % kBcLanes), oa, ob);
}

__global__ void k_create_center_dv(GridView g, const PolyDev* __restrict__ pd, const DevState* __restrict__ st, int lv,
                                   const int32_t* oa, const int32_t* ob) {
  if (!st->run) return;
  const LevelDesc L = st->lv[lv];
  const int64_t n = (int64_t)L.cnum * kBcLanes;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (int64_t)gridDim.x * blockDim.x)
    create_center_lane(g, pd, L.cstart, (int)(u / kBcLanes), (int)(u % kBcLanes), oa, ob);
}

// kernel_DivGrid_createBorderSC, :512-607: the border children of one
// refined cell, in border order.
// Border children: lane i of a refined cell's group creates the child of
// border i (its slot follows from the kinds of borders 0..i-1), so a cell's
// children are built side by side.
__device__ __forceinline__ void border_child(const GridView& g, const PolyDev* __restrict__ pd, int lo, int t, int i,
                                             const int32_t* oc, const int32_t* od) {
  const int P = lo + t;
  if (g.c_rnum[P] == 0) return;
  const int bnP = g.c_bnum[P];
  if (i >= bnP) return;
  const int kind = child_kind(g, P, bnP, i);
  if (!kind) return;
  const int p = g.c_poly[P];
  const PolyDev& D = pd[p];
  if (!D.refining) return;
  const int f = D.cs - lo;
  int co = D.ncb + D.U + (oc[t] - oc[f]);
  int eo = D.nbb0 + D.nv + (od[t] - od[f]);
  for (int j = 0; j < i; j++) {
    const int k = child_kind(g, P, bnP, j);
    co += k ? 1 : 0;
    eo += k;
  }
  const int pbs = g.c_bstart[P];
  const int cbs = g.c_bstart[g.c_subcell[P]];  // centre child's borders
  const int ue = pbs + i;
  const int b = cbs + i;
  int vv[6];
  if (kind == 6) {
    const int pt = g.b_partner[ue];
    const int psub = g.c_subcell[g.b_parent[pt]];
    const int pse = g.c_bstart[psub] + g.b_id[pt];
    vv[0] = g.b_v0[ue];
    vv[1] = g.b_v1[pse];
    vv[2] = g.b_v0[pse];
    vv[3] = g.b_v1[ue];
    vv[4] = g.b_v1[b];
    vv[5] = g.b_v0[b];
  } else {
    vv[0] = g.b_v0[ue];
    vv[1] = g.b_v1[ue];
    vv[2] = g.b_v1[b];
    vv[3] = g.b_v0[b];
  }
  for (int k = 0; k < kind; k++) {
    const int nb = eo + k;
    const int a = vv[k], z = vv[(k + 1) % kind];
    g.b_id[nb] = k;
    g.b_v0[nb] = a;
    g.b_v1[nb] = z;
    g.b_parent[nb] = co;
    g.b_partner[nb] = -1;
    g.b_subcell[nb] = -1;
    store3(g.b_n + 3 * (int64_t)nb, arc_normal(vtx(g, a), vtx(g, z)));
  }
  g.c_bnum[co] = kind;
  g.c_bstart[co] = eo;
  g.c_rstart[co] = -1;
  g.c_rnum[co] = -1;
  g.c_subcell[co] = -1;
  g.c_upcell[co] = P;
  g.c_type[co] = kUndefined;
  g.c_poly[co] = p;
  g.c_root[co] = g.c_root[P];
  g.b_subcell[ue] = co;
  if (kind == 6) g.b_subcell[g.b_partner[ue]] = co;
}

__global__ void k_border_children(GridView g, const PolyDev* __restrict__ pd, int lo, int n, const int32_t* oc,
                                  const int32_t* od) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int t = (int)(u / kBcLanes);
  if (t < n) border_child(g, pd, lo, t, (int)(u % kBcLanes), oc, od);
}

__global__ void k_border_children_dv(GridView g, const PolyDev* __restrict__ pd, const DevState* __restrict__ st,
                                     int lv, const int32_t* oc, const int32_t* od) {
  if (!st->run) return;
  const LevelDesc L = st->lv[lv];
  const int64_t n = (int64_t)L.cnum * kBcLanes;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (int64_t)gridDim.x * blockDim.x)
    border_child(g, pd, L.cstart, (int)(u / kBcLanes), (int)(u % kBcLanes), oc, od);
}

// device_setBorderPartner, :42-58: link `edge` with the border of `cell`
// running the opposite way; the lower index writes both sides.
__device__ __forceinline__ void link_border(const GridView& g, int cell, int edge) {
  const int bs = g.c_bstart[cell], bn = g.c_bnum[cell];
  const int a = g.b_v0[edge], z = g.b_v1[edge];
  for (int i = 0; i < bn; i++) {
    const int j = bs + i;
    if (a == g.b_v1[j] && z == g.b_v0[j] && edge < j) {
      g.b_partner[edge] = j;
      g.b_partner[j] = edge;
    }
  }
}

__device__ __forceinline__ void link_or_clear(const GridView& g, int te, int edge) {
  const int sc = g.b_subcell[te];
  if (sc != -1)
    link_border(g, sc, edge);
  else
    g.b_partner[edge] = -1;
}

// kernel_DivGrid_setSCPartner, :610-682, over the new level's cells (the
// centre children are skipped).
__device__ __forceinline__ void sc_partner_cell(const GridView& g, const PolyDev* __restrict__ pd, int X) {
  const PolyDev& D = pd[g.c_poly[X]];
  if (X - D.ncb < D.U) return;
  const int Pu = g.c_upcell[X];
  const int ub = g.c_bnum[Pu], ubs = g.c_bstart[Pu];
  int upedge = -1;
  for (int i = 0; i < ub; i++)
    if (g.b_subcell[ubs + i] == X) {
      upedge = ubs + i;
      break;
    }
  if (upedge == -1) return;
  const int xb = g.c_bstart[X];
  const int upid = g.b_id[upedge];
  if (g.c_bnum[X] == 6) {
    const int pe = g.b_partner[upedge];
    const int pc = g.b_parent[pe];
    const int u = g.c_bnum[pc], pbs = g.c_bstart[pc];
    const int pid = g.b_id[pe];
    link_or_clear(g, pbs + (pid + 1) % u, xb);
    int te = g.c_bstart[g.c_subcell[pc]] + pid;
    g.b_partner[xb + 1] = te;
    g.b_partner[te] = xb + 1;
    link_or_clear(g, pbs + (pid + u - 1) % u, xb + 2);
    link_or_clear(g, ubs + (upid + 1) % ub, xb + 3);
    te = g.c_bstart[g.c_subcell[Pu]] + upid;
    g.b_partner[xb + 4] = te;
    g.b_partner[te] = xb + 4;
    link_or_clear(g, ubs + (upid + ub - 1) % ub, xb + 5);
  } else if (g.c_bnum[X] == 4) {
    link_or_clear(g, ubs + (upid + 1) % ub, xb + 1);
    const int te = g.c_bstart[g.c_subcell[Pu]] + upid;
    g.b_partner[xb + 2] = te;
    g.b_partner[te] = xb + 2;
    link_or_clear(g, ubs + (upid + ub - 1) % ub, xb + 3);
  }
}

__global__ void k_sc_partner(GridView g, const PolyDev* __restrict__ pd, int lo, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) sc_partner_cell(g, pd, lo + t);
}

// over the cells of level lv + 1, the level just created
__global__ void k_sc_partner_dv(GridView g, const PolyDev* __restrict__ pd, const DevState* __restrict__ st, int lv) {
  if (!st->run) return;
  const LevelDesc L = st->lv[lv + 1];
  FOR_LEVEL_CELLS(L, t) sc_partner_cell(g, pd, L.cstart + t);
}

// --- dispatch (:1700-1784) ---
// Refs of each polygon: [roff, roff + R) in the edge-major ref array.
__global__ void k_poly_refs(PolyDev* pd, int npoly, const int32_t* off, const int32_t* cnt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npoly) return;
  PolyDev& P = pd[p];
  const int last = P.e0 + P.N - 1;
  P.roff = off[P.e0];
  P.R = off[last] + cnt[last] - P.roff;
}

// head flags at the first ref of every edge but each polygon's first
// (:971-981); an edge without refs shares its slot with the next edge.
__global__ void k_edge_heads(const int32_t* off, const int32_t* epoly, const PolyDev* __restrict__ pd, int Ntot,
                             int32_t* flag) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Ntot) return;
  const PolyDev& P = pd[epoly[e]];
  if (e == P.e0) return;
  const int o = off[e];
  if (o < P.roff + P.R) flag[o] = 1;
}

// ref -> global edge id from the inclusive scan of head flags
// (kernel_EC2CS_adjustEFlag, :984-994), counted from the polygon's first ref.
__global__ void k_ref_edges(const int32_t* rcell, const int32_t* incl, const int32_t* cpoly,
                            const PolyDev* __restrict__ pd, const DevState* st, int32_t* redge) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st->R) return;
  const PolyDev& P = pd[cpoly[rcell[i]]];
  const int base = P.roff > 0 ? incl[P.roff - 1] : 0;
  redge[i] = P.e0 + (incl[i] - base);
}

__global__ void k_count_cells(const int32_t* rcell, const DevState* st, int32_t* ccnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st->R) return;
  atomicAdd(&ccnt[rcell[i]], 1);
}

__global__ void k_scatter_refs(const int32_t* rcell, const int32_t* redge, const DevState* st, int32_t* fill,
                               int32_t* refs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st->R) return;
  const int pos = atomicAdd(&fill[rcell[i]], 1);
  refs[pos] = redge[i];
}

// kernel_EC2CS_setCEInfo (:1030-1043) for bottom cells, plus the ascending
// order of each run.  One wave per cell: each lane ranks its elements
// against the whole run (stable by position), staged in LDS when it fits.
constexpr int kSortCap = 1024;
__global__ void __launch_bounds__(256) k_finish_cells(GridView g, const PolyDev* __restrict__ pd, int ncell,
                                                      const int32_t* ccnt, const int32_t* coff,
                                                      const int32_t* __restrict__ src) {
  __shared__ int32_t lds[4][kSortCap];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + w;
  int n = 0, s = 0;
  bool bottom = false;
  if (t < ncell) {
    const PolyDev& P = pd[g.c_poly[t]];
    bottom = t >= P.bcs && t < P.bcs + P.bcn;
    if (bottom) {
      n = ccnt[t];
      s = coff[t];
    }
  }
  const bool in_lds = n <= kSortCap;
  if (in_lds)
    for (int k = lane; k < n; k += 64) lds[w][k] = src[s + k];
  __syncthreads();
  for (int k = lane; k < n; k += 64) {
    const int v = in_lds ? lds[w][k] : src[s + k];
    int rank = 0;
    for (int j = 0; j < n; j++) {
      const int u = in_lds ? lds[w][j] : src[s + j];
      rank += (u < v) || (u == v && j < k);
    }
    g.refs[s + rank] = v;
  }
  if (bottom && lane == 0) {
    g.c_rnum[t] = n;
    g.c_rstart[t] = s;
  }
}

__global__ void k_fill(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_copy(int32_t* d, const int32_t* s, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}

// ---------------------------------------------------------------------------
// Device-driven build: the host loop's per-level decisions as kernels
// ---------------------------------------------------------------------------
constexpr int kCtlBlock = 256;
constexpr int kDvGrid = 2048;  // blocks of the grid-stride launches

// Block-wide exclusive prefix of an int64 per thread (kCtlBlock threads);
// *total receives the block's sum.
__device__ __forceinline__ long long block_excl64(long long v, long long* total) {
  __shared__ long long ws[kCtlBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  long long inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const long long u = __shfl_up(inc, d, 64);
    if (lane >= d) inc += u;
  }
  if (lane == 63) ws[w] = inc;
  __syncthreads();
  long long base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kCtlBlock / 64; i++) {
    if (i < w) base += ws[i];
    tot += ws[i];
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

// Build start: the state cleared (as a host-side memset of it would), the
// polygon records from their template, level 0 of every polygon.
__global__ void __launch_bounds__(kCtlBlock) k_dv_start(DevState* st, PolyDev* pd, const PolyDev* __restrict__ pd0,
                                                        int np, LevelDesc* plv, int32_t* plvn, int32_t* pmarks,
                                                        float* gamma) {
  int32_t* w = (int32_t*)st;
  for (int i = threadIdx.x; i < (int)(sizeof(DevState) / 4); i += kCtlBlock) w[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int64_t V0 = (int64_t)kGrid0Verts * np, B0 = (int64_t)kGrid0Borders * np, C0 = (int64_t)kGrid0Cells * np;
    st->lv[0] = LevelDesc{0, (int32_t)V0, 0, (int32_t)B0, 0, (int32_t)C0, 0};
    st->gnum = 1;
    st->tot_v = V0;
    st->tot_b = B0;
    st->tot_c = C0;
    st->run = 1;
  }
  for (int p = threadIdx.x; p < np; p += kCtlBlock) {
    pd[p] = pd0[p];
    plv[(int64_t)p * kMaxLevels] = LevelDesc{kGrid0Verts * p, kGrid0Verts, kGrid0Borders * p, kGrid0Borders,
                                             kGrid0Cells * p, kGrid0Cells, 0};
    plvn[p] = 1;
    pmarks[p] = 0;
    gamma[p] = 0.f;
  }
}

// One copy of the level-0 grid per polygon (block = polygon); borders are
// (parent, partner, v0, v1, subcell, id), cells (bstart, bnum, rstart, rnum,
// upcell, subcell, type), all local to the copy.
__global__ void __launch_bounds__(kCtlBlock) k_dv_grid0(GridView g, const double* __restrict__ v0,
                                                        const double* __restrict__ n0, const int32_t* __restrict__ b0,
                                                        const int32_t* __restrict__ c0) {
  const int p = blockIdx.x;
  const int vb = kGrid0Verts * p, bb = kGrid0Borders * p, cb = kGrid0Cells * p;
  double* vw = const_cast<double*>(g.v);
  for (int i = threadIdx.x; i < 3 * kGrid0Verts; i += kCtlBlock) vw[3 * (int64_t)vb + i] = v0[i];
  for (int i = threadIdx.x; i < 3 * kGrid0Borders; i += kCtlBlock) g.b_n[3 * (int64_t)bb + i] = n0[i];
  for (int i = threadIdx.x; i < kGrid0Borders; i += kCtlBlock) {
    const int32_t* r = b0 + 6 * i;
    g.b_parent[bb + i] = r[0] + cb;
    g.b_partner[bb + i] = r[1] < 0 ? -1 : r[1] + bb;
    g.b_v0[bb + i] = r[2] + vb;
    g.b_v1[bb + i] = r[3] + vb;
    g.b_subcell[bb + i] = r[4];
    g.b_id[bb + i] = r[5];
  }
  for (int i = threadIdx.x; i < kGrid0Cells; i += kCtlBlock) {
    const int32_t* r = c0 + 7 * i;
    g.c_bstart[cb + i] = r[0] + bb;
    g.c_bnum[cb + i] = r[1];
    g.c_rstart[cb + i] = r[2];
    g.c_rnum[cb + i] = r[3];
    g.c_upcell[cb + i] = r[4];
    g.c_subcell[cb + i] = r[5];
    g.c_type[cb + i] = r[6];
    g.c_poly[cb + i] = p;
    g.c_root[cb + i] = i;
  }
}

// Top of level lv (createAHHG's loop, :1983-1992): whether this level's
// walks use walk records, the stop of polygons with a fixed depth, the stop
// of every polygon at the last level; the loop ends when none refines.
__global__ void __launch_bounds__(kCtlBlock) k_lv_begin(DevState* st, PolyDev* pd, int np, int lv, int maxd,
                                                        int32_t* pmarks, int64_t wrec_cap) {
  if (!st->run) return;
  __syncthreads();  // every thread has read run before thread 0 may clear it
  if (threadIdx.x == 0) {
    const LevelDesc GL = st->lv[lv];
    if (GL.cnum > 0 && lv < 32 && (int64_t)GL.cstart + GL.cnum <= wrec_cap &&
        st->extra_next >= kWalkRecStepsPerCell * (double)GL.cnum)
      st->rec_levels |= 1u << lv;
  }
  int any = 0;
  for (int p = threadIdx.x; p < np; p += kCtlBlock) {
    PolyDev& P = pd[p];
    if (P.refining && ((P.depth > 0 && lv == P.depth - 1) || lv == maxd - 1)) {
      P.refining = 0;
      P.L = lv;
      P.use_prev = 0;
    }
    if (P.refining) {
      any = 1;
      pmarks[p] += 1;
    }
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0 && !any) st->run = 0;
}

// After level lv's plan (divGrid's host part and the density rule,
// :1595-1605): the mean walk length, the density rule per polygon, the
// next level's offsets per polygon, the arena and int32 checks, the next
// level's descriptor and the expected walk steps there.
__global__ void __launch_bounds__(kCtlBlock) k_lv_step(PolyDev* pd, int np, DevState* st, int lv, int autod,
                                                       float* gamma, int64_t vcap, int64_t bcap, int64_t ccap) {
  if (!st->run) return;
  __shared__ double sh_m;
  __shared__ int sh_ok;
  // walkers of this level's mark walk and its steps per walker
  long long wk = 0;
  for (int p = threadIdx.x; p < np; p += kCtlBlock) wk += pd[p].refining ? pd[p].N : 0;
  long long walkers;
  (void)block_excl64(wk, &walkers);
  if (threadIdx.x == 0) {
    const unsigned long long steps = st->steps_mark;
    sh_m = walkers > 0 ? (double)(steps - st->steps_before) / (double)walkers : 0.0;
    st->steps_before = steps;
  }
  // density rule, then offsets in polygon order (a block scan per 256 polygons)
  const int64_t tv = st->tot_v, tb = st->tot_b, tc = st->tot_c;
  long long cv = 0, cb = 0, cc = 0, wn = 0;
  for (int base = 0; base < np; base += kCtlBlock) {
    const int p = base + threadIdx.x;
    long long a = 0, b = 0, d = 0;
    if (p < np) {
      PolyDev& P = pd[p];
      if (P.refining && autod && P.depth == 0) {
        const float k = P.U * 1.0 / P.N;
        gamma[p] = k;
        if (k > 0.1 || (k < 0.1 && (0.1 - k) < 0.01)) {
          P.refining = 0;
          P.L = lv;
          P.use_prev = lv != 0 ? 1 : 0;
        }
      }
      if (P.refining) {
        a = P.nv;
        b = (long long)P.nv + P.nbb;
        d = (long long)P.U + P.nbc;
        wn += P.N;
      }
    }
    long long ta, tb2, td;
    const long long ea = block_excl64(a, &ta);
    const long long eb = block_excl64(b, &tb2);
    const long long ed = block_excl64(d, &td);
    if (p < np && pd[p].refining) {
      PolyDev& P = pd[p];
      P.nvb = (int32_t)(tv + cv + ea);
      P.nbb0 = (int32_t)(tb + cb + eb);
      P.ncb = (int32_t)(tc + cc + ed);
    }
    cv += ta;
    cb += tb2;
    cc += td;
  }
  long long wnext;
  (void)block_excl64(wn, &wnext);
  if (threadIdx.x == 0) {
    sh_ok = 1;
    if (cc == 0) {
      st->run = 0;  // nothing to create: the loop ends here
      sh_ok = 0;
    } else if (tv + cv > INT32_MAX || tb + cb > INT32_MAX || tc + cc > INT32_MAX) {
      st->limit = 1;
      st->abort = 1;
      st->run = 0;
      sh_ok = 0;
    } else if (tv + cv > vcap || tb + cb > bcap || tc + cc > ccap) {
      st->overflow = 1;
      st->need_v = tv + cv;
      st->need_b = tb + cb;
      st->need_c = tc + cc;
      st->abort = 1;
      st->run = 0;
      sh_ok = 0;
    }
    if (sh_ok) {
      st->lv[lv + 1] = LevelDesc{(int32_t)tv, (int32_t)cv, (int32_t)tb, (int32_t)cb, (int32_t)tc, (int32_t)cc, 0};
      st->gnum = lv + 2;
      st->tot_v = tv + cv;
      st->tot_b = tb + cb;
      st->tot_c = tc + cc;
      st->extra_next = (double)wnext * 2.0 * fmax(sh_m - 1.0, 0.0);
    }
  }
}

// Every polygon still refining moves to the level just created.
__global__ void k_lv_advance(PolyDev* pd, int np, const DevState* __restrict__ st, int lv, LevelDesc* plv,
                             int32_t* plvn) {
  if (!st->run) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  PolyDev& P = pd[p];
  if (!P.refining) return;
  LevelDesc* L = plv + (int64_t)p * kMaxLevels;
  L[lv].rnum = P.U;
  L[lv + 1] = LevelDesc{P.nvb, P.nv, P.nbb0, P.nv + P.nbb, P.ncb, P.U + P.nbc, 0};
  plvn[p] = lv + 2;
  P.pbs = P.bs;
  P.plbs = P.lbs;
  P.lbs = P.lbs + P.bn;
  P.cs = P.ncb;
  P.cn = P.U + P.nbc;
  P.bs = P.nbb0;
  P.bn = P.nv + P.nbb;
}

// End of the enqueued levels: a loop still running means levels are left.
__global__ void k_dv_loop_end(DevState* st) {
  if (st->run) st->abort = 1;
  st->tot_c32 = (int32_t)st->tot_c;
}

// Bottom segments and the density ratio of polygons that never refine.
__global__ void k_dv_bottom(PolyDev* pd, int np, DevState* st, const LevelDesc* __restrict__ plv, float* gamma,
                            int maxd) {
  if (st->abort) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  PolyDev& P = pd[p];
  if ((P.depth == 0 && maxd == 1) || P.depth == 1) gamma[p] = (float)(42 * 1.0 / P.N);
  const LevelDesc B = plv[(int64_t)p * kMaxLevels + P.L];
  P.bcs = B.cstart;
  P.bcn = B.cnum;
  P.refining = 0;
  atomicMax(&st->maxL, P.L);
}

// Refs of each polygon after the count walk, and the check that they fit.
__global__ void k_dv_poly_refs(PolyDev* pd, int np, const DevState* __restrict__ st, const int32_t* off,
                               const int32_t* cnt, LevelDesc* plv) {
  if (st->abort) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  PolyDev& P = pd[p];
  const int last = P.e0 + P.N - 1;
  P.roff = off[P.e0];
  P.R = off[last] + cnt[last] - P.roff;
  plv[(int64_t)p * kMaxLevels + P.L].rnum = P.R;
}

// Write phase of the device-driven dispatch, first part: an edge of a
// polygon that stopped after marking its bottom level (use_prev), whose
// clean mark walk there fitted the walk slots, gets its cells (in walk
// order, the write walk's) from the slots; every other edge is flagged for
// the write walk.
__global__ void k_dv_slot_copy(const int32_t* __restrict__ epoly, const PolyDev* __restrict__ pd, int Ntot,
                               const int32_t* __restrict__ slot, const int32_t* __restrict__ cnt,
                               const int32_t* __restrict__ off, int32_t* __restrict__ rcell,
                               int32_t* __restrict__ flag, const DevState* __restrict__ st) {
  if (st->abort) return;  // the refs may not fit: nothing is written
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Ntot) return;
  const int n = cnt[e];
  const bool copy = pd[epoly[e]].use_prev && n <= kSlotK && slot[e] != kSlotDirty;
  flag[e] = copy ? 0 : 1;
  if (copy) {
    int32_t* o = rcell + off[e];
    for (int k = 0; k < n; k++) o[k] = slot[(int64_t)k * Ntot + e];
  }
}

// The flagged edges in edge order (pos: exclusive scan of the flags).
__global__ void k_dv_walk_list(const int32_t* __restrict__ flag, const int32_t* __restrict__ pos, int Ntot,
                               int32_t* __restrict__ list, const DevState* __restrict__ st) {
  if (st->abort) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < Ntot && flag[e]) list[pos[e]] = e;
}

__global__ void k_dv_check_refs(DevState* st, int64_t rcap) {
  if (st->abort) return;
  const int32_t R = st->R;
  if (R < 0) {
    st->limit = 1;
    st->abort = 1;
  } else if ((int64_t)R + 1 > rcap) {
    st->overflow = 1;
    st->need_r = (int64_t)R + 1;
    st->abort = 1;
  } else {
    st->R_ok = R;
  }
}

// p[0, *n) = v, with *n read on the device (grid-stride).
__global__ void k_fill_dv(int32_t* p, const int32_t* __restrict__ n, int32_t v, const DevState* __restrict__ st) {
  if (st->abort) return;
  const int m = *n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) p[i] = v;
}

// d[0, *n) = s[0, *n).
__global__ void k_copy_dv(int32_t* d, const int32_t* __restrict__ s, const int32_t* __restrict__ n,
                          const DevState* __restrict__ st) {
  if (st->abort) return;
  const int m = *n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) d[i] = s[i];
}

__global__ void k_dv_edge_heads(const int32_t* off, const int32_t* epoly, const PolyDev* __restrict__ pd, int Ntot,
                                int32_t* flag, const DevState* __restrict__ st) {
  if (st->abort) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Ntot) return;
  const PolyDev& P = pd[epoly[e]];
  if (e == P.e0) return;
  const int o = off[e];
  if (o < P.roff + P.R) flag[o] = 1;
}

// ref -> global edge id (as k_ref_edges), and the ref's cell as a sort key.
__global__ void k_dv_ref_edges(const int32_t* rcell, const int32_t* incl, const int32_t* cpoly,
                               const PolyDev* __restrict__ pd, const DevState* __restrict__ st, int32_t* redge) {
  if (st->abort) return;
  const int R = st->R_ok;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < R; i += gridDim.x * blockDim.x) {
    const PolyDev& P = pd[cpoly[rcell[i]]];
    const int base = P.roff > 0 ? incl[P.roff - 1] : 0;
    redge[i] = P.e0 + (incl[i] - base);
  }
}

// Runs of equal cells in the sorted refs: the run of cell k spans
// [start[k], end[k]) (both left 0 for a cell without refs).
__global__ void k_dv_cell_runs(const uint32_t* __restrict__ key, const DevState* __restrict__ st,
                               int32_t* __restrict__ start, int32_t* __restrict__ end) {
  if (st->abort) return;
  const int R = st->R_ok;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < R; i += gridDim.x * blockDim.x) {
    const uint32_t k = key[i];
    if (i == 0 || key[i - 1] != k) start[k] = i;
    if (i == R - 1 || key[i + 1] != k) end[k] = i + 1;
  }
}

// Refs per cell from its run: cnt[k] = end[k] - start[k], in place of end.
__global__ void k_dv_run_counts(const int32_t* __restrict__ start, int32_t* __restrict__ end,
                                const DevState* __restrict__ st) {
  if (st->abort) return;
  const int C = st->tot_c32;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < C; k += gridDim.x * blockDim.x) end[k] -= start[k];
}

// kernel_EC2CS_setCEInfo (:1030-1043) for bottom cells: the refs are in
// cell-major order already (a stable sort by cell of the edge-major refs,
// so ascending edge order within each cell).
__global__ void k_dv_finish_cells(GridView g, const PolyDev* __restrict__ pd, const DevState* __restrict__ st,
                                  const int32_t* ccnt, const int32_t* coff) {
  if (st->abort) return;
  const int C = st->tot_c32;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < C; t += gridDim.x * blockDim.x) {
    const PolyDev& P = pd[g.c_poly[t]];
    if (t >= P.bcs && t < P.bcs + P.bcn) {
      g.c_rnum[t] = ccnt[t];
      g.c_rstart[t] = coff[t];
    }
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
int set_err(gahhg_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  c->err = buf;
  return code;
}

int hip_check(gahhg_ctx* c, int s, const char* what) {
  if (s == hipSuccess) return GAHHG_OK;
  return set_err(c, GAHHG_E_HIP, "%s: %s", what, hipGetErrorString((hipError_t)s));
}

int dalloc(gahhg_ctx* c, void** p, int64_t bytes) {
  *p = nullptr;
  if (bytes <= 0) bytes = 256;
  hipError_t s = hipMalloc(p, (size_t)bytes);
  if (s != hipSuccess) {
    *p = nullptr;
    return set_err(c, GAHHG_E_NOMEM, "hipMalloc(%lld): %s", (long long)bytes, hipGetErrorString(s));
  }
  return GAHHG_OK;
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

int use_buf(gahhg_ctx* c, DevBuf& b, int64_t bytes, void** active) {
  bytes = std::max<int64_t>(bytes, 256);
  if (b.p == nullptr || bytes > b.cap) {
    (void)hipStreamSynchronize(c->stream);
    free_buf(b);
    const int64_t n = bytes + bytes / 8;  // headroom for the next, slightly larger build
    if (dalloc(c, &b.p, n)) return GAHHG_E_NOMEM;
    b.cap = n;
  }
  *active = b.p;
  return GAHHG_OK;
}

void free_buf(DevBuf& b) {
  dfree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

int fill_i32(gahhg_ctx* c, int32_t* p, int64_t n, int32_t v) {
  if (n <= 0) return GAHHG_OK;
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, p, n, v);
  return hip_check(c, hipGetLastError(), "fill");
}

int copy_i32(gahhg_ctx* c, int32_t* d, const int32_t* s, int64_t n) {
  if (n <= 0) return GAHHG_OK;
  hipLaunchKernelGGL(k_copy, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, d, s, n);
  return hip_check(c, hipGetLastError(), "copy");
}

GridView view(gahhg_ctx* c) {
  GridView v;
  v.v = c->g.v;
  v.b_parent = c->g.b_parent;
  v.b_partner = c->g.b_partner;
  v.b_v0 = c->g.b_v0;
  v.b_v1 = c->g.b_v1;
  v.b_subcell = c->g.b_subcell;
  v.b_id = c->g.b_id;
  v.b_n = c->g.b_n;
  v.c_bstart = c->g.c_bstart;
  v.c_bnum = c->g.c_bnum;
  v.c_rstart = c->g.c_rstart;
  v.c_rnum = c->g.c_rnum;
  v.c_upcell = c->g.c_upcell;
  v.c_subcell = c->g.c_subcell;
  v.c_type = c->g.c_type;
  v.c_poly = c->g.c_poly;
  v.c_root = c->g.c_root;
  v.refs = c->g.refs;
  return v;
}

#define CK_SYNC(c)                                                                      \
  do {                                                                                  \
    int _rc = hip_check((c), hipStreamSynchronize((c)->stream), "stream sync");         \
    if (_rc != GAHHG_OK) return _rc;                                                    \
  } while (0)

// Grows a device array preserving `keep` leading elements (checkStorage,
// :1820-1910).  The stream must be idle.
template <typename T>
static int grow(gahhg_ctx* c, T** p, int64_t keep, int64_t need, int64_t newcap) {
  (void)need;
  CK_SYNC(c);
  T* q = nullptr;
  if (dalloc(c, (void**)&q, newcap * (int64_t)sizeof(T))) return GAHHG_E_NOMEM;
  if (*p && keep > 0) {
    hipError_t s = hipMemcpy(q, *p, (size_t)keep * sizeof(T), hipMemcpyDeviceToDevice);
    if (s != hipSuccess) {
      dfree(q);
      return hip_check(c, s, "grow copy");
    }
  }
  dfree(*p);
  *p = q;
  return GAHHG_OK;
}

static int64_t next_cap(int64_t cap, int64_t need) {
  int64_t n = std::max<int64_t>(cap, 1024);
  while (n < need) n *= 2;
  return n;
}

static int ensure_v(gahhg_ctx* c, int64_t used, int64_t need) {
  if (need <= c->g.vcap) return GAHHG_OK;
  int64_t nc = next_cap(c->g.vcap, need);
  if (grow(c, &c->g.v, 3 * used, 3 * need, 3 * nc)) return GAHHG_E_NOMEM;
  c->g.vcap = nc;
  return GAHHG_OK;
}

static int ensure_b(gahhg_ctx* c, int64_t used, int64_t need) {
  if (need <= c->g.bcap) return GAHHG_OK;
  int64_t nc = next_cap(c->g.bcap, need);
  DeviceGrid& g = c->g;
  if (grow(c, &g.b_parent, used, need, nc) || grow(c, &g.b_partner, used, need, nc) ||
      grow(c, &g.b_v0, used, need, nc) || grow(c, &g.b_v1, used, need, nc) ||
      grow(c, &g.b_subcell, used, need, nc) || grow(c, &g.b_id, used, need, nc) ||
      grow(c, &g.b_n, 3 * used, 3 * need, 3 * nc))
    return GAHHG_E_NOMEM;
  g.bcap = nc;
  return GAHHG_OK;
}

static int ensure_c(gahhg_ctx* c, int64_t used, int64_t need) {
  if (need <= c->g.ccap) return GAHHG_OK;
  int64_t nc = next_cap(c->g.ccap, need);
  DeviceGrid& g = c->g;
  if (grow(c, &g.c_bstart, used, need, nc) || grow(c, &g.c_bnum, used, need, nc) ||
      grow(c, &g.c_rstart, used, need, nc) || grow(c, &g.c_rnum, used, need, nc) ||
      grow(c, &g.c_upcell, used, need, nc) || grow(c, &g.c_subcell, used, need, nc) ||
      grow(c, &g.c_type, used, need, nc) || grow(c, &g.c_poly, used, need, nc) ||
      grow(c, &g.c_root, used, need, nc))
    return GAHHG_E_NOMEM;
  g.ccap = nc;
  return GAHHG_OK;
}

// Walk records for cells [0, need); the first `used` are kept on growth.
static int ensure_wrec(gahhg_ctx* c, int64_t used, int64_t need) {
  if (need <= c->wrec_cap) return GAHHG_OK;
  const int64_t nc = next_cap(c->wrec_cap, need);
  WalkRec* p = nullptr;
  if (dalloc(c, (void**)&p, nc * (int64_t)sizeof(WalkRec))) return GAHHG_E_NOMEM;
  if (used > 0 && c->d_wrec) {
    const int rc = hip_check(c, hipMemcpyAsync(p, c->d_wrec, used * sizeof(WalkRec), hipMemcpyDeviceToDevice, c->stream),
                             "walk record copy");
    if (rc) {
      dfree(p);
      return rc;
    }
  }
  CK_SYNC(c);
  dfree(c->d_wrec);
  c->d_wrec = p;
  c->wrec_cap = nc;
  return GAHHG_OK;
}

static int ensure_flags(gahhg_ctx* c, int64_t need) {
  if (need <= c->fcap) return GAHHG_OK;
  CK_SYNC(c);
  int64_t nc = next_cap(c->fcap, need);
  int32_t** bufs[8] = {&c->d_f0, &c->d_f1, &c->d_o0, &c->d_o1, &c->d_f2, &c->d_f3, &c->d_o2, &c->d_o3};
  for (int i = 0; i < 8; i++) {
    dfree(*bufs[i]);
    *bufs[i] = nullptr;
  }
  for (int i = 0; i < 8; i++)
    if (dalloc(c, (void**)bufs[i], nc * 4)) return GAHHG_E_NOMEM;
  c->fcap = nc;
  return GAHHG_OK;
}

static int ensure_refs(gahhg_ctx* c, int64_t R) {
  if (R <= c->r2cap && R <= c->g.rcap) return GAHHG_OK;
  CK_SYNC(c);
  int64_t nc = next_cap(std::max(c->r2cap, c->g.rcap), R);
  dfree(c->d_rcell);
  dfree(c->d_rflag);
  dfree(c->d_redge);
  dfree(c->d_rscan);
  dfree(c->g.refs);
  c->d_rcell = c->d_rflag = c->d_redge = c->d_rscan = c->g.refs = nullptr;
  if (dalloc(c, (void**)&c->d_rcell, nc * 4) || dalloc(c, (void**)&c->d_rflag, nc * 4) ||
      dalloc(c, (void**)&c->d_redge, nc * 4) || dalloc(c, (void**)&c->d_rscan, nc * 4) ||
      dalloc(c, (void**)&c->g.refs, nc * 4))
    return GAHHG_E_NOMEM;
  c->r2cap = nc;
  c->g.rcap = nc;
  return GAHHG_OK;
}

static int sync_state(gahhg_ctx* c) {
  hipError_t s = hipMemcpyAsync(c->h_st, c->d_st, sizeof(DevState), hipMemcpyDeviceToHost, c->stream);
  if (s == hipSuccess) s = hipStreamSynchronize(c->stream);
  return hip_check(c, s, "state readback");
}

static int pd_download(gahhg_ctx* c) {
  hipError_t s = hipMemcpyAsync(c->pd.data(), c->d_pd, c->pd.size() * sizeof(PolyDev), hipMemcpyDeviceToHost,
                                c->stream);
  if (s == hipSuccess) s = hipStreamSynchronize(c->stream);
  return hip_check(c, s, "polygon state readback");
}

static int pd_upload(gahhg_ctx* c) {
  return hip_check(c, hipMemcpyAsync(c->d_pd, c->pd.data(), c->pd.size() * sizeof(PolyDev), hipMemcpyHostToDevice,
                                     c->stream),
                   "polygon state upload");
}

#define CK(x)                        \
  do {                               \
    int _rc = (x);                   \
    if (_rc != GAHHG_OK) return _rc; \
  } while (0)
#define LAUNCH_CK(name) CK(hip_check(c, hipGetLastError(), name))

// GPISP::createAHHGInit + createAHHG (:1947-2015) + dispatchSELast
// (:1700-1784) for every polygon of the context at once.  Level l of every
// polygon still refining is processed by the same launches; one host
// synchronisation per level applies the density rule per polygon and sizes
// the next level.
// Device arrays for a hierarchy read from a file: sized for its counts (the
// contents are copied in by the caller), per-edge records and scratch.
int attach_loaded_grid(gahhg_ctx* c, int64_t V, int64_t B, int64_t C, int64_t R) {
  const int64_t Nt = c->n_edges;
  if (Nt > 0) {
    hipLaunchKernelGGL(k_edges, dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, (const double*)c->d_pv,
                       (const int32_t*)c->d_pe, (int)Nt, c->d_edge);
    CK(hip_check(c, hipGetLastError(), "edge records"));
  }
  CK(ensure_v(c, 0, std::max<int64_t>(V, 1)));
  CK(ensure_b(c, 0, std::max<int64_t>(B, 1)));
  CK(ensure_c(c, 0, std::max<int64_t>(C, 1)));
  CK(ensure_flags(c, std::max<int64_t>(Nt, C) + 1));
  CK(ensure_refs(c, R + 1));
  return GAHHG_OK;
}

int build_hierarchy(gahhg_ctx* c) {
  const int np = (int)c->polys.size();
  const int64_t Nt = c->n_edges;
  const int maxd = c->prm.max_depth;
  const bool autod = c->prm.auto_depth != 0;

  // per-edge records (p0, p1, plane normal) for every walk and leaf test
  hipLaunchKernelGGL(k_edges, dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, (const double*)c->d_pv,
                     (const int32_t*)c->d_pe, (int)Nt, c->d_edge);
  LAUNCH_CK("edge records");

  // level 0: one copy of the 42-cell grid per polygon
  std::vector<double> v0(3 * kGrid0Verts), bn0(3 * kGrid0Borders);
  std::vector<int32_t> bi0(6 * kGrid0Borders), ci0(7 * kGrid0Cells);
  build_grid0(v0.data(), bi0.data(), bn0.data(), ci0.data());
  const int64_t V0 = (int64_t)kGrid0Verts * np, B0 = (int64_t)kGrid0Borders * np, C0 = (int64_t)kGrid0Cells * np;
  // capacities (createAHHGInit :1949-1953, summed over the batch)
  CK(ensure_v(c, 0, std::max<int64_t>(Nt, 160) + V0));
  CK(ensure_b(c, 0, std::max<int64_t>(4 * Nt, 480) + B0));
  CK(ensure_c(c, 0, std::max<int64_t>(Nt, 84) + C0));
  CK(ensure_flags(c, std::max<int64_t>(4 * Nt, 480) + C0));
  {
    std::vector<double> hv(3 * V0), hn(3 * B0);
    std::vector<int32_t> cols[7], bcols[6];
    for (auto& x : cols) x.resize(C0);
    for (auto& x : bcols) x.resize(B0);
    std::vector<int32_t> cpoly(C0), croot(C0);
    for (int p = 0; p < np; p++) {
      const int vb = kGrid0Verts * p, bb = kGrid0Borders * p, cb = kGrid0Cells * p;
      std::copy(v0.begin(), v0.end(), hv.begin() + 3 * (int64_t)vb);
      std::copy(bn0.begin(), bn0.end(), hn.begin() + 3 * (int64_t)bb);
      for (int i = 0; i < kGrid0Borders; i++) {
        bcols[0][bb + i] = bi0[6 * i] + cb;                                   // parent
        bcols[1][bb + i] = bi0[6 * i + 1] < 0 ? -1 : bi0[6 * i + 1] + bb;     // partner
        bcols[2][bb + i] = bi0[6 * i + 2] + vb;                               // v0
        bcols[3][bb + i] = bi0[6 * i + 3] + vb;                               // v1
        bcols[4][bb + i] = bi0[6 * i + 4];                                    // subcell (-1)
        bcols[5][bb + i] = bi0[6 * i + 5];                                    // id
      }
      for (int i = 0; i < kGrid0Cells; i++) {
        cols[0][cb + i] = ci0[7 * i] + bb;
        for (int k = 1; k < 7; k++) cols[k][cb + i] = ci0[7 * i + k];
        cpoly[cb + i] = p;
        croot[cb + i] = i;
      }
    }
    hipError_t s = hipMemcpy(c->g.v, hv.data(), hv.size() * 8, hipMemcpyHostToDevice);
    if (s == hipSuccess) s = hipMemcpy(c->g.b_n, hn.data(), hn.size() * 8, hipMemcpyHostToDevice);
    int32_t* bd[6] = {c->g.b_parent, c->g.b_partner, c->g.b_v0, c->g.b_v1, c->g.b_subcell, c->g.b_id};
    for (int k = 0; k < 6 && s == hipSuccess; k++) s = hipMemcpy(bd[k], bcols[k].data(), B0 * 4, hipMemcpyHostToDevice);
    int32_t* cd[7] = {c->g.c_bstart, c->g.c_bnum, c->g.c_rstart, c->g.c_rnum, c->g.c_upcell, c->g.c_subcell, c->g.c_type};
    for (int k = 0; k < 7 && s == hipSuccess; k++) s = hipMemcpy(cd[k], cols[k].data(), C0 * 4, hipMemcpyHostToDevice);
    if (s == hipSuccess) s = hipMemcpy(c->g.c_poly, cpoly.data(), C0 * 4, hipMemcpyHostToDevice);
    if (s == hipSuccess) s = hipMemcpy(c->g.c_root, croot.data(), C0 * 4, hipMemcpyHostToDevice);
    CK(hip_check(c, s, "grid0 upload"));
  }

  // polygon records
  c->pd.assign(np, PolyDev());
  c->plv.assign(np, std::vector<LevelDesc>());
  c->pgamma.assign(np, 0.f);
  int64_t e0 = 0;
  for (int p = 0; p < np; p++) {
    PolyDev& P = c->pd[p];
    memset(&P, 0, sizeof(P));
    P.e0 = (int32_t)e0;
    P.N = c->polys[p].N;
    P.cell0 = kGrid0Cells * p;
    P.refining = maxd > 1 ? 1 : 0;
    P.L = 0;
    P.cs = kGrid0Cells * p;
    P.cn = kGrid0Cells;
    P.bs = kGrid0Borders * p;
    P.bn = kGrid0Borders;
    P.lbs = 0;
    P.tbl = -1;
    P.tbins = 0;
    P.depth = p < (int)c->depth_override.size() ? c->depth_override[p] : 0;
    if (P.depth == 1) P.refining = 0;  // level 0 is the bottom; nothing is marked
    c->plv[p].push_back(LevelDesc{kGrid0Verts * p, kGrid0Verts, kGrid0Borders * p, kGrid0Borders, kGrid0Cells * p,
                                  kGrid0Cells, 0});
    e0 += P.N;
  }
  c->glv.clear();
  c->glv.push_back(LevelDesc{0, (int32_t)V0, 0, (int32_t)B0, 0, (int32_t)C0, 0});
  c->tot_v = V0;
  c->tot_b = B0;
  c->tot_c = C0;
  CK(pd_upload(c));
  const bool learned = c->learn.on;
  if (learned) {
    CK(learn_prepare(c));
    c->learn.h_dec.assign(np, 1);
    c->learn.rule_stop.assign(np, -1);
  }
  memset(c->h_st, 0, sizeof(DevState));
  CK(hip_check(c, hipMemcpyAsync(c->d_st, c->h_st, sizeof(DevState), hipMemcpyHostToDevice, c->stream), "state init"));
  CK(fill_i32(c, c->d_loc_cur, Nt, 0));
  CK(fill_i32(c, c->d_loc_prev, Nt, 0));
  if (c->ecap < Nt) {  // per-edge counts: written by the mark walks, read by the dispatch
    dfree(c->d_ecnt);
    c->d_ecnt = nullptr;
    CK(dalloc(c, (void**)&c->d_ecnt, Nt * 4));
    c->ecap = Nt;
  }

  GridView g = view(c);
  c->pmarks.assign(c->pd.size(), 0);
  // levels with walk records (bit l), and the next level's expected steps
  // beyond each edge's first cell: from the mark walk's cells per walking
  // edge at this level, m, the next level's is about 1 + 2 (m - 1) (the
  // cells halve; most edges stay in one)
  uint32_t rec_levels = 0;
  double extra_next = 0.0, m_level = 0.0;
  unsigned long long steps_before = 0;
  int lv = 0;
  for (;; lv++) {
    const LevelDesc GL = c->glv[lv];
    // walk records of this level's cells (complete since the level was
    // created; the mark walk below and the dispatch read them), when the
    // walks here are expected to be long enough to gain from them
    if (GL.cnum > 0 && lv < 32 && extra_next >= kWalkRecStepsPerCell * (double)GL.cnum) {
      CK(ensure_wrec(c, GL.cstart, (int64_t)GL.cstart + GL.cnum));
      hipLaunchKernelGGL(k_walk_records, dim3(blocks_for(GL.cnum, kRecBlock)), dim3(kRecBlock), 0, c->stream, g,
                         GL.cstart, GL.cnum, c->d_wrec);
      LAUNCH_CK("walk records");
      rec_levels |= 1u << lv;
    }
    bool any = false;
    for (const PolyDev& P : c->pd) any |= P.refining != 0;
    // polygons with a fixed depth stop before marking their last level
    // (the level loop of createAHHG with autoDepth off, :1983-1992)
    bool stopped = false;
    for (PolyDev& P : c->pd)
      if (P.refining && P.depth > 0 && lv == P.depth - 1) {
        P.refining = 0;
        P.L = lv;
        P.use_prev = 0;
        stopped = true;
      }
    if (stopped) CK(pd_upload(c));  // the level's kernels read the device copy
    any = false;
    for (const PolyDev& P : c->pd) any |= P.refining != 0;
    if (!any) break;
    if (lv == maxd - 1) {
      // the last level is created but never marked (createAHHG :1983)
      for (PolyDev& P : c->pd)
        if (P.refining) {
          P.refining = 0;
          P.L = lv;
          P.use_prev = 0;
        }
      break;
    }
    if (c->prm.verbose) printf("Subdivide level %d on GPU.\n", lv);
    for (size_t p = 0; p < c->pd.size(); p++) c->pmarks[p] += c->pd[p].refining ? 1 : 0;
    CK(ensure_flags(c, (int64_t)GL.cnum + 1));
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(GL.cnum, kBlock)), dim3(kBlock), 0, c->stream, g, GL.cstart, GL.cnum,
                       lv == 0 ? 1 : 0);
    LAUNCH_CK("zero");
    if ((rec_levels >> lv) & 1u)
      hipLaunchKernelGGL((k_walk<kMark, true>), dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, g,
                         (const WalkRec*)c->d_wrec, rec_levels, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                         (const PolyDev*)c->d_pd, (int)Nt, lv, c->d_loc_cur, c->d_loc_prev, (const int32_t*)nullptr,
                         c->d_ecnt, (int32_t*)nullptr, c->d_st, 0, (int32_t*)nullptr, (const int32_t*)nullptr);
    else
      hipLaunchKernelGGL((k_walk<kMark, false>), dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, g,
                         (const WalkRec*)c->d_wrec, rec_levels, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                         (const PolyDev*)c->d_pd, (int)Nt, lv, c->d_loc_cur, c->d_loc_prev, (const int32_t*)nullptr,
                         c->d_ecnt, (int32_t*)nullptr, c->d_st, 0, (int32_t*)nullptr, (const int32_t*)nullptr);
    LAUNCH_CK("mark walk");
    hipLaunchKernelGGL(k_plan, dim3(blocks_for(GL.cnum, kBlock)), dim3(kBlock), 0, c->stream, g, GL.cstart, GL.cnum,
                       c->d_f0, c->d_f1, c->d_f2, c->d_f3);
    LAUNCH_CK("plan");
    CK(scan_i32(c, c->d_f0, c->d_o0, GL.cnum, nullptr, nullptr, 0));
    CK(scan_i32(c, c->d_f1, c->d_o1, GL.cnum, nullptr, nullptr, 0));
    CK(scan_i32(c, c->d_f2, c->d_o2, GL.cnum, nullptr, nullptr, 0));
    CK(scan_i32(c, c->d_f3, c->d_o3, GL.cnum, nullptr, nullptr, 0));
    hipLaunchKernelGGL(k_plan_totals, dim3(blocks_for(np, kBlock)), dim3(kBlock), 0, c->stream, c->d_pd, np,
                       GL.cstart, (const int32_t*)c->d_f0, (const int32_t*)c->d_f1, (const int32_t*)c->d_f2,
                       (const int32_t*)c->d_f3, (const int32_t*)c->d_o0, (const int32_t*)c->d_o1,
                       (const int32_t*)c->d_o2, (const int32_t*)c->d_o3, (const DevState*)nullptr, 0);
    LAUNCH_CK("plan totals");
    // the model runs only at levels where some polygon may be inside its
    // window (crossed ratio within a factor 4 of the floor at the previous
    // level); below it every polygon refines anyway
    bool model_level = false;
    if (learned)
      for (int p = 0; p < np; p++) {
        const PolyDev& P = c->pd[p];
        model_level |= P.refining && P.depth == 0 && c->pgamma[p] * 4.0f >= kLearnFloor;
      }
    if (model_level) {
      CK(learn_level(c, lv, GL.cstart, GL.cnum, c->learn.h_dec.data()));
    } else if (learned) {
      std::fill(c->learn.h_dec.begin(), c->learn.h_dec.end(), 1);
    }
    CK(hip_check(c, hipMemcpyAsync(&c->h_st->steps_mark, &c->d_st->steps_mark, sizeof(unsigned long long),
                                   hipMemcpyDeviceToHost, c->stream),
                 "walk steps"));
    CK(pd_download(c));  // also completes the decisions' and the step count's copies
    {
      int64_t walkers = 0;
      for (const PolyDev& P : c->pd) walkers += P.refining ? P.N : 0;
      const unsigned long long steps = c->h_st->steps_mark;
      const double m = walkers > 0 ? (double)(steps - steps_before) / (double)walkers : 0.0;
      steps_before = steps;
      m_level = m;
    }
    // density rule per polygon (:1597-1604) and the next level's layout
    int64_t nv_all = 0, nb_all = 0, nc_all = 0;
    for (int p = 0; p < np; p++) {
      PolyDev& P = c->pd[p];
      if (!P.refining) continue;
      if (learned && P.depth == 0) {
        // the model decides only inside the window it was trained on:
        // refine while crossed cells / N < kLearnFloor (one to two levels
        // before the density rule would stop), never more than kLearnAbove
        // levels past the level where the density rule stops
        float k = P.U * 1.0 / P.N;
        c->pgamma[p] = k;
        int& rs = c->learn.rule_stop[p];
        if (rs < 0 && (k > 0.1 || (k < 0.1 && (0.1 - k) < 0.01))) rs = lv;
        bool stop;
        if (k < kLearnFloor)
          stop = false;
        else if (rs >= 0 && lv >= rs + kLearnAbove)
          stop = true;
        else
          stop = !c->learn.h_dec[p];
        if (stop) {
          P.refining = 0;
          P.L = lv;
          P.use_prev = lv != 0 ? 1 : 0;
          continue;
        }
      } else if (autod && P.depth == 0) {
        float k = P.U * 1.0 / P.N;
        c->pgamma[p] = k;
        if (k > 0.1 || (k < 0.1 && (0.1 - k) < 0.01)) {
          P.refining = 0;
          P.L = lv;
          P.use_prev = lv != 0 ? 1 : 0;
          continue;
        }
      }
      P.nvb = (int32_t)(c->tot_v + nv_all);
      P.nbb0 = (int32_t)(c->tot_b + nb_all);
      P.ncb = (int32_t)(c->tot_c + nc_all);
      nv_all += P.nv;
      nb_all += (int64_t)P.nv + P.nbb;
      nc_all += (int64_t)P.U + P.nbc;
    }
    if (nc_all == 0) break;
    {
      // the next level walks the edges of the polygons that go on refining
      int64_t walkers = 0;
      for (const PolyDev& P : c->pd) walkers += P.refining ? P.N : 0;
      extra_next = (double)walkers * 2.0 * std::max(m_level - 1.0, 0.0);
    }
    if (c->tot_v + nv_all > INT32_MAX || c->tot_b + nb_all > INT32_MAX || c->tot_c + nc_all > INT32_MAX)
      return set_err(c, GAHHG_E_LIMIT, "build: level %d exceeds int32 indexing", lv + 1);
    CK(ensure_v(c, c->tot_v, c->tot_v + nv_all));
    CK(ensure_b(c, c->tot_b, c->tot_b + nb_all));
    CK(ensure_c(c, c->tot_c, c->tot_c + nc_all));
    CK(pd_upload(c));
    g = view(c);
    hipLaunchKernelGGL(k_create_center, dim3(blocks_for((int64_t)GL.cnum * kBcLanes, kBlock)), dim3(kBlock), 0,
                       c->stream, g,
                       (const PolyDev*)c->d_pd, GL.cstart, GL.cnum, (const int32_t*)c->d_o0, (const int32_t*)c->d_o1);
    LAUNCH_CK("create center");
    hipLaunchKernelGGL(k_border_children, dim3(blocks_for((int64_t)GL.cnum * kBcLanes, kBlock)), dim3(kBlock), 0,
                       c->stream, g,
                       (const PolyDev*)c->d_pd, GL.cstart, GL.cnum, (const int32_t*)c->d_o2, (const int32_t*)c->d_o3);
    LAUNCH_CK("border children");
    const LevelDesc NL{(int32_t)c->tot_v, (int32_t)nv_all, (int32_t)c->tot_b, (int32_t)nb_all, (int32_t)c->tot_c,
                       (int32_t)nc_all, 0};
    hipLaunchKernelGGL(k_sc_partner, dim3(blocks_for(NL.cnum, kBlock)), dim3(kBlock), 0, c->stream, g,
                       (const PolyDev*)c->d_pd, NL.cstart, NL.cnum);
    LAUNCH_CK("border child partners");
    c->glv.push_back(NL);
    c->tot_v += nv_all;
    c->tot_b += nb_all;
    c->tot_c += nc_all;
    // advance every refining polygon to its new level
    for (int p = 0; p < np; p++) {
      PolyDev& P = c->pd[p];
      if (!P.refining) continue;
      c->plv[p].back().rnum = P.U;
      c->plv[p].push_back(LevelDesc{P.nvb, P.nv, P.nbb0, P.nv + P.nbb, P.ncb, P.U + P.nbc, 0});
      P.pbs = P.bs;
      P.plbs = P.lbs;
      P.lbs = P.lbs + P.bn;
      P.cs = P.ncb;
      P.cn = P.U + P.nbc;
      P.bs = P.nbb0;
      P.bn = P.nv + P.nbb;
    }
    CK(pd_upload(c));
  }
  for (int p = 0; p < np; p++)
    if ((c->pd[p].depth == 0 && maxd == 1) || c->pd[p].depth == 1) c->pgamma[p] = (float)(42 * 1.0 / c->pd[p].N);
  c->gnum = (int32_t)c->glv.size();

  // bottom segments
  for (int p = 0; p < np; p++) {
    PolyDev& P = c->pd[p];
    const LevelDesc& B = c->plv[p][P.L];
    P.bcs = B.cstart;
    P.bcn = B.cnum;
    P.refining = 0;
  }
  CK(pd_upload(c));

  // dispatchSELast for every polygon at its bottom level
  if (c->ecap < Nt) {
    dfree(c->d_ecnt);
    c->d_ecnt = nullptr;
    CK(dalloc(c, (void**)&c->d_ecnt, Nt * 4));
    c->ecap = Nt;
  }
  CK(ensure_flags(c, std::max<int64_t>(Nt, c->tot_c) + 1));
  g = view(c);
  // dispatch walks at each polygon's bottom level: the edges of polygons
  // whose bottom level has walk records in one launch, the others in another
  bool bottom_rec = false, bottom_plain = false;
  for (const PolyDev& P : c->pd) {
    const bool r = P.L < 32 && ((rec_levels >> P.L) & 1u);
    bottom_rec |= r;
    bottom_plain |= !r;
  }
  if (bottom_rec)
    hipLaunchKernelGGL((k_walk<kCount, true>), dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, g,
                       (const WalkRec*)c->d_wrec, rec_levels, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                       (const PolyDev*)c->d_pd, (int)Nt, 0, c->d_loc_cur, c->d_loc_prev, (const int32_t*)nullptr,
                       c->d_ecnt, (int32_t*)nullptr, c->d_st, 0, (int32_t*)nullptr, (const int32_t*)nullptr);
  if (bottom_plain)
    hipLaunchKernelGGL((k_walk<kCount, false>), dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, g,
                       (const WalkRec*)c->d_wrec, rec_levels, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                       (const PolyDev*)c->d_pd, (int)Nt, 0, c->d_loc_cur, c->d_loc_prev, (const int32_t*)nullptr,
                       c->d_ecnt, (int32_t*)nullptr, c->d_st, 0, (int32_t*)nullptr, (const int32_t*)nullptr);
  LAUNCH_CK("count walk");
  CK(scan_i32(c, c->d_ecnt, c->d_o0, Nt, nullptr, &c->d_st->R, 0));
  hipLaunchKernelGGL(k_poly_refs, dim3(blocks_for(np, kBlock)), dim3(kBlock), 0, c->stream, c->d_pd, np,
                     (const int32_t*)c->d_o0, (const int32_t*)c->d_ecnt);
  LAUNCH_CK("polygon refs");
  CK(sync_state(c));
  CK(pd_download(c));
  const int64_t R = c->h_st->R;
  if (R < 0) return set_err(c, GAHHG_E_LIMIT, "dispatch: ref count overflow");
  c->tot_r = R;
  CK(ensure_refs(c, R + 1));
  g = view(c);
  if (bottom_rec)
    hipLaunchKernelGGL((k_walk<kWrite, true>), dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, g,
                       (const WalkRec*)c->d_wrec, rec_levels, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                       (const PolyDev*)c->d_pd, (int)Nt, 0, c->d_loc_cur, c->d_loc_prev, (const int32_t*)c->d_o0,
                       (int32_t*)nullptr, c->d_rcell, c->d_st, 0, (int32_t*)nullptr, (const int32_t*)nullptr);
  if (bottom_plain)
    hipLaunchKernelGGL((k_walk<kWrite, false>), dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, g,
                       (const WalkRec*)c->d_wrec, rec_levels, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                       (const PolyDev*)c->d_pd, (int)Nt, 0, c->d_loc_cur, c->d_loc_prev, (const int32_t*)c->d_o0,
                       (int32_t*)nullptr, c->d_rcell, c->d_st, 0, (int32_t*)nullptr, (const int32_t*)nullptr);
  LAUNCH_CK("write walk");
  CK(fill_i32(c, c->d_rflag, R, 0));
  hipLaunchKernelGGL(k_edge_heads, dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, (const int32_t*)c->d_o0,
                     (const int32_t*)c->d_epoly, (const PolyDev*)c->d_pd, (int)Nt, c->d_rflag);
  LAUNCH_CK("edge heads");
  // R refs: every buffer indexed by ref slot holds at least R + 1 entries
  CK(scan_i32(c, c->d_rflag, c->d_rscan, R, nullptr, nullptr, 1));
  hipLaunchKernelGGL(k_ref_edges, dim3(blocks_for(R, kBlock)), dim3(kBlock), 0, c->stream, (const int32_t*)c->d_rcell,
                     (const int32_t*)c->d_rscan, (const int32_t*)c->g.c_poly, (const PolyDev*)c->d_pd,
                     (const DevState*)c->d_st, c->d_redge);
  LAUNCH_CK("ref edges");
  // counting sort into cell-major order
  const int64_t C = c->tot_c;
  CK(fill_i32(c, c->d_f1, C, 0));
  hipLaunchKernelGGL(k_count_cells, dim3(blocks_for(R, kBlock)), dim3(kBlock), 0, c->stream,
                     (const int32_t*)c->d_rcell, (const DevState*)c->d_st, c->d_f1);
  LAUNCH_CK("count cells");
  CK(scan_i32(c, c->d_f1, c->d_o1, C, nullptr, nullptr, 0));
  CK(copy_i32(c, c->d_f0, c->d_o1, C));
  hipLaunchKernelGGL(k_scatter_refs, dim3(blocks_for(R, kBlock)), dim3(kBlock), 0, c->stream,
                     (const int32_t*)c->d_rcell, (const int32_t*)c->d_redge, (const DevState*)c->d_st, c->d_f0,
                     c->d_rflag);
  LAUNCH_CK("scatter refs");
  hipLaunchKernelGGL(k_finish_cells, dim3(blocks_for(C, 4)), dim3(256), 0, c->stream, g, (const PolyDev*)c->d_pd,
                     (int)C, (const int32_t*)c->d_f1, (const int32_t*)c->d_o1, (const int32_t*)c->d_rflag);
  LAUNCH_CK("finish cells");
  for (int p = 0; p < np; p++) c->plv[p][c->pd[p].L].rnum = c->pd[p].R;
  CK(sync_state(c));
  return GAHHG_OK;
}


// ---------------------------------------------------------------------------
// Device-driven build: host orchestration
// ---------------------------------------------------------------------------
void drop_build_graph(gahhg_ctx* c) {
  DevBuildState& d = c->dvb;
  if (d.exec) (void)hipGraphExecDestroy(d.exec);
  if (d.graph) (void)hipGraphDestroy(d.graph);
  d.exec = nullptr;
  d.graph = nullptr;
  d.key.clear();
  d.graph_levels = 0;
}

void free_dev_build(gahhg_ctx* c) {
  drop_build_graph(c);
  DevBuildState& d = c->dvb;
  void* ptrs[] = {d.g0_v, d.g0_n, d.g0_b, d.g0_c, d.pd0, d.plv, d.plvn, d.gamma, d.pmarks, d.hist, d.csums, d.sscr};
  for (void* q : ptrs) dfree(q);
  if (d.hbuf) (void)hipHostFree(d.hbuf);
  d = DevBuildState();
}

namespace {

// Arena of one device-driven build: what each array holds.
struct DvArena {
  int64_t v, b, c, r;
  int end_bit;  // key bits of the ref sort (cells < 2^end_bit)
};

int dv_upload(gahhg_ctx* c, void* dst, const void* src, int64_t bytes) {
  return hip_check(c, hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice), "device build upload");
}

// The build's scan scratch for scans of up to n elements.
int dv_scan_scratch(gahhg_ctx* c, int64_t n) {
  DevBuildState& d = c->dvb;
  const int64_t need = scan_scratch_entries(n);
  if (d.sscr_cap >= need) return GAHHG_OK;
  CK_SYNC(c);
  dfree(d.sscr);
  d.sscr = nullptr;
  d.sscr_cap = 0;
  CK(dalloc(c, (void**)&d.sscr, need * 4));
  d.sscr_cap = need;
  return GAHHG_OK;
}

// Everything the enqueued build reads or writes, sized before anything is
// enqueued (nothing is allocated while the build runs or is recorded).
int dv_prepare(gahhg_ctx* c, const int64_t* want, DvArena* a) {
  DevBuildState& d = c->dvb;
  const int np = (int)c->polys.size();
  const int64_t Nt = c->n_edges;
  const int64_t V0 = (int64_t)kGrid0Verts * np, B0 = (int64_t)kGrid0Borders * np, C0 = (int64_t)kGrid0Cells * np;
  // the reference's initial capacities (createAHHGInit :1949-1953, summed
  // over the batch), or what the last build of this polygon set used plus
  // an eighth, or what an overflowed attempt asked for
  int64_t ev = std::max<int64_t>(Nt, 160) + V0, eb = std::max<int64_t>(4 * Nt, 480) + B0;
  int64_t ec = std::max<int64_t>(Nt, 84) + C0, er = std::max<int64_t>(2 * Nt, 1024);
  if (d.shape_n == Nt && d.shape_np == np) {
    ev = std::max(ev, d.used_v + d.used_v / 8);
    eb = std::max(eb, d.used_b + d.used_b / 8);
    ec = std::max(ec, d.used_c + d.used_c / 8);
    er = std::max(er, d.used_r + d.used_r / 8 + 1);
  }
  ev = std::max(ev, want[0]);
  eb = std::max(eb, want[1]);
  ec = std::max(ec, want[2]);
  er = std::max(er, want[3]);
  CK(ensure_v(c, 0, ev));
  CK(ensure_b(c, 0, eb));
  CK(ensure_c(c, 0, ec));
  CK(ensure_flags(c, std::max<int64_t>(Nt, c->g.ccap) + 1));
  CK(ensure_refs(c, er));
  // walk records for at most 4M cells (1.4 GB): deeper levels of larger
  // builds walk without them
  CK(ensure_wrec(c, 0, std::min<int64_t>(c->g.ccap, 1 << 22)));
  if (c->ecap < Nt) {
    dfree(c->d_ecnt);
    c->d_ecnt = nullptr;
    CK(dalloc(c, (void**)&c->d_ecnt, Nt * 4));
    c->ecap = Nt;
  }
  if (c->wslot_cap < Nt) {
    CK_SYNC(c);
    dfree(c->d_wslot);
    c->d_wslot = nullptr;
    CK(dalloc(c, (void**)&c->d_wslot, std::max<int64_t>(Nt, 1) * kSlotK * 4));
    c->wslot_cap = Nt;
  }
  a->v = c->g.vcap;
  a->b = c->g.bcap;
  a->c = c->g.ccap;
  a->r = std::min(c->r2cap, c->g.rcap);
  a->end_bit = 1;
  while ((1LL << a->end_bit) < a->c) a->end_bit++;
  // radix histograms and every scan's scratch, reserved now
  const int64_t hn = radix_hist_entries(a->r);
  if (d.hist_cap < hn) {
    CK_SYNC(c);
    dfree(d.hist);
    d.hist = nullptr;
    d.hist_cap = 0;
    CK(dalloc(c, (void**)&d.hist, hn * 4));
    d.hist_cap = hn;
  }
  CK(dv_scan_scratch(c, std::max<int64_t>(Nt, hn)));
  if (!d.csums) CK(dalloc(c, (void**)&d.csums, kScanColsSums * 4));
  // per-polygon arrays
  if (d.pcap < np) {
    CK_SYNC(c);
    void* ptrs[] = {d.pd0, d.plv, d.plvn, d.gamma, d.pmarks};
    for (void* q : ptrs) dfree(q);
    d.pd0 = nullptr, d.plv = nullptr, d.plvn = nullptr, d.gamma = nullptr, d.pmarks = nullptr;
    d.pcap = 0;
    d.pd0_host.clear();
    CK(dalloc(c, (void**)&d.pd0, (int64_t)np * (int64_t)sizeof(PolyDev)));
    CK(dalloc(c, (void**)&d.plv, (int64_t)np * kMaxLevels * (int64_t)sizeof(LevelDesc)));
    CK(dalloc(c, (void**)&d.plvn, (int64_t)np * 4));
    CK(dalloc(c, (void**)&d.gamma, (int64_t)np * 4));
    CK(dalloc(c, (void**)&d.pmarks, (int64_t)np * 4));
    d.pcap = np;
  }
  if (d.hcap < np) {
    CK_SYNC(c);
    if (d.hbuf) (void)hipHostFree(d.hbuf);
    d.hbuf = nullptr;
    d.hcap = 0;
    const size_t hb = (size_t)np * (sizeof(PolyDev) + kMaxLevels * sizeof(LevelDesc) + 12);
    if (hipHostMalloc((void**)&d.hbuf, hb, hipHostMallocDefault) != hipSuccess) {
      d.hbuf = nullptr;
      return set_err(c, GAHHG_E_NOMEM, "hipHostMalloc(%zu)", hb);
    }
    d.hcap = np;
  }
  // level-0 template, once per context
  if (!d.g0_v) {
    std::vector<double> v0(3 * kGrid0Verts), bn0(3 * kGrid0Borders);
    std::vector<int32_t> bi0(6 * kGrid0Borders), ci0(7 * kGrid0Cells);
    build_grid0(v0.data(), bi0.data(), bn0.data(), ci0.data());
    CK(dalloc(c, (void**)&d.g0_v, (int64_t)v0.size() * 8));
    CK(dalloc(c, (void**)&d.g0_n, (int64_t)bn0.size() * 8));
    CK(dalloc(c, (void**)&d.g0_b, (int64_t)bi0.size() * 4));
    CK(dalloc(c, (void**)&d.g0_c, (int64_t)ci0.size() * 4));
    CK(dv_upload(c, d.g0_v, v0.data(), (int64_t)v0.size() * 8));
    CK(dv_upload(c, d.g0_n, bn0.data(), (int64_t)bn0.size() * 8));
    CK(dv_upload(c, d.g0_b, bi0.data(), (int64_t)bi0.size() * 4));
    CK(dv_upload(c, d.g0_c, ci0.data(), (int64_t)ci0.size() * 4));
  }
  // polygon records at the start of a build (as build_hierarchy sets them)
  const int maxd = c->prm.max_depth;
  std::vector<PolyDev> t((size_t)np);
  int64_t e0 = 0;
  for (int p = 0; p < np; p++) {
    PolyDev& P = t[(size_t)p];
    memset(&P, 0, sizeof(P));
    P.e0 = (int32_t)e0;
    P.N = c->polys[p].N;
    P.cell0 = kGrid0Cells * p;
    P.refining = maxd > 1 ? 1 : 0;
    P.cs = kGrid0Cells * p;
    P.cn = kGrid0Cells;
    P.bs = kGrid0Borders * p;
    P.bn = kGrid0Borders;
    P.tbl = -1;
    P.depth = p < (int)c->depth_override.size() ? c->depth_override[p] : 0;
    if (P.depth == 1) P.refining = 0;
    e0 += P.N;
  }
  if (d.pd0_host.size() != t.size() || memcmp(d.pd0_host.data(), t.data(), t.size() * sizeof(PolyDev)) != 0) {
    CK_SYNC(c);  // a replay in flight may still read the old template
    CK(dv_upload(c, d.pd0, t.data(), (int64_t)t.size() * (int64_t)sizeof(PolyDev)));
    d.pd0_host = std::move(t);
  }
  return GAHHG_OK;
}

unsigned dv_grid(int64_t n) { return (unsigned)std::min<int64_t>(blocks_for(n, kBlock), kDvGrid); }

// Level 0 of every polygon, the state and the polygon records.
int dv_enqueue_start(gahhg_ctx* c) {
  DevBuildState& d = c->dvb;
  const int np = (int)c->polys.size();
  const int64_t Nt = c->n_edges;
  hipLaunchKernelGGL(k_edges, dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, (const double*)c->d_pv,
                     (const int32_t*)c->d_pe, (int)Nt, c->d_edge);
  hipLaunchKernelGGL(k_dv_start, dim3(1), dim3(kCtlBlock), 0, c->stream, c->d_st, c->d_pd, (const PolyDev*)d.pd0, np,
                     d.plv, d.plvn, d.pmarks, d.gamma);
  hipLaunchKernelGGL(k_dv_grid0, dim3(np), dim3(kCtlBlock), 0, c->stream, view(c), (const double*)d.g0_v,
                     (const double*)d.g0_n, (const int32_t*)d.g0_b, (const int32_t*)d.g0_c);
  CK(hip_check(c, hipGetLastError(), "device build start"));
  CK(fill_i32(c, c->d_loc_cur, Nt, 0));
  return fill_i32(c, c->d_loc_prev, Nt, 0);
}

// Iteration lv of createAHHG's level loop (mark, plan, density rule,
// create), every launch a no-op once the loop has ended.
int dv_enqueue_level(gahhg_ctx* c, const DvArena& a, int lv) {
  DevBuildState& d = c->dvb;
  const int np = (int)c->polys.size();
  const int64_t Nt = c->n_edges;
  const GridView g = view(c);
  const unsigned cg = dv_grid(a.c);
  hipLaunchKernelGGL(k_lv_begin, dim3(1), dim3(kCtlBlock), 0, c->stream, c->d_st, c->d_pd, np, lv,
                     (int)c->prm.max_depth, d.pmarks, c->wrec_cap);
  hipLaunchKernelGGL(k_walk_records_dv, dim3((unsigned)std::min<int64_t>(blocks_for(a.c, kRecBlock), kDvGrid)),
                     dim3(kRecBlock), 0, c->stream, g, (const DevState*)c->d_st, lv, c->d_wrec);
  hipLaunchKernelGGL(k_zero_dv, dim3(cg), dim3(kBlock), 0, c->stream, g, (const DevState*)c->d_st, lv);
  hipLaunchKernelGGL((k_walk<kMark, true>), dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, g,
                     (const WalkRec*)c->d_wrec, 0u, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                     (const PolyDev*)c->d_pd, (int)Nt, lv, c->d_loc_cur, c->d_loc_prev, (const int32_t*)nullptr,
                     c->d_ecnt, (int32_t*)nullptr, c->d_st, 1, c->d_wslot, (const int32_t*)nullptr);
  hipLaunchKernelGGL((k_walk<kMark, false>), dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, g,
                     (const WalkRec*)c->d_wrec, 0u, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                     (const PolyDev*)c->d_pd, (int)Nt, lv, c->d_loc_cur, c->d_loc_prev, (const int32_t*)nullptr,
                     c->d_ecnt, (int32_t*)nullptr, c->d_st, 1, c->d_wslot, (const int32_t*)nullptr);
  hipLaunchKernelGGL(k_plan_dv, dim3(cg), dim3(kBlock), 0, c->stream, g, (const DevState*)c->d_st, lv, c->d_f0,
                     c->d_f1, c->d_f2, c->d_f3);
  CK(hip_check(c, hipGetLastError(), "device build level"));
  ScanCols sc = {{c->d_f0, c->d_f1, c->d_f2, c->d_f3}, {c->d_o0, c->d_o1, c->d_o2, c->d_o3}, {nullptr, nullptr, nullptr, nullptr}};
  CK(scan_cols_dev(c, sc, 4, &c->d_st->lv[lv].cnum, 0, d.csums));
  hipLaunchKernelGGL(k_plan_totals, dim3(blocks_for(np, kBlock)), dim3(kBlock), 0, c->stream, c->d_pd, np, 0,
                     (const int32_t*)c->d_f0, (const int32_t*)c->d_f1, (const int32_t*)c->d_f2,
                     (const int32_t*)c->d_f3, (const int32_t*)c->d_o0, (const int32_t*)c->d_o1,
                     (const int32_t*)c->d_o2, (const int32_t*)c->d_o3, (const DevState*)c->d_st, lv);
  hipLaunchKernelGGL(k_lv_step, dim3(1), dim3(kCtlBlock), 0, c->stream, c->d_pd, np, c->d_st, lv,
                     c->prm.auto_depth != 0 ? 1 : 0, d.gamma, a.v, a.b, a.c);
  hipLaunchKernelGGL(k_create_center_dv, dim3(dv_grid(a.c * kBcLanes)), dim3(kBlock), 0, c->stream, g,
                     (const PolyDev*)c->d_pd,
                     (const DevState*)c->d_st, lv, (const int32_t*)c->d_o0, (const int32_t*)c->d_o1);
  hipLaunchKernelGGL(k_border_children_dv, dim3(dv_grid(a.c * kBcLanes)), dim3(kBlock), 0, c->stream, g,
                     (const PolyDev*)c->d_pd,
                     (const DevState*)c->d_st, lv, (const int32_t*)c->d_o2, (const int32_t*)c->d_o3);
  hipLaunchKernelGGL(k_sc_partner_dv, dim3(cg), dim3(kBlock), 0, c->stream, g, (const PolyDev*)c->d_pd,
                     (const DevState*)c->d_st, lv);
  hipLaunchKernelGGL(k_lv_advance, dim3(blocks_for(np, kBlock)), dim3(kBlock), 0, c->stream, c->d_pd, np,
                     (const DevState*)c->d_st, lv, d.plv, d.plvn);
  return hip_check(c, hipGetLastError(), "device build level");
}

// dispatchSELast (:1700-1784) for every polygon at its bottom level, first
// part: bottom segments, the count walk and the refs of each polygon (their
// total in st->R).
int dv_enqueue_count(gahhg_ctx* c) {
  DevBuildState& d = c->dvb;
  const int np = (int)c->polys.size();
  const int64_t Nt = c->n_edges;
  const GridView g = view(c);
  hipLaunchKernelGGL(k_dv_loop_end, dim3(1), dim3(1), 0, c->stream, c->d_st);
  hipLaunchKernelGGL(k_dv_bottom, dim3(blocks_for(np, kBlock)), dim3(kBlock), 0, c->stream, c->d_pd, np, c->d_st,
                     (const LevelDesc*)d.plv, d.gamma, (int)c->prm.max_depth);
  for (int w = 0; w < 2; w++) {
    if (w == 0)
      hipLaunchKernelGGL((k_walk<kCount, true>), dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, g,
                         (const WalkRec*)c->d_wrec, 0u, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                         (const PolyDev*)c->d_pd, (int)Nt, 0, c->d_loc_cur, c->d_loc_prev, (const int32_t*)nullptr,
                         c->d_ecnt, (int32_t*)nullptr, c->d_st, 1, (int32_t*)nullptr, (const int32_t*)nullptr);
    else
      hipLaunchKernelGGL((k_walk<kCount, false>), dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, g,
                         (const WalkRec*)c->d_wrec, 0u, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                         (const PolyDev*)c->d_pd, (int)Nt, 0, c->d_loc_cur, c->d_loc_prev, (const int32_t*)nullptr,
                         c->d_ecnt, (int32_t*)nullptr, c->d_st, 1, (int32_t*)nullptr, (const int32_t*)nullptr);
  }
  CK(hip_check(c, hipGetLastError(), "count walk"));
  CK(scan_i32_with(c, c->d_ecnt, c->d_o0, Nt, nullptr, &c->d_st->R, 0, d.sscr));
  hipLaunchKernelGGL(k_dv_poly_refs, dim3(blocks_for(np, kBlock)), dim3(kBlock), 0, c->stream, c->d_pd, np,
                     (const DevState*)c->d_st, (const int32_t*)c->d_o0, (const int32_t*)c->d_ecnt, d.plv);
  return hip_check(c, hipGetLastError(), "polygon refs");
}

// Second part: the check that the refs fit, the write walk, the refs in
// cell-major order, then the classification of `levels` levels.
int dv_enqueue_write(gahhg_ctx* c, const DvArena& a, int levels) {
  DevBuildState& d = c->dvb;
  const int64_t Nt = c->n_edges;
  const GridView g = view(c);
  const unsigned rg = dv_grid(a.r), cg = dv_grid(a.c);
  hipLaunchKernelGGL(k_dv_check_refs, dim3(1), dim3(1), 0, c->stream, c->d_st, a.r);
  // edges whose cells the walk slots hold are copied; the write walk walks
  // the others, listed in edge order
  hipLaunchKernelGGL(k_dv_slot_copy, dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream,
                     (const int32_t*)c->d_epoly, (const PolyDev*)c->d_pd, (int)Nt, (const int32_t*)c->d_wslot,
                     (const int32_t*)c->d_ecnt, (const int32_t*)c->d_o0, c->d_rcell, c->d_f0,
                     (const DevState*)c->d_st);
  CK(scan_i32_with(c, c->d_f0, c->d_f1, Nt, nullptr, &c->d_st->nwalk, 0, d.sscr));
  hipLaunchKernelGGL(k_dv_walk_list, dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream,
                     (const int32_t*)c->d_f0, (const int32_t*)c->d_f1, (int)Nt, c->d_f2, (const DevState*)c->d_st);
  for (int w = 0; w < 2; w++) {
    if (w == 0)
      hipLaunchKernelGGL((k_walk<kWrite, true>), dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, g,
                         (const WalkRec*)c->d_wrec, 0u, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                         (const PolyDev*)c->d_pd, (int)Nt, 0, c->d_loc_cur, c->d_loc_prev, (const int32_t*)c->d_o0,
                         (int32_t*)nullptr, c->d_rcell, c->d_st, 1, (int32_t*)nullptr, (const int32_t*)c->d_f2);
    else
      hipLaunchKernelGGL((k_walk<kWrite, false>), dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream, g,
                         (const WalkRec*)c->d_wrec, 0u, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                         (const PolyDev*)c->d_pd, (int)Nt, 0, c->d_loc_cur, c->d_loc_prev, (const int32_t*)c->d_o0,
                         (int32_t*)nullptr, c->d_rcell, c->d_st, 1, (int32_t*)nullptr, (const int32_t*)c->d_f2);
  }
  hipLaunchKernelGGL(k_fill_dv, dim3(rg), dim3(kBlock), 0, c->stream, c->d_rflag, (const int32_t*)&c->d_st->R_ok, 0,
                     (const DevState*)c->d_st);
  hipLaunchKernelGGL(k_dv_edge_heads, dim3(blocks_for(Nt, kBlock)), dim3(kBlock), 0, c->stream,
                     (const int32_t*)c->d_o0, (const int32_t*)c->d_epoly, (const PolyDev*)c->d_pd, (int)Nt,
                     c->d_rflag, (const DevState*)c->d_st);
  CK(hip_check(c, hipGetLastError(), "write walk"));
  ScanCols s1 = {{c->d_rflag, nullptr, nullptr, nullptr}, {c->d_rscan, nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr, nullptr}};
  CK(scan_cols_dev(c, s1, 1, &c->d_st->R_ok, 1, d.csums));
  hipLaunchKernelGGL(k_dv_ref_edges, dim3(rg), dim3(kBlock), 0, c->stream, (const int32_t*)c->d_rcell,
                     (const int32_t*)c->d_rscan, (const int32_t*)c->g.c_poly, (const PolyDev*)c->d_pd,
                     (const DevState*)c->d_st, c->d_redge);
  // the refs in cell-major order by a stable sort of (cell, edge) over the
  // edge-major refs, then refs per cell from the sorted runs and their
  // offsets by a scan
  uint32_t *kres = nullptr, *vres = nullptr;
  CK(radix_sort_pairs(c, (uint32_t*)c->d_rcell, (uint32_t*)c->d_redge, (uint32_t*)c->d_rflag, (uint32_t*)c->d_rscan,
                      &c->d_st->R_ok, a.r, a.end_bit, d.hist, &kres, &vres, d.sscr));
  hipLaunchKernelGGL(k_fill_dv, dim3(cg), dim3(kBlock), 0, c->stream, c->d_f0, (const int32_t*)&c->d_st->tot_c32, 0,
                     (const DevState*)c->d_st);
  hipLaunchKernelGGL(k_fill_dv, dim3(cg), dim3(kBlock), 0, c->stream, c->d_f1, (const int32_t*)&c->d_st->tot_c32, 0,
                     (const DevState*)c->d_st);
  hipLaunchKernelGGL(k_dv_cell_runs, dim3(rg), dim3(kBlock), 0, c->stream, (const uint32_t*)kres,
                     (const DevState*)c->d_st, c->d_f0, c->d_f1);
  hipLaunchKernelGGL(k_dv_run_counts, dim3(cg), dim3(kBlock), 0, c->stream, (const int32_t*)c->d_f0, c->d_f1,
                     (const DevState*)c->d_st);
  CK(hip_check(c, hipGetLastError(), "ref counts"));
  ScanCols s2 = {{c->d_f1, nullptr, nullptr, nullptr}, {c->d_o1, nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr, nullptr}};
  CK(scan_cols_dev(c, s2, 1, &c->d_st->tot_c32, 0, d.csums));
  hipLaunchKernelGGL(k_copy_dv, dim3(rg), dim3(kBlock), 0, c->stream, c->g.refs, (const int32_t*)vres,
                     (const int32_t*)&c->d_st->R_ok, (const DevState*)c->d_st);
  hipLaunchKernelGGL(k_dv_finish_cells, dim3(cg), dim3(kBlock), 0, c->stream, g, (const PolyDev*)c->d_pd,
                     (const DevState*)c->d_st, (const int32_t*)c->d_f1, (const int32_t*)c->d_o1);
  CK(hip_check(c, hipGetLastError(), "cell refs"));
  return classify_dev_enqueue(c, levels);
}

// What a recorded build depends on: every pointer and size baked into it.
std::vector<uintptr_t> dv_key(gahhg_ctx* c, const DvArena& a, int levels) {
  const DevBuildState& d = c->dvb;
  const DeviceGrid& g = c->g;
  std::vector<uintptr_t> k = {
      (uintptr_t)c->polys.size(), (uintptr_t)c->n_edges, (uintptr_t)c->prm.max_depth, (uintptr_t)c->prm.auto_depth,
      (uintptr_t)levels, (uintptr_t)a.v, (uintptr_t)a.b, (uintptr_t)a.c, (uintptr_t)a.r, (uintptr_t)a.end_bit,
      (uintptr_t)c->wrec_cap, (uintptr_t)c->fcap, (uintptr_t)d.hist_cap,
      (uintptr_t)c->stream, (uintptr_t)c->d_pv, (uintptr_t)c->d_pe, (uintptr_t)c->d_edge, (uintptr_t)c->d_epoly,
      (uintptr_t)c->d_pd, (uintptr_t)c->d_st, (uintptr_t)c->d_loc_cur, (uintptr_t)c->d_loc_prev,
      (uintptr_t)c->d_f0, (uintptr_t)c->d_f1, (uintptr_t)c->d_f2, (uintptr_t)c->d_f3, (uintptr_t)c->d_o0,
      (uintptr_t)c->d_o1, (uintptr_t)c->d_o2, (uintptr_t)c->d_o3, (uintptr_t)c->d_ecnt, (uintptr_t)c->d_wslot,
      (uintptr_t)c->d_rcell, (uintptr_t)c->d_rflag, (uintptr_t)c->d_redge, (uintptr_t)c->d_rscan,
      (uintptr_t)c->d_wrec, (uintptr_t)d.g0_v, (uintptr_t)d.g0_n, (uintptr_t)d.g0_b, (uintptr_t)d.g0_c,
      (uintptr_t)d.pd0, (uintptr_t)d.plv, (uintptr_t)d.plvn, (uintptr_t)d.gamma, (uintptr_t)d.pmarks,
      (uintptr_t)d.hist, (uintptr_t)d.csums, (uintptr_t)d.sscr, (uintptr_t)d.sscr_cap, (uintptr_t)g.v, (uintptr_t)g.b_parent, (uintptr_t)g.b_partner,
      (uintptr_t)g.b_v0, (uintptr_t)g.b_v1, (uintptr_t)g.b_subcell, (uintptr_t)g.b_id, (uintptr_t)g.b_n,
      (uintptr_t)g.c_bstart, (uintptr_t)g.c_bnum, (uintptr_t)g.c_rstart, (uintptr_t)g.c_rnum, (uintptr_t)g.c_upcell,
      (uintptr_t)g.c_subcell, (uintptr_t)g.c_type, (uintptr_t)g.c_poly, (uintptr_t)g.c_root, (uintptr_t)g.refs};
  return k;
}

// Levels to enqueue before the first look at whether the loop still runs:
// those of the last build of this polygon set, else an estimate (a level
// per factor 4 of edges per polygon; the density rule stops near there).
int dv_depth_hint(gahhg_ctx* c) {
  const DevBuildState& d = c->dvb;
  const int maxd = c->prm.max_depth;
  if (d.shape_n == c->n_edges && d.shape_np == (int)c->polys.size() && d.used_levels > 0)
    return std::min(maxd, d.used_levels);
  int64_t nmax = 1;
  for (const Polygon& p : c->polys) nmax = std::max<int64_t>(nmax, p.N);
  int h = 2;
  while ((1LL << (2 * (h - 1))) < nmax) h++;
  return std::min(maxd, h);
}

}  // namespace

int build_classify_dev(gahhg_ctx* c, bool* replayed, int32_t* attempts) {
  DevBuildState& d = c->dvb;
  const int np = (int)c->polys.size();
  const int maxd = c->prm.max_depth;
  *replayed = false;
  *attempts = 0;
  int64_t want[4] = {0, 0, 0, 0};
  DvArena a{};
  int levels = 0;
  bool done = false;
  for (int attempt = 0; attempt < 16 && !done; attempt++) {
    *attempts = attempt + 1;
    *replayed = false;
    CK(dv_prepare(c, want, &a));
    const int glv = d.graph_levels;
    std::vector<uintptr_t> key = dv_key(c, a, glv);
    if (d.exec && glv > 0 && key == d.key) {
      CK(hip_check(c, hipGraphLaunch(d.exec, c->stream), "build graph launch"));
      levels = glv;
      *replayed = true;
    } else {
      // eager: levels up to the hint, then one word read back per further
      // pair of levels until the loop has ended
      CK(dv_enqueue_start(c));
      int lv = 0, upto = dv_depth_hint(c);
      for (;;) {
        for (; lv < upto; lv++) CK(dv_enqueue_level(c, a, lv));
        if (lv >= maxd) break;
        int32_t run = 0;
        CK(hip_check(c, hipMemcpyAsync(&c->h_st->run, &c->d_st->run, 4, hipMemcpyDeviceToHost, c->stream),
                     "loop state"));
        CK_SYNC(c);
        run = c->h_st->run;
        if (!run) break;
        upto = std::min(maxd, lv + 2);
      }
      levels = lv;
      CK(dv_enqueue_count(c));
      // the refs' total is known now: grow the ref arrays to it here rather
      // than let the device stop the build short
      CK(hip_check(c, hipMemcpyAsync(&c->h_st->R, &c->d_st->R, 4, hipMemcpyDeviceToHost, c->stream), "ref count"));
      CK(hip_check(c, hipMemcpyAsync(&c->h_st->abort, &c->d_st->abort, 4, hipMemcpyDeviceToHost, c->stream),
                   "ref count"));
      CK_SYNC(c);
      if (!c->h_st->abort && c->h_st->R >= 0 && (int64_t)c->h_st->R + 1 > a.r) {
        const int64_t R = (int64_t)c->h_st->R + 1;
        CK(ensure_refs(c, R + R / 8));
        a.r = std::min(c->r2cap, c->g.rcap);
        const int64_t hn = radix_hist_entries(a.r);
        if (d.hist_cap < hn) {
          dfree(d.hist);
          d.hist = nullptr;
          d.hist_cap = 0;
          CK(dalloc(c, (void**)&d.hist, hn * 4));
          d.hist_cap = hn;
        }
        CK(dv_scan_scratch(c, std::max<int64_t>(c->n_edges, hn)));
      }
      CK(dv_enqueue_write(c, a, levels));
    }
    // the state and the per-polygon results, one wait for all
    char* hb = d.hbuf;
    CK(hip_check(c, hipMemcpyAsync(c->h_st, c->d_st, sizeof(DevState), hipMemcpyDeviceToHost, c->stream), "readback"));
    CK(hip_check(c, hipMemcpyAsync(hb, c->d_pd, (size_t)np * sizeof(PolyDev), hipMemcpyDeviceToHost, c->stream),
                 "readback"));
    hb += (size_t)np * sizeof(PolyDev);
    CK(hip_check(c, hipMemcpyAsync(hb, d.plv, (size_t)np * kMaxLevels * sizeof(LevelDesc), hipMemcpyDeviceToHost,
                                   c->stream),
                 "readback"));
    hb += (size_t)np * kMaxLevels * sizeof(LevelDesc);
    CK(hip_check(c, hipMemcpyAsync(hb, d.plvn, (size_t)np * 4, hipMemcpyDeviceToHost, c->stream), "readback"));
    CK(hip_check(c, hipMemcpyAsync(hb + (size_t)np * 4, d.pmarks, (size_t)np * 4, hipMemcpyDeviceToHost, c->stream),
                 "readback"));
    CK(hip_check(c, hipMemcpyAsync(hb + (size_t)np * 8, d.gamma, (size_t)np * 4, hipMemcpyDeviceToHost, c->stream),
                 "readback"));
    CK_SYNC(c);
    const DevState& s = *c->h_st;
    if (s.limit) return set_err(c, GAHHG_E_LIMIT, "build: a level or the refs exceed int32 indexing");
    if (s.overflow) {
      // grow what the attempt ran out of and start again
      drop_build_graph(c);
      // (a level's need is a lower bound of the hierarchy's: the deeper
      // levels grow it further, so the arena grows four-fold past it)
      if (s.need_v > a.v || s.need_b > a.b || s.need_c > a.c) {
        want[0] = std::max(want[0], 4 * s.need_v);
        want[1] = std::max(want[1], 4 * s.need_b);
        want[2] = std::max(want[2], 4 * s.need_c);
      }
      want[3] = std::max(want[3], s.need_r + s.need_r / 8);
      continue;
    }
    if (s.abort) {
      // a replay whose level loop needed more levels than it recorded
      drop_build_graph(c);
      d.used_levels = maxd;
      continue;
    }
    done = true;
  }
  if (!done) return set_err(c, GAHHG_E_STATE, "build: the arena did not converge");
  // host mirrors: levels, polygon records and their level segments
  const DevState& s = *c->h_st;
  const PolyDev* hpd = (const PolyDev*)d.hbuf;
  const LevelDesc* hplv = (const LevelDesc*)(d.hbuf + (size_t)np * sizeof(PolyDev));
  const int32_t* hplvn = (const int32_t*)((const char*)hplv + (size_t)np * kMaxLevels * sizeof(LevelDesc));
  c->pd.assign(hpd, hpd + np);
  c->plv.assign((size_t)np, std::vector<LevelDesc>());
  for (int p = 0; p < np; p++) c->plv[p].assign(hplv + (int64_t)p * kMaxLevels, hplv + (int64_t)p * kMaxLevels + hplvn[p]);
  c->pmarks.assign(hplvn + np, hplvn + 2 * np);
  const float* hg = (const float*)(hplvn + 2 * np);
  c->pgamma.assign(hg, hg + np);
  c->glv.assign(s.lv, s.lv + s.gnum);
  c->gnum = s.gnum;
  c->tot_v = s.tot_v;
  c->tot_b = s.tot_b;
  c->tot_c = s.tot_c;
  c->tot_r = s.R;
  c->und_snap = (int64_t)s.und_first;
  c->und_final = (int64_t)s.undefined_leaves;
  // a polygon set built before by this context (same shape) is likely to be
  // built again: only then is the build recorded, so that a one-off build
  // does not pay for the capture and instantiation
  const bool again = d.shape_n == c->n_edges && d.shape_np == np && d.used_levels > 0;
  d.used_v = s.tot_v;
  d.used_b = s.tot_b;
  d.used_c = s.tot_c;
  d.used_r = (int64_t)s.R + 1;
  d.used_levels = s.gnum;
  d.shape_n = c->n_edges;
  d.shape_np = np;
  // record the build for the next one of this polygon set (the levels it
  // needed; a polygon set that needs more stops the replay short, and the
  // next build runs eagerly and records again)
  if (!*replayed && c->prm.build_graph && again) {
    drop_build_graph(c);
    const int gl = s.gnum;
    hipGraph_t gr = nullptr;
    hipGraphExec_t ex = nullptr;
    CK(hip_check(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "build capture"));
    int rc = dv_enqueue_start(c);
    for (int lv = 0; lv < gl && !rc; lv++) rc = dv_enqueue_level(c, a, lv);
    if (!rc) rc = dv_enqueue_count(c);
    if (!rc) rc = dv_enqueue_write(c, a, gl);
    const hipError_t e = hipStreamEndCapture(c->stream, &gr);
    if (!rc) rc = hip_check(c, e, "build capture end");
    if (!rc) rc = hip_check(c, hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0), "build graph instantiate");
    if (rc) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    d.graph = gr;
    d.exec = ex;
    d.graph_levels = gl;
    d.key = dv_key(c, a, gl);
  }
  return GAHHG_OK;
}

// Loads this file's code object onto the current device (context set-up).
void load_code_build() {
  hipFuncAttributes a;
  (void)hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_zero));
}

}  // namespace gahhg
