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

struct WalkDiag {
  unsigned long long steps, partner_missing, cap_hits, unlocated, bad_index;
};

enum { kMark = 0, kCount = 1, kWrite = 2 };

// One polygon edge across the cells of one level: markNECell (:1153-1289,
// MODE kMark), calcRefNum (:685-818, kCount), writeRef (:820-955, kWrite).
template <int MODE>
__device__ void walk_edge(const GridView& g, const double* __restrict__ edge, int e, int level,
                          int ncell0, int32_t* buf, int32_t* out_cells, int nb_valid,
                          WalkDiag& d) {
  const double* er = edge + 9 * (int64_t)e;
  const V3 p0 = load3(er), p1 = load3(er + 3), ne = load3(er + 6);
  int cur = -1;
  if (level == 0) {
    for (int i = 0; i < ncell0; i++) {
      cur = inside_cell(g, p0, i);
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
    cur = sc >= 0 ? inside_cell(g, p0, sc) : -1;
    if (cur == -1) {
      const int pbs = g.c_bstart[par], pbn = g.c_bnum[par];
      for (int i = 0; i < pbn; i++) {
        cur = g.b_subcell[pbs + i];
        if (cur == -1) continue;
        cur = inside_cell(g, p0, cur);
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
    if (MODE == kMark) g.c_rnum[cur] = kMarked;
    if (MODE == kWrite) out_cells[steps] = cur;
    steps++;
    int intertype = 0, info = 0;
    const int out = seg_exit_border(g, ne, p0, p1, cur, in, &intertype, &info);
    if (out == -1) break;
    if (intertype == kInterX || intertype == kInterT1) {
      const int pt = g.b_partner[out];
      if (pt == -1) {
        d.partner_missing++;
        break;
      }
      in = pt;
      cur = g.b_parent[in];
    } else if (intertype == kInterT2) {
      // step a travel-epsilon past the grid vertex and test the two cells
      // that share it (:753-792)
      V3 cp = vtx(g, info == 0 ? g.b_v0[out] : g.b_v1[out]);
      V3 ed = unit(sub(p1, p0));
      cp = add(cp, mul(ed, kTravelEps));
      const int u = g.c_bnum[cur];
      const int pt = g.b_partner[out];
      if (pt != -1) {
        int r = inside_cell(g, cp, g.b_parent[pt]);
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
        int r = inside_cell(g, cp, g.b_parent[po]);
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
        // markNECell derives `in` from the new cell's bstart (:1279-1285);
        // the resulting index is kept, and treated as -1 outside the array.
        const int nbs = g.c_bstart[nc];
        const int kk = t > 0 ? (out - nbs + 1) % u : (out - nbs + u - 1) % u;
        const int idx = nbs + kk;
        if (idx < 0 || idx >= nb_valid) {
          d.bad_index++;
          in = -1;
        } else {
          in = g.b_partner[idx];
        }
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

template <int MODE>
__global__ void __launch_bounds__(kBlock) k_walk(GridView g, const double* __restrict__ edge, int N,
                                                 int level, int ncell0, int32_t* buf,
                                                 int32_t* buf_prev, const int32_t* off,
                                                 int32_t* cnt_out, int32_t* rcell, DevState* st) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  // borders that exist while this level is walked (marking OVERLAP guard)
  const int nb_valid = st->lv[level].bstart + st->lv[level].bnum;
  WalkDiag d = {0, 0, 0, 0, 0};
  if (e < N) {
    if (MODE == kMark) buf_prev[e] = buf[e];
    int32_t* oc = (MODE == kWrite) ? rcell + off[e] : nullptr;
    walk_edge<MODE>(g, edge, e, level, ncell0, buf, oc, nb_valid, d);
    if (MODE == kCount) cnt_out[e] = (int32_t)d.steps;
  }
  unsigned long long s = wave_sum_u64(d.steps);
  unsigned long long pm = wave_sum_u64(d.partner_missing);
  unsigned long long ch = wave_sum_u64(d.cap_hits);
  unsigned long long un = wave_sum_u64(d.unlocated);
  unsigned long long bi = wave_sum_u64(d.bad_index);
  if ((threadIdx.x & 63) == 0) {
    if (MODE == kMark && s) atomicAdd(&st->steps_mark, s);
    if (MODE == kCount && s) atomicAdd(&st->steps_dispatch, s);
    if (MODE != kWrite) {  // the write walk repeats the count walk exactly
      if (pm) atomicAdd(&st->partner_missing, pm);
      if (ch) atomicAdd(&st->cap_hits, ch);
      if (un) atomicAdd(&st->unlocated, un);
      if (bi) atomicAdd(&st->bad_index, bi);
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

// kernel_DivGridFast_zeroRefCount, :1127-1150
__global__ void k_zero(GridView g, const DevState* st, int lv) {
  const LevelDesc L = st->lv[lv];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.cnum) return;
  const int cid = L.cstart + t;
  g.c_rnum[cid] = 0;
  if (L.cstart == 0) {
    g.c_rstart[cid] = -1;
    g.c_subcell[cid] = -1;
    g.c_upcell[cid] = -1;
    g.c_type[cid] = kUndefined;
    const int bs = g.c_bstart[cid], bn = g.c_bnum[cid];
    for (int i = 0; i < bn; i++) g.b_subcell[bs + i] = -1;
  }
}

// kernel_DivGrid_calcCenterOffsets flags, :381-404
__global__ void k_refine_flags(GridView g, const DevState* st, int lv, int32_t* f0, int32_t* f1) {
  const LevelDesc L = st->lv[lv];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.cnum) return;
  const int cid = L.cstart + t;
  const bool r = g.c_rnum[cid] != 0;
  f0[t] = r ? g.c_bnum[cid] : 0;
  f1[t] = r ? 1 : 0;
}

// kernel_DivGrid_createCenter, :406-470
__global__ void k_create_center(GridView g, const DevState* st, int lv, const int32_t* o0,
                                const int32_t* o1) {
  const LevelDesc L = st->lv[lv];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.cnum) return;
  const int pid = L.cstart + t;
  if (g.c_rnum[pid] == 0) return;
  const int vbase = L.vstart + L.vnum;
  const int bbase = L.bstart + L.bnum;
  const int cbase = L.cstart + L.cnum;
  const int bs = g.c_bstart[pid], bn = g.c_bnum[pid];
  V3 center;
  if (bn == 4) {
    center = mul(add(vtx(g, g.b_v0[bs]), vtx(g, g.b_v1[bs])), 0.5);
  } else {
    center = v3(0.0, 0.0, 0.0);
    for (int i = 0; i < bn; i++) center = add(center, vtx(g, g.b_v0[bs + i]));
    center = divk(center, bn);
  }
  center = unit(center);
  const int vo = vbase + o0[t];
  const int bo = bbase + o0[t];
  const int child = cbase + o1[t];
  double* vw = const_cast<double*>(g.v);
  for (int i = 0; i < bn; i++) store3(vw + 3 * (int64_t)(vo + i), unit(mul(add(vtx(g, g.b_v0[bs + i]), center), 0.5)));
  for (int i = 0; i < bn; i++) {
    const int b = bo + i;
    g.b_id[b] = i;
    g.b_parent[b] = child;
    g.b_partner[b] = -1;
    g.b_subcell[b] = -1;
    g.b_v0[b] = vo + i;
    g.b_v1[b] = vo + (i + 1) % bn;
    store3(g.b_n + 3 * (int64_t)b, arc_normal(vtx(g, vo + i), vtx(g, vo + (i + 1) % bn)));
  }
  g.c_bstart[child] = bo;
  g.c_bnum[child] = bn;
  g.c_rstart[child] = -1;
  g.c_rnum[child] = -1;
  g.c_upcell[child] = pid;
  g.c_subcell[child] = -1;
  g.c_type[child] = kUndefined;
  g.c_subcell[pid] = child;
}

// Which border child (if any) centre-child border b produces:
// 6 = hexagon shared with the refined neighbour (created by the lower id),
// 4 = quad against an unrefined or missing neighbour, 0 = none.
// kernel_DivGrid_calcBorderSCOffsets, :473-509
__device__ __forceinline__ int border_child_kind(const GridView& g, int b, int* upedge_out) {
  const int P = g.c_upcell[g.b_parent[b]];
  const int id = g.b_id[b];
  const int ue = g.c_bstart[P] + id;
  *upedge_out = ue;
  if (g.c_bnum[P] == 4 && id == 0) return 0;
  const int pt = g.b_partner[ue];
  if (pt != -1 && g.c_subcell[g.b_parent[pt]] != -1 && P < g.b_parent[pt]) return 6;
  if (pt == -1 || g.c_subcell[g.b_parent[pt]] == -1) return 4;
  return 0;
}

// t indexes the centre-child borders of level lv+1: [bbase, bbase + nv)
__global__ void k_bsc_flags(GridView g, int bbase, int nv, int32_t* f0, int32_t* f1) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nv) return;
  int ue;
  const int k = border_child_kind(g, bbase + t, &ue);
  f0[t] = k;
  f1[t] = k ? 1 : 0;
}

// kernel_DivGrid_createBorderSC, :512-607
__global__ void k_bsc_create(GridView g, int bbase, int nv, int cbase, int U, const int32_t* o0,
                             const int32_t* o1) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nv) return;
  const int b = bbase + t;
  int ue;
  const int kind = border_child_kind(g, b, &ue);
  if (!kind) return;
  const int eo = bbase + nv + o0[t];
  const int co = cbase + U + o1[t];
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
  for (int i = 0; i < kind; i++) {
    const int nb = eo + i;
    const int a = vv[i], z = vv[(i + 1) % kind];
    g.b_id[nb] = i;
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
  g.c_upcell[co] = g.b_parent[ue];
  g.c_type[co] = kUndefined;
  g.b_subcell[ue] = co;
  if (kind == 6) g.b_subcell[g.b_partner[ue]] = co;
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

// kernel_DivGrid_setSCPartner, :610-682 (border children [start, start+c0))
__global__ void k_sc_partner(GridView g, int start, const DevState* st) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= st->c0) return;
  const int X = start + t;
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

// Level ranges of level lv+1 once its border children exist.
__global__ void k_level_next(DevState* st, int lv, int vbase, int nv, int bbase, int cbase, int U) {
  LevelDesc n;
  n.vstart = vbase;
  n.vnum = nv;
  n.bstart = bbase;
  n.bnum = nv + st->e0;
  n.cstart = cbase;
  n.cnum = U + st->c0;
  n.rnum = 0;
  st->lv[lv + 1] = n;
  st->lv[lv].rnum = U;
}

// --- dispatch (:1700-1784) ---
// head flags at the first ref of every edge except edge 0 (:971-981); an
// edge without refs shares its slot with the next edge.
__global__ void k_edge_heads(const int32_t* off, int N, const DevState* st, int32_t* flag) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < 1 || e >= N) return;
  const int o = off[e];
  if (o < st->R) flag[o] = 1;
}

__global__ void k_count_cells(const int32_t* rcell, const DevState* st, int cstart, int32_t* ccnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st->R) return;
  atomicAdd(&ccnt[rcell[i] - cstart], 1);
}

__global__ void k_scatter_refs(const int32_t* rcell, const int32_t* redge, const DevState* st,
                               int cstart, int32_t* fill, int32_t* refs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st->R) return;
  const int pos = atomicAdd(&fill[rcell[i] - cstart], 1);
  refs[pos] = redge[i];
}

// kernel_EC2CS_setCEInfo (:1030-1043) plus the ascending order of each run.
// One wave per cell: each lane ranks its elements against the whole run
// (stable by position), staged in LDS when the run fits.
constexpr int kSortCap = 1024;
__global__ void __launch_bounds__(256) k_finish_cells(GridView g, int cstart, int cnum, const int32_t* ccnt,
                                                      const int32_t* coff, const int32_t* __restrict__ src) {
  __shared__ int32_t lds[4][kSortCap];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + w;
  int n = 0, s = 0;
  if (t < cnum) {
    n = ccnt[t];
    s = coff[t];
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
  if (t < cnum && lane == 0) {
    g.c_rnum[cstart + t] = n;
    g.c_rstart[cstart + t] = s;
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
      grow(c, &g.c_type, used, need, nc))
    return GAHHG_E_NOMEM;
  g.ccap = nc;
  return GAHHG_OK;
}

static int ensure_flags(gahhg_ctx* c, int64_t need) {
  if (need <= c->fcap) return GAHHG_OK;
  CK_SYNC(c);
  int64_t nc = next_cap(c->fcap, need);
  dfree(c->d_f0);
  dfree(c->d_f1);
  dfree(c->d_o0);
  dfree(c->d_o1);
  c->d_f0 = c->d_f1 = c->d_o0 = c->d_o1 = nullptr;
  if (dalloc(c, (void**)&c->d_f0, nc * 4) || dalloc(c, (void**)&c->d_f1, nc * 4) ||
      dalloc(c, (void**)&c->d_o0, nc * 4) || dalloc(c, (void**)&c->d_o1, nc * 4))
    return GAHHG_E_NOMEM;
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
  dfree(c->g.refs);
  c->d_rcell = c->d_rflag = c->d_redge = c->g.refs = nullptr;
  if (dalloc(c, (void**)&c->d_rcell, nc * 4) || dalloc(c, (void**)&c->d_rflag, nc * 4) ||
      dalloc(c, (void**)&c->d_redge, nc * 4) || dalloc(c, (void**)&c->g.refs, nc * 4))
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

#define CK(x)                   \
  do {                          \
    int _rc = (x);              \
    if (_rc != GAHHG_OK) return _rc; \
  } while (0)
#define LAUNCH_CK(name) CK(hip_check(c, hipGetLastError(), name))

// GPISP::createAHHGInit + createAHHG (:1947-2015) + dispatchSELast.
int build_hierarchy(gahhg_ctx* c) {
  const int N = c->poly.N;
  const int maxd = c->prm.max_depth;
  const bool autod = c->prm.auto_depth != 0;

  // level 0 grid
  std::vector<double> v0(3 * kGrid0Verts), bn0(3 * kGrid0Borders);
  std::vector<int32_t> bi0(6 * kGrid0Borders), ci0(7 * kGrid0Cells);
  build_grid0(v0.data(), bi0.data(), bn0.data(), ci0.data());

  // capacities (createAHHGInit :1949-1953)
  CK(ensure_v(c, 0, std::max(N, 160)));
  CK(ensure_b(c, 0, std::max<int64_t>(4LL * N, 480)));
  CK(ensure_c(c, 0, std::max(N, 84)));
  CK(ensure_flags(c, std::max<int64_t>(4LL * N, 480)));

  {  // upload grid-0 (copyGrid0HostToDevice, :1532-1555) as SoA
    std::vector<int32_t> col(kGrid0Borders);
    auto up_b = [&](int32_t* dst, int k) {
      for (int i = 0; i < kGrid0Borders; i++) col[i] = bi0[6 * i + k];
      return hipMemcpy(dst, col.data(), kGrid0Borders * 4, hipMemcpyHostToDevice);
    };
    std::vector<int32_t> colc(kGrid0Cells);
    auto up_c = [&](int32_t* dst, int k) {
      for (int i = 0; i < kGrid0Cells; i++) colc[i] = ci0[7 * i + k];
      return hipMemcpy(dst, colc.data(), kGrid0Cells * 4, hipMemcpyHostToDevice);
    };
    hipError_t s = hipMemcpy(c->g.v, v0.data(), v0.size() * 8, hipMemcpyHostToDevice);
    if (s == hipSuccess) s = hipMemcpy(c->g.b_n, bn0.data(), bn0.size() * 8, hipMemcpyHostToDevice);
    int32_t* bd[6] = {c->g.b_parent, c->g.b_partner, c->g.b_v0, c->g.b_v1, c->g.b_subcell, c->g.b_id};
    for (int k = 0; k < 6 && s == hipSuccess; k++) s = up_b(bd[k], k);
    int32_t* cd[7] = {c->g.c_bstart, c->g.c_bnum, c->g.c_rstart, c->g.c_rnum, c->g.c_upcell, c->g.c_subcell, c->g.c_type};
    for (int k = 0; k < 7 && s == hipSuccess; k++) s = up_c(cd[k], k);
    CK(hip_check(c, s, "grid0 upload"));
  }

  // per-edge records (p0, p1, plane normal) for every walk and leaf test
  hipLaunchKernelGGL(k_edges, dim3(blocks_for(N, kBlock)), dim3(kBlock), 0, c->stream,
                     (const double*)c->d_pv, (const int32_t*)c->d_pe, N, c->d_edge);
  LAUNCH_CK("edge records");

  memset(c->lv, 0, sizeof(c->lv));
  c->lv[0] = LevelDesc{0, kGrid0Verts, 0, kGrid0Borders, 0, kGrid0Cells, 0};
  memset(c->h_st, 0, sizeof(DevState));
  c->h_st->lv[0] = c->lv[0];
  CK(hip_check(c, hipMemcpy(c->d_st, c->h_st, sizeof(DevState), hipMemcpyHostToDevice), "state init"));
  CK(fill_i32(c, c->d_loc_cur, N, 0));
  CK(fill_i32(c, c->d_loc_prev, N, 0));

  CK(hip_check(c, hipEventRecord(c->ev[0], c->stream), "event"));

  c->gnum = maxd;
  c->gamma = 0.f;
  bool early = false;
  GridView g = view(c);
  for (int lv = 0; lv < c->gnum - 1; lv++) {
    if (c->prm.verbose) printf("Subdivide level %d on GPU.\n", lv);
    // upper bound of this level's cell count (exact for level 0)
    int64_t cbound = c->lv[lv].cnum;
    // cell flags of this level, centre-border flags (<= 6 per cell) and the
    // next level's cell flags (<= 7 per cell) share these buffers
    CK(ensure_flags(c, 7 * cbound + 1));
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(cbound, kBlock)), dim3(kBlock), 0, c->stream, g, c->d_st, lv);
    LAUNCH_CK("zero");
    hipLaunchKernelGGL(k_walk<kMark>, dim3(blocks_for(N, kBlock)), dim3(kBlock), 0, c->stream, g,
                       (const double*)c->d_edge, N, lv, kGrid0Cells, c->d_loc_cur, c->d_loc_prev,
                       (const int32_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr, c->d_st);
    LAUNCH_CK("mark walk");
    hipLaunchKernelGGL(k_refine_flags, dim3(blocks_for(cbound, kBlock)), dim3(kBlock), 0, c->stream,
                       g, (const DevState*)c->d_st, lv, c->d_f0, c->d_f1);
    LAUNCH_CK("refine flags");
    const int32_t* dn = &c->d_st->lv[lv].cnum;
    CK(scan_i32(c, c->d_f0, c->d_o0, cbound, dn, &c->d_st->nv, 0));
    CK(scan_i32(c, c->d_f1, c->d_o1, cbound, dn, &c->d_st->U, 0));
    CK(sync_state(c));
    c->lv[lv] = c->h_st->lv[lv];
    const int U = c->h_st->U;
    const int64_t nv = c->h_st->nv;
    if (autod) {
      // auto-depth rule, :1597-1604
      float k = U * 1.0 / N;
      c->gamma = k;
      if (k > 0.1 || (k < 0.1 && (0.1 - k) < 0.01)) {
        early = true;
        c->gnum = lv + 1;
        break;
      }
    }
    const LevelDesc L = c->lv[lv];
    const int64_t vbase = (int64_t)L.vstart + L.vnum;
    const int64_t bbase = (int64_t)L.bstart + L.bnum;
    const int64_t cbase = (int64_t)L.cstart + L.cnum;
    if (vbase + nv > INT32_MAX || bbase + 7 * nv > INT32_MAX || cbase + U + nv > INT32_MAX)
      return set_err(c, GAHHG_E_LIMIT, "build: level %d exceeds int32 indexing", lv + 1);
    CK(ensure_v(c, vbase, vbase + nv));
    CK(ensure_b(c, bbase, bbase + 7 * nv));
    CK(ensure_c(c, cbase, cbase + U + nv));
    g = view(c);
    hipLaunchKernelGGL(k_create_center, dim3(blocks_for(L.cnum, kBlock)), dim3(kBlock), 0, c->stream,
                       g, (const DevState*)c->d_st, lv, (const int32_t*)c->d_o0, (const int32_t*)c->d_o1);
    LAUNCH_CK("create center");
    hipLaunchKernelGGL(k_bsc_flags, dim3(blocks_for(nv, kBlock)), dim3(kBlock), 0, c->stream, g,
                       (int)bbase, (int)nv, c->d_f0, c->d_f1);
    LAUNCH_CK("border child flags");
    CK(scan_i32(c, c->d_f0, c->d_o0, nv, nullptr, &c->d_st->e0, 0));
    CK(scan_i32(c, c->d_f1, c->d_o1, nv, nullptr, &c->d_st->c0, 0));
    hipLaunchKernelGGL(k_bsc_create, dim3(blocks_for(nv, kBlock)), dim3(kBlock), 0, c->stream, g,
                       (int)bbase, (int)nv, (int)cbase, U, (const int32_t*)c->d_o0,
                       (const int32_t*)c->d_o1);
    LAUNCH_CK("border child create");
    hipLaunchKernelGGL(k_sc_partner, dim3(blocks_for(nv, kBlock)), dim3(kBlock), 0, c->stream, g,
                       (int)(cbase + U), (const DevState*)c->d_st);
    LAUNCH_CK("border child partners");
    hipLaunchKernelGGL(k_level_next, dim3(1), dim3(1), 0, c->stream, c->d_st, lv, (int)vbase, (int)nv,
                       (int)bbase, (int)cbase, U);
    LAUNCH_CK("level table");
    // host-side upper bound for the next level until the next readback
    c->lv[lv + 1] = LevelDesc{(int)vbase, (int)nv, (int)bbase, (int)(7 * nv), (int)cbase, (int)(U + nv), 0};
  }
  if (maxd == 1) c->gamma = (float)(42 * 1.0 / N);

  // dispatchSELast at level L = gnum - 1
  const int Lb = c->gnum - 1;
  int32_t* par = (early && c->gnum != 1) ? c->d_loc_prev : c->d_loc_cur;
  CK(sync_state(c));
  for (int l = 0; l < c->gnum; l++) c->lv[l] = c->h_st->lv[l];
  const LevelDesc LB = c->lv[Lb];
  if (c->ecap < N) {
    dfree(c->d_ecnt);
    c->d_ecnt = nullptr;
    CK(dalloc(c, (void**)&c->d_ecnt, (int64_t)N * 4));
    c->ecap = N;
  }
  CK(ensure_flags(c, std::max<int64_t>(N, LB.cnum) + 1));
  g = view(c);
  hipLaunchKernelGGL(k_walk<kCount>, dim3(blocks_for(N, kBlock)), dim3(kBlock), 0, c->stream, g,
                     (const double*)c->d_edge, N, Lb, kGrid0Cells, par, (int32_t*)nullptr,
                     (const int32_t*)nullptr, c->d_ecnt, (int32_t*)nullptr, c->d_st);
  LAUNCH_CK("count walk");
  CK(scan_i32(c, c->d_ecnt, c->d_o0, N, nullptr, &c->d_st->R, 0));
  CK(sync_state(c));
  const int64_t R = c->h_st->R;
  if (R < 0) return set_err(c, GAHHG_E_LIMIT, "dispatch: ref count overflow");
  CK(ensure_refs(c, R + 1));
  g = view(c);
  hipLaunchKernelGGL(k_walk<kWrite>, dim3(blocks_for(N, kBlock)), dim3(kBlock), 0, c->stream, g,
                     (const double*)c->d_edge, N, Lb, kGrid0Cells, par, (int32_t*)nullptr,
                     (const int32_t*)c->d_o0, (int32_t*)nullptr, c->d_rcell, c->d_st);
  LAUNCH_CK("write walk");
  // ref -> edge id: inclusive scan of head flags
  CK(fill_i32(c, c->d_rflag, R, 0));
  hipLaunchKernelGGL(k_edge_heads, dim3(blocks_for(N, kBlock)), dim3(kBlock), 0, c->stream,
                     (const int32_t*)c->d_o0, N, (const DevState*)c->d_st, c->d_rflag);
  LAUNCH_CK("edge heads");
  CK(scan_i32(c, c->d_rflag, c->d_redge, R, nullptr, nullptr, 1));
  // counting sort into cell-major order
  CK(fill_i32(c, c->d_f1, LB.cnum, 0));
  hipLaunchKernelGGL(k_count_cells, dim3(blocks_for(R, kBlock)), dim3(kBlock), 0, c->stream,
                     (const int32_t*)c->d_rcell, (const DevState*)c->d_st, LB.cstart, c->d_f1);
  LAUNCH_CK("count cells");
  CK(scan_i32(c, c->d_f1, c->d_o1, LB.cnum, nullptr, nullptr, 0));
  CK(copy_i32(c, c->d_f0, c->d_o1, LB.cnum));
  hipLaunchKernelGGL(k_scatter_refs, dim3(blocks_for(R, kBlock)), dim3(kBlock), 0, c->stream,
                     (const int32_t*)c->d_rcell, (const int32_t*)c->d_redge, (const DevState*)c->d_st,
                     LB.cstart, c->d_f0, c->d_rflag);
  LAUNCH_CK("scatter refs");
  hipLaunchKernelGGL(k_finish_cells, dim3(blocks_for(LB.cnum, 4)), dim3(256), 0, c->stream, g,
                     LB.cstart, LB.cnum, (const int32_t*)c->d_f1, (const int32_t*)c->d_o1,
                     (const int32_t*)c->d_rflag);
  LAUNCH_CK("finish cells");
  CK(hip_check(c, hipEventRecord(c->ev[1], c->stream), "event"));
  CK(sync_state(c));
  c->lv[Lb].rnum = (int32_t)R;
  return GAHHG_OK;
}

}  // namespace gahhg
