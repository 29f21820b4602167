// gahhg_classify.hip -- leaf labelling on gfx950.
//
// GPISP::classifyAHHG / GAHHG::classifyBottomGrid
// (source/GPISPAHHG.cu:1787-1817,2018-2059) with kernels
// setBottomNECell (:1046-1071), setBottomDECell (:1074-1124),
// setDECell (:1292-1363), setIECell (:1366-1392), setTopECell (:1395-1419).
//
// The reference's propagation kernels read neighbour types while other
// threads write them, so which cells end up labelled depends on scheduling.
// Here every propagation pass reads a snapshot taken before the pass
// (double-buffered types), and the adoption passes repeat until nothing
// changes.  The values written are the reference's; only coverage is made
// deterministic (and complete where the reference's single pass stops).
#include <float.h>
#include <stdio.h>

#include <algorithm>

#include "gahhg_internal.h"

#pragma clang fp contract(off)

namespace gahhg {

constexpr int kCBlock = 256;

static inline unsigned cblocks(int64_t n) {
  return (unsigned)std::max<int64_t>((n + kCBlock - 1) / kCBlock, 1);
}

// blocks for G lanes per item
static inline unsigned gblocks(int64_t n, int G) {
  return (unsigned)std::max<int64_t>((n * G + kCBlock - 1) / kCBlock, 1);
}

// Lanes per cell for the per-ref loops of a level: the smallest power of two
// >= a quarter of the mean refs per bottom cell, in [1, 64] (small groups
// keep many cells per wave where cells hold few refs).
__host__ __device__ int group_for_refs(double mean_refs) {
  int G = 1;
  while (G < 64 && G < 0.25 * mean_refs) G *= 2;
  return G;
}

__device__ __forceinline__ V3 cvtx(const GridView& g, int i) { return load3(g.v + 3 * (int64_t)i); }

// Leaf centre as the query and classification kernels compute it: the mean
// of the border start vertices, normalised (:1062-1066, :1481-1485).
__device__ __forceinline__ V3 mean_center(const GridView& g, int cid) {
  const int bs = g.c_bstart[cid], bn = g.c_bnum[cid];
  V3 c = v3(0.0, 0.0, 0.0);
  for (int i = 0; i < bn; i++) c = add(c, cvtx(g, g.b_v0[bs + i]));
  c = divk(c, bn);
  return unit(c);
}

// Cells of one level whose polygon bottoms out there (bottom) or deeper
// (upper); every polygon of the batch follows its own reference sequence.
__device__ __forceinline__ int cell_phase(const GridView& g, const PolyDev* pd, int cid, int lv) {
  const int L = pd[g.c_poly[cid]].L;
  return L == lv ? 0 : (L > lv ? 1 : 2);
}

__device__ __forceinline__ double shfl_d(double v, int src) {
  return __hiloint2double(__shfl(__double2hiint(v), src, 64), __shfl(__double2loint(v), src, 64));
}

struct Group {
  int lane, gl, gbase, G;
  unsigned long long mask;
};

__device__ __forceinline__ Group group_of(int G) {
  Group r;
  r.G = G;
  r.lane = threadIdx.x & 63;
  r.gl = r.lane & (G - 1);
  r.gbase = r.lane & ~(G - 1);
  r.mask = G == 64 ? ~0ull : (((1ull << G) - 1ull) << r.gbase);
  return r;
}

// The feature device_getCPProp (:61-171) picks among one cell's refs.
struct CpChoice {
  int slot;        // position in the refs of the chosen edge; -1: none chosen
  int nv;          // its nearest polygon vertex, -1 when the nearest point is inside the edge
  double mindist;  // squared chord distance from the centre to that point
  double side;     // dot(edge normal, centre) of the chosen edge
};

// device_getCPProp's loop over refs[0, rn) with a group of G lanes (G a
// power of two up to 64): the lanes evaluate the nearest feature of the
// refs in parallel (the per-ref part of the loop), then every lane of the
// group replays the reference's selection over those values in ref order,
// so the chosen feature -- including the shared-vertex tie rule, whose
// comparison needs only each ref's own values -- is the reference's.
// Conditions are uniform within a group; every lane returns the choice.
__device__ __forceinline__ CpChoice cp_select(const int32_t* __restrict__ refs, int rn, const int32_t* __restrict__ pe,
                                              const double* __restrict__ edge, V3 center, const Group& gr) {
  const int G = gr.G;
  double mindist = FLT_MAX;
  int best = -1;            // ref slot of the chosen feature
  double best_side = 0.0;   // dot(n0, center) of the chosen edge
  double best_lean = 0.0;
  int nv_index = -1;
  for (int base = 0; base < rn; base += G) {
    const int i = base + gr.gl;
    double curdist = 0.0, side = 0.0, lean = 0.0;
    int cur_nv = -1;
    if (i < rn) {
      const int se = refs[i];
      const double* er = edge + 9 * (int64_t)se;
      const V3 v0 = load3(er), v1 = load3(er + 3), n0 = load3(er + 6);
      V3 n1 = unit(cross(center, n0));
      V3 inter = unit(cross(n0, n1));
      if (lt0(dot(inter, v0))) inter = neg(inter);
      V3 t0 = unit(sub(v0, inter));
      V3 t1 = unit(sub(v1, inter));
      const double t3 = dot(t0, t1);
      if (lt0(t3)) {
        curdist = sqlen(sub(inter, center));
        cur_nv = -1;
      } else {
        curdist = sqlen(sub(center, v0));
        const double vd = sqlen(sub(center, v1));
        if (vd < curdist) {
          curdist = vd;
          cur_nv = pe[2 * se + 1];
        } else {
          cur_nv = pe[2 * se];
        }
      }
      side = dot(n0, center);
      if (cur_nv != -1) {
        // the tie rule's lean of this edge: its far end towards the centre,
        // seen from the shared vertex.  The vertices are v0 = pv[pe[2 se]]
        // and v1 = pv[pe[2 se + 1]] (the edge record holds copies), so
        // pv[cur_nv] and pv[other end] are taken from there.
        const V3 pnv = cur_nv == pe[2 * se] ? v0 : v1;
        const V3 oth = (pe[2 * se] == cur_nv) ? v1 : v0;
        const V3 a1 = unit(sub(oth, pnv));
        lean = dot(a1, sub(center, pnv));
      }
    }
    // Only the refs at the chunk's least distance can change the choice:
    // one at a larger distance either loses to that ref later in the chunk
    // or (at or past the current minimum) changes nothing, and the
    // shared-vertex rule links refs of one nearest vertex, which are at the
    // same distance.  So the selection is replayed over those refs alone,
    // and a chunk whose least distance exceeds the current one is skipped.
    double cmin = i < rn ? curdist : __builtin_inf();
    for (int o = 1; o < G; o <<= 1)
      cmin = fmin(cmin, __hiloint2double(__shfl_xor(__double2hiint(cmin), o, G), __shfl_xor(__double2loint(cmin), o, G)));
    if (!(cmin <= mindist)) continue;  // uniform within the group
    unsigned long long cand = __ballot(i < rn && curdist == cmin) & gr.mask;
    while (cand) {
      const int src = __ffsll((long long)cand) - 1;
      cand &= cand - 1;
      const int j = src - gr.gbase;
      const double cd = shfl_d(curdist, src), sd = shfl_d(side, src), ln = shfl_d(lean, src);
      const int cn = __shfl(cur_nv, src, 64);
      if (nv_index != -1 && nv_index == cn) {
        const double a = best_side, b = sd;
        if ((le0(a) && ge0(b)) || (ge0(a) && le0(b))) {
          if (ln > best_lean) {
            mindist = cd;
            best = base + j;
            best_side = sd;
            best_lean = ln;
            nv_index = cn;
            continue;
          }
        }
      }
      if (cd < mindist) {
        mindist = cd;
        best = base + j;
        best_side = sd;
        best_lean = ln;
        nv_index = cn;
      }
    }
  }
  return CpChoice{best, nv_index, mindist, best_side};
}

// The type getCPProp gives: inside when the centre is on the inner side of
// the chosen edge.  With no feature chosen (all distances NaN) the reference
// keeps a zero normal, whose side test fails.
__device__ __forceinline__ int cp_type(const CpChoice& ch) { return gt0(ch.slot >= 0 ? ch.side : 0.0) ? kIn : kOut; }

// setBottomNECell (:1046-1071): a group of G lanes per bottom boundary leaf
// (G chosen from the level's mean refs per leaf), typed by getCPProp at the
// leaf's centre.
__global__ void __launch_bounds__(kCBlock) k_ne(GridView g, const PolyDev* __restrict__ pd, const double* pv,
                                                const int32_t* pe, const double* edge, int cstart, int cnum, int lv,
                                                int G) {
  const Group gr = group_of(G);
  const int w = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G);
  if (w >= cnum) return;
  const int cid = cstart + w;
  if (cell_phase(g, pd, cid, lv) != 0) return;
  const int rn = g.c_rnum[cid];
  if (rn == 0) return;
  const V3 center = mean_center(g, cid);
  const CpChoice ch = cp_select(g.refs + g.c_rstart[cid], rn, pe, edge, center, gr);
  if (gr.gl == 0) g.c_type[cid] = cp_type(ch);
}

// Known-answer hook: getCPProp's choice at given centres over given ref
// lists (case k: refs[rs_rn[2k], + rs_rn[2k+1]), centre centers[3k..]).
__global__ void __launch_bounds__(kCBlock) k_kat_cp_prop(const double* pv, const int32_t* pe, const double* edge,
                                                         const int32_t* refs, const int32_t* rs_rn,
                                                         const double* centers, int n, int G, int32_t* out3,
                                                         double* mindist) {
  const Group gr = group_of(G);
  const int w = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G);
  if (w >= n) return;
  const V3 center = load3(centers + 3 * (int64_t)w);
  const CpChoice ch = cp_select(refs + rs_rn[2 * w], rs_rn[2 * w + 1], pe, edge, center, gr);
  if (gr.gl == 0) {
    out3[3 * w] = cp_type(ch);
    out3[3 * w + 1] = ch.slot;
    out3[3 * w + 2] = ch.nv;
    mindist[w] = ch.mindist;
  }
}

__global__ void k_snap(const int32_t* type, int cstart, int cnum, int32_t* snap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < cnum) snap[t] = type[cstart + t];
}

// Crossing parity from the centre of `from` to the midpoint of border `mb`
// over the refs of `from` (dev_isInterSeg, :174-205), giving the label on
// the far side; a lane group splits the refs and the count is the popcount
// of the group's ballots.
__device__ int parity_label_group(const GridView& g, const double* __restrict__ edge, int from, int mb,
                                  int from_type, const Group& gr) {
  const V3 m = unit(mul(add(cvtx(g, g.b_v0[mb]), cvtx(g, g.b_v1[mb])), 0.5));
  const V3 c = mean_center(g, from);
  const V3 n1 = arc_normal(c, m);
  int cnt = 0;
  const int rs = g.c_rstart[from], rn = g.c_rnum[from];
  for (int base = 0; base < rn; base += gr.G) {
    const int i = base + gr.gl;
    bool x = false;
    if (i < rn) {
      const double* er = edge + 9 * (int64_t)g.refs[rs + i];
      const V3 u0 = load3(er), u1 = load3(er + 3), n0 = load3(er + 6);
      x = !same_side(dot(n0, c), dot(n0, m)) && !same_side(dot(n1, u0), dot(n1, u1));
    }
    cnt += __popcll(__ballot(x) & gr.mask);
  }
  return (cnt % 2 == 0) ? from_type : 3 - from_type;
}

// setBottomDECell (a lane group per cell): every border whose neighbour was
// labelled in the snapshot decides in turn; the last one wins, as in the
// reference loop.
__device__ __forceinline__ void de_bottom_cell(const GridView& g, const PolyDev* __restrict__ pd, const double* edge,
                                               int cstart, int t, int lv, const int32_t* snap, const Group& gr) {
  if (snap[t] != kUndefined) return;
  const int cid = cstart + t;
  if (cell_phase(g, pd, cid, lv) != 0) return;
  // the reference loop lets every labelled neighbour decide in border
  // order, the last one's label standing: the last labelled neighbour
  // alone gives it (found from the end; one parity count, not one per
  // labelled neighbour)
  const int bs = g.c_bstart[cid], bn = g.c_bnum[cid];
  for (int i = bn - 1; i >= 0; i--) {
    const int pe_ = g.b_partner[bs + i];
    if (pe_ == -1) continue;
    const int nb = g.b_parent[pe_];
    const int nt = snap[nb - cstart];
    if (nt == kUndefined) continue;
    const int lab = parity_label_group(g, edge, nb, pe_, nt, gr);
    if (gr.gl == 0) g.c_type[cid] = lab;
    return;
  }
}

__global__ void __launch_bounds__(kCBlock) k_de_bottom(GridView g, const PolyDev* __restrict__ pd,
                                                       const double* edge, int cstart, int cnum, int lv,
                                                       const int32_t* snap, int G) {
  const Group gr = group_of(G);
  const int t = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G);
  if (t < cnum) de_bottom_cell(g, pd, edge, cstart, t, lv, snap, gr);
}

// setIECell round: an unlabelled leaf adopts the label of the first
// neighbour that is an empty labelled cell.
__device__ __forceinline__ void ie_cell(const GridView& g, int cstart, int t, const int32_t* snap, DevState* st) {
  const int cid = cstart + t;
  if (snap[t] != kUndefined || g.c_subcell[cid] != -1) return;
  const int bs = g.c_bstart[cid], bn = g.c_bnum[cid];
  for (int i = 0; i < bn; i++) {
    const int pe_ = g.b_partner[bs + i];
    if (pe_ == -1) continue;
    const int pc = g.b_parent[pe_];
    const int pt = snap[pc - cstart];
    if (g.c_rnum[pc] == 0 && pt != kUndefined) {
      g.c_type[cid] = pt;
      st->changed = 1;
      return;
    }
  }
}

__global__ void __launch_bounds__(kCBlock) k_ie(GridView g, int cstart, int cnum, const int32_t* snap,
                                                DevState* st) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < cnum) ie_cell(g, cstart, t, snap, st);
}

// setDECell: an empty unlabelled cell above the bottom takes its label from
// the leaf reached through the first refined neighbour.  The lookup is one
// lane per cell; the crossing parity over the leaf's refs (for leaves with
// refs) is counted by the whole wave, one pending cell after another, its
// 64 lanes splitting the refs (most cells of a level need none, and a leaf
// holds tens of refs that one lane would fetch in turn).
__device__ __forceinline__ void de_upper_wave(const GridView& g, const PolyDev* __restrict__ pd, const double* edge,
                                              int cstart, int64_t t, int64_t cnum, int lv) {
  int cid = -1, target = -1, mb = -1, ttype = 0;
  bool need = false;
  if (t < cnum) {
    cid = cstart + (int)t;
    if (cell_phase(g, pd, cid, lv) == 1 && g.c_rnum[cid] == 0 && g.c_type[cid] == kUndefined) {
      const int bs = g.c_bstart[cid], bn = g.c_bnum[cid];
      int pe_ = -1, pc = -1;
      for (int i = 0; i < bn; i++) {
        pe_ = g.b_partner[bs + i];
        if (pe_ != -1 && g.c_subcell[g.b_parent[pe_]] != -1) {
          pc = g.b_parent[pe_];
          break;
        }
      }
      if (pc != -1 && (target = g.b_subcell[pe_]) >= 0) {
        while (g.c_subcell[target] != -1) target = g.c_subcell[target];
        ttype = g.c_type[target];
        if (g.c_rnum[target] == 0) {
          g.c_type[cid] = ttype;
        } else {
          need = true;
          mb = pe_;
        }
      }
    }
  }
  const Group gr = group_of(64);
  unsigned long long pend = __ballot(need);
  while (pend) {
    const int src = __ffsll((long long)pend) - 1;
    pend &= pend - 1;
    const int lab = parity_label_group(g, edge, __shfl(target, src, 64), __shfl(mb, src, 64),
                                       __shfl(ttype, src, 64), gr);
    if (gr.lane == src) g.c_type[cid] = lab;
  }
}

// Every lane of a wave runs de_upper_wave the same number of times (the
// loop bound is the wave's first cell).
__global__ void __launch_bounds__(kCBlock) k_de_upper(GridView g, const PolyDev* __restrict__ pd,
                                                      const double* edge, int cstart, int cnum, int lv) {
  const int64_t w0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63);
  de_upper_wave(g, pd, edge, cstart, w0 + (threadIdx.x & 63), cnum, lv);
}

// setTopECell on one polygon's 42 level-0 cells (block = polygon),
// repeated to a fixpoint.
__global__ void k_top(GridView g, const PolyDev* __restrict__ pd, const DevState* __restrict__ gate) {
  __shared__ int snap[64];
  __shared__ int changed;
  if (gate && gate->abort) return;
  const int base = pd[blockIdx.x].cell0;
  const int cnum = kGrid0Cells;
  const int t = threadIdx.x;
  for (;;) {
    if (t < cnum) snap[t] = g.c_type[base + t];
    if (t == 0) changed = 0;
    __syncthreads();
    if (t < cnum && g.c_rnum[base + t] == 0 && snap[t] == kUndefined) {
      const int bs = g.c_bstart[base + t], bn = g.c_bnum[base + t];
      for (int i = 0; i < bn; i++) {
        const int pe_ = g.b_partner[bs + i];
        if (pe_ != -1 && snap[g.b_parent[pe_] - base] != kUndefined) {
          g.c_type[base + t] = snap[g.b_parent[pe_] - base];
          changed = 1;
          break;
        }
      }
    }
    __syncthreads();
    const int ch = changed;
    __syncthreads();
    if (!ch) break;
  }
}

// Counts leaves of one level left unlabelled (diagnostic).
__global__ void k_count_undef(GridView g, int cstart, int cnum, DevState* st) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnum) return;
  const int cid = cstart + t;
  if (g.c_subcell[cid] == -1 && g.c_type[cid] == kUndefined) atomicAdd(&st->undefined_leaves, 1ULL);
}

// ---- device-driven classification: sizes, lane groups and the fixpoint on
// the device (every kernel returns at once when the build stopped short or
// the level is below every polygon's bottom) ----
__device__ __forceinline__ bool cls_skip(const DevState* st, int lv) { return st->abort || lv > st->maxL; }

#define FOR_LEVEL(L, t) \
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < (L).cnum; t += gridDim.x * blockDim.x)

// Lane-group size of level lv (as the host version chooses it) and the
// fixpoint list cleared.
__global__ void __launch_bounds__(256) k_cls_level(const PolyDev* __restrict__ pd, int np, DevState* st, int lv) {
  if (cls_skip(st, lv)) return;
  __shared__ long long sr[256], sl[256];
  long long r = 0, l = 0;
  for (int p = threadIdx.x; p < np; p += 256)
    if (pd[p].L == lv) {
      r += pd[p].R;
      l += pd[p].bcn;
    }
  sr[threadIdx.x] = r;
  sl[threadIdx.x] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long refs = 0, leaves = 0;
    for (int i = 0; i < 256; i++) refs += sr[i], leaves += sl[i];
    st->cls_G[lv] = group_for_refs(leaves > 0 ? (double)refs / leaves : 1.0);
    st->ie_n = 0;
  }
}

__global__ void __launch_bounds__(kCBlock, 6) k_ne_dv(GridView g, const PolyDev* __restrict__ pd, const double* pv,
                                                   const int32_t* pe, const double* edge, const DevState* st, int lv) {
  if (cls_skip(st, lv)) return;
  const int G = st->cls_G[lv];
  const Group gr = group_of(G);
  const LevelDesc L = st->lv[lv];
  const int groups = (int)((int64_t)gridDim.x * blockDim.x / G);
  for (int w = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G); w < L.cnum; w += groups) {
    const int cid = L.cstart + w;
    if (cell_phase(g, pd, cid, lv) != 0) continue;
    const int rn = g.c_rnum[cid];
    if (rn == 0) continue;
    const V3 center = mean_center(g, cid);
    const CpChoice ch = cp_select(g.refs + g.c_rstart[cid], rn, pe, edge, center, gr);
    if (gr.gl == 0) g.c_type[cid] = cp_type(ch);
  }
}

__global__ void k_snap_dv(const int32_t* type, const DevState* st, int lv, int32_t* snap) {
  if (cls_skip(st, lv)) return;
  const LevelDesc L = st->lv[lv];
  FOR_LEVEL(L, t) snap[t] = type[L.cstart + t];
}

__global__ void __launch_bounds__(kCBlock) k_de_bottom_dv(GridView g, const PolyDev* __restrict__ pd,
                                                          const double* edge, const DevState* st, int lv,
                                                          const int32_t* snap) {
  if (cls_skip(st, lv)) return;
  const int G = st->cls_G[lv];
  const Group gr = group_of(G);
  const LevelDesc L = st->lv[lv];
  const int groups = (int)((int64_t)gridDim.x * blockDim.x / G);
  for (int t = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G); t < L.cnum; t += groups)
    de_bottom_cell(g, pd, edge, L.cstart, t, lv, snap, gr);
}

__global__ void __launch_bounds__(kCBlock) k_de_upper_dv(GridView g, const PolyDev* __restrict__ pd,
                                                         const double* edge, const DevState* st, int lv) {
  if (cls_skip(st, lv)) return;
  const LevelDesc L = st->lv[lv];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t w0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); w0 < L.cnum; w0 += stride)
    de_upper_wave(g, pd, edge, L.cstart, w0 + (threadIdx.x & 63), L.cnum, lv);
}

// setIECell's first round over the level.
__global__ void __launch_bounds__(kCBlock) k_ie_dv(GridView g, DevState* st, int lv, const int32_t* snap) {
  if (cls_skip(st, lv)) return;
  const LevelDesc L = st->lv[lv];
  FOR_LEVEL(L, t) ie_cell(g, L.cstart, t, snap, st);
}

// The level's leaves still unlabelled after that round: counted (as the
// host version counts them) and listed for the fixpoint.
__global__ void k_ie_list(GridView g, DevState* st, int lv, int32_t* list) {
  if (cls_skip(st, lv)) return;
  const LevelDesc L = st->lv[lv];
  FOR_LEVEL(L, t) {
    const int cid = L.cstart + t;
    if (g.c_subcell[cid] == -1 && g.c_type[cid] == kUndefined) {
      list[atomicAdd(&st->ie_n, 1)] = cid;
      atomicAdd(&st->und_first, 1ULL);
    }
  }
}

// setIECell's remaining rounds to the fixpoint, by one workgroup over the
// listed leaves: each round reads every listed leaf's neighbours (phase 1:
// nothing is written, so the reads see the types at the round's start, the
// snapshot of the host version's rounds) and then writes the new types
// (phase 2); it ends after a round that changed nothing.
constexpr int kFixBlock = 1024;
__global__ void __launch_bounds__(kFixBlock) k_ie_fix(GridView g, const DevState* st, int lv, const int32_t* list,
                                                      int32_t* nt) {
  if (cls_skip(st, lv)) return;
  const int n = st->ie_n;
  __shared__ int changed;
  for (;;) {
    if (threadIdx.x == 0) changed = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kFixBlock) {
      const int cid = list[i];
      int ty = g.c_type[cid];
      if (ty == kUndefined) {
        const int bs = g.c_bstart[cid], bn = g.c_bnum[cid];
        for (int k = 0; k < bn; k++) {
          const int pe_ = g.b_partner[bs + k];
          if (pe_ == -1) continue;
          const int pc = g.b_parent[pe_];
          const int pt = g.c_type[pc];
          if (g.c_rnum[pc] == 0 && pt != kUndefined) {
            ty = pt;
            changed = 1;
            break;
          }
        }
      }
      nt[i] = ty;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kFixBlock) g.c_type[list[i]] = nt[i];
    __syncthreads();
    const int ch = changed;
    __syncthreads();
    if (!ch) break;
  }
}

__global__ void k_count_undef_dv(GridView g, DevState* st) {
  if (st->abort) return;
  const int C = st->tot_c32;
  for (int cid = blockIdx.x * blockDim.x + threadIdx.x; cid < C; cid += gridDim.x * blockDim.x)
    if (g.c_subcell[cid] == -1 && g.c_type[cid] == kUndefined) atomicAdd(&st->undefined_leaves, 1ULL);
}

#define CCK(x)                       \
  do {                               \
    int _rc = (x);                   \
    if (_rc != GAHHG_OK) return _rc; \
  } while (0)

__global__ void k_clear_changed(DevState* st) { st->changed = 0; }
__global__ void k_clear_undef(DevState* st) { st->undefined_leaves = 0; }

// setIECell to a fixpoint.  Rounds run in batches; the change flag is
// cleared before the last round of a batch, so a clear flag after the batch
// means the last round changed nothing.  Returns the number of leaves still
// unlabelled after the first round via *after_first.
static int ie_fixpoint(gahhg_ctx* c, int lv, int32_t* snap, unsigned long long* after_first) {
  const LevelDesc L = c->glv[lv];
  GridView g = view(c);
  bool first = true;
  for (int batch = 0; batch < 1 << 20; batch++) {
    const int rounds = first ? 1 : 4;
    for (int r = 0; r < rounds; r++) {
      if (r == rounds - 1) hipLaunchKernelGGL(k_clear_changed, dim3(1), dim3(1), 0, c->stream, c->d_st);
      hipLaunchKernelGGL(k_snap, dim3(cblocks(L.cnum)), dim3(kCBlock), 0, c->stream,
                         (const int32_t*)c->g.c_type, L.cstart, L.cnum, snap);
      hipLaunchKernelGGL(k_ie, dim3(cblocks(L.cnum)), dim3(kCBlock), 0, c->stream, g, L.cstart, L.cnum,
                         (const int32_t*)snap, c->d_st);
    }
    if (first) {
      hipLaunchKernelGGL(k_clear_undef, dim3(1), dim3(1), 0, c->stream, c->d_st);
      hipLaunchKernelGGL(k_count_undef, dim3(cblocks(L.cnum)), dim3(kCBlock), 0, c->stream, g, L.cstart,
                         L.cnum, c->d_st);
    }
    CCK(hip_check(c, hipGetLastError(), "classify fixpoint"));
    CCK(hip_check(c, hipMemcpyAsync(c->h_st, c->d_st, sizeof(DevState), hipMemcpyDeviceToHost, c->stream),
                  "readback"));
    CCK(hip_check(c, hipStreamSynchronize(c->stream), "sync"));
    if (first) *after_first = c->h_st->undefined_leaves;
    first = false;
    if (!c->h_st->changed) return GAHHG_OK;
  }
  return set_err(c, GAHHG_E_STATE, "classification did not converge");
}

int run_kat_cp_prop(gahhg_ctx* c, const int32_t* d_refs, const int32_t* d_rs_rn, const double* d_centers, int64_t n,
                    int G, int32_t* d_out3, double* d_mindist) {
  if (n <= 0) return GAHHG_OK;
  hipLaunchKernelGGL(k_kat_cp_prop, dim3(gblocks((int)n, G)), dim3(kCBlock), 0, c->stream, (const double*)c->d_pv,
                     (const int32_t*)c->d_pe, (const double*)c->d_edge, d_refs, d_rs_rn, d_centers, (int)n, G, d_out3,
                     d_mindist);
  return hip_check(c, hipGetLastError(), "kat cp_prop kernel");
}

// classifyAHHG (:2018-2059) with every size and lane-group choice on the
// device: levels [0, levels) from the deepest up (the ones below every
// polygon's bottom return at once), then setTopECell and the count of
// leaves left unlabelled.  Enqueued only; nothing waits.
int classify_dev_enqueue(gahhg_ctx* c, int levels) {
  const int np = (int)c->polys.size();
  const GridView g = view(c);
  const PolyDev* pd = c->d_pd;
  DevState* st = c->d_st;
  int32_t* snap = c->d_f0;
  int32_t* list = c->d_f2;
  int32_t* nt = c->d_f3;
  const unsigned cg = (unsigned)std::min<int64_t>(std::max<int64_t>((c->g.ccap + kCBlock - 1) / kCBlock, 1), 2048);
  const unsigned gg = (unsigned)std::min<int64_t>(gblocks(c->g.ccap, 64), 2048);  // lane-group launches
  for (int lv = levels - 1; lv >= 0; lv--) {
    hipLaunchKernelGGL(k_cls_level, dim3(1), dim3(256), 0, c->stream, pd, np, st, lv);
    hipLaunchKernelGGL(k_ne_dv, dim3(gg), dim3(kCBlock), 0, c->stream, g, pd, (const double*)c->d_pv,
                       (const int32_t*)c->d_pe, (const double*)c->d_edge, (const DevState*)st, lv);
    hipLaunchKernelGGL(k_snap_dv, dim3(cg), dim3(kCBlock), 0, c->stream, (const int32_t*)c->g.c_type,
                       (const DevState*)st, lv, snap);
    hipLaunchKernelGGL(k_de_bottom_dv, dim3(gg), dim3(kCBlock), 0, c->stream, g, pd, (const double*)c->d_edge,
                       (const DevState*)st, lv, (const int32_t*)snap);
    hipLaunchKernelGGL(k_de_upper_dv, dim3(cg), dim3(kCBlock), 0, c->stream, g, pd, (const double*)c->d_edge,
                       (const DevState*)st, lv);
    hipLaunchKernelGGL(k_snap_dv, dim3(cg), dim3(kCBlock), 0, c->stream, (const int32_t*)c->g.c_type,
                       (const DevState*)st, lv, snap);
    hipLaunchKernelGGL(k_ie_dv, dim3(cg), dim3(kCBlock), 0, c->stream, g, st, lv, (const int32_t*)snap);
    hipLaunchKernelGGL(k_ie_list, dim3(cg), dim3(kCBlock), 0, c->stream, g, st, lv, list);
    hipLaunchKernelGGL(k_ie_fix, dim3(1), dim3(kFixBlock), 0, c->stream, g, (const DevState*)st, lv,
                       (const int32_t*)list, nt);
    CCK(hip_check(c, hipGetLastError(), "classify level"));
  }
  hipLaunchKernelGGL(k_top, dim3(np), dim3(64), 0, c->stream, g, pd, (const DevState*)st);
  hipLaunchKernelGGL(k_count_undef_dv, dim3(cg), dim3(kCBlock), 0, c->stream, g, st);
  return hip_check(c, hipGetLastError(), "classify top");
}

int classify_hierarchy(gahhg_ctx* c) {
  const int np = (int)c->polys.size();
  int maxL = 0;
  for (const PolyDev& P : c->pd) maxL = std::max(maxL, (int)P.L);
  GridView g = view(c);
  int32_t* snap = c->d_f0;
  for (int lv = 0; lv <= maxL; lv++)
    if (c->fcap < c->glv[lv].cnum) return set_err(c, GAHHG_E_STATE, "classify: scratch too small");
  unsigned long long und_snap = 0, und1 = 0;
  const PolyDev* pd = c->d_pd;
  for (int lv = maxL; lv >= 0; lv--) {
    if (c->prm.verbose) printf("Classify level %d cells.\n", lv);
    const LevelDesc L = c->glv[lv];
    // lane group per cell from the mean refs of the leaves with refs of the
    // polygons whose bottom level this is
    int64_t refs = 0, leaves = 0;
    for (const PolyDev& P : c->pd)
      if (P.L == lv) {
        refs += P.R;
        leaves += P.bcn;
      }
    const int G = group_for_refs(leaves > 0 ? (double)refs / leaves : 1.0);
    // bottom polygons: setBottomNECell, then setBottomDECell on a snapshot
    hipLaunchKernelGGL(k_ne, dim3(gblocks(L.cnum, G)), dim3(kCBlock), 0, c->stream, g, pd, (const double*)c->d_pv,
                       (const int32_t*)c->d_pe, (const double*)c->d_edge, L.cstart, L.cnum, lv, G);
    hipLaunchKernelGGL(k_snap, dim3(cblocks(L.cnum)), dim3(kCBlock), 0, c->stream, (const int32_t*)c->g.c_type,
                       L.cstart, L.cnum, snap);
    hipLaunchKernelGGL(k_de_bottom, dim3(gblocks(L.cnum, G)), dim3(kCBlock), 0, c->stream, g, pd,
                       (const double*)c->d_edge, L.cstart, L.cnum, lv, (const int32_t*)snap, G);
    // deeper polygons: setDECell
    hipLaunchKernelGGL(k_de_upper, dim3(cblocks(L.cnum)), dim3(kCBlock), 0, c->stream, g, pd,
                       (const double*)c->d_edge, L.cstart, L.cnum, lv);
    CCK(hip_check(c, hipGetLastError(), "classify level"));
    CCK(ie_fixpoint(c, lv, snap, &und1));
    und_snap += und1;
  }
  hipLaunchKernelGGL(k_top, dim3(np), dim3(64), 0, c->stream, g, pd, (const DevState*)nullptr);
  hipLaunchKernelGGL(k_clear_undef, dim3(1), dim3(1), 0, c->stream, c->d_st);
  hipLaunchKernelGGL(k_count_undef, dim3(cblocks(c->tot_c)), dim3(kCBlock), 0, c->stream, g, 0, (int)c->tot_c,
                     c->d_st);
  CCK(hip_check(c, hipGetLastError(), "classify top"));
  CCK(hip_check(c, hipMemcpyAsync(c->h_st, c->d_st, sizeof(DevState), hipMemcpyDeviceToHost, c->stream),
                "readback"));
  CCK(hip_check(c, hipStreamSynchronize(c->stream), "sync"));
  c->und_snap = (int64_t)und_snap;
  c->und_final = (int64_t)c->h_st->undefined_leaves;
  return GAHHG_OK;
}

// Loads this file's code object onto the current device (context set-up).
void load_code_classify() {
  hipFuncAttributes a;
  (void)hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_snap));
}

}  // namespace gahhg
