// gahhg_learn.hip -- learned refinement depth inside the build.
//
// The reference refines a polygon while the crossed cells of a level stay
// below about 0.09 N (divGrid, source/GPISPAHHG.cu:1595-1605).  With a model
// set, a polygon without a fixed depth is refined instead while
//
//   sum_r w_r (2^y_r(stop) - 2^y_r(refine)) * query_scale
//       >  ns_edge * N + ns_cell * (cells of the next level) + ns_level
//
// where, per top-level cell r of the polygon, w_r is the number of points
// expected there (from a hint sample, else the expected total spread by
// edge share) and y_r(k) is the model's log2 query time per point (ns) if
// refinement stops at this level (k = 0) or after one more (k = 1).  The
// features, the model and the decision run on the device after the level's
// plan; the decisions come back with the level's one host synchronisation.
// Hierarchies stay reference-structured (one bottom level per polygon), so
// the oracle built at the chosen depth reproduces them exactly.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "gahhg_internal.h"

#pragma clang fp contract(off)

namespace gahhg {

// A named namespace (not an anonymous one) so that profilers show the
// kernels' names.
namespace learn_kernels {

// The reference's inside test: every border plane n.p >= -eps.
__device__ __forceinline__ bool inside(const GridView& g, V3 q, int cid) {
  const int bs = g.c_bstart[cid], bn = g.c_bnum[cid];
  for (int k = 0; k < bn; k++)
    if (lt0(dot(load3(g.b_n + 3 * (int64_t)(bs + k)), q))) return false;
  return true;
}

// Points per (polygon, top-level cell) of the hint sample, added to the
// block's copy of the counters (kStripes copies of npoly x 42); the top-level
// cell is the first level-0 cell (index order) whose planes all pass, the
// same test the query's level-0 scan applies (level-0 geometry is the same
// for every polygon: polygon 0's cells are used, staged in LDS).
__global__ void k_hint_hist(GridView g, const PolyDev* __restrict__ pd, const double* __restrict__ xyz,
                            const int32_t* __restrict__ poly, int64_t n, int npoly, uint32_t* __restrict__ hist,
                            int32_t* __restrict__ loc) {
  __shared__ Level0Lds l0;
  level0_load(g, pd[0].cell0, &l0);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int key = -1;
  if (t < n) {
    int cell = -1;
    const int p = poly ? poly[t] : 0;
    if (p >= 0 && p < npoly) {
      const int i = level0_cell(&l0, load3(xyz + 3 * t));
      if (i >= 0) {
        key = p * kGrid0Cells + i;
        cell = pd[p].cell0 + i;
      }
    }
    loc[t] = cell;
  }
  wave_add<uint32_t>(hist + (int64_t)(blockIdx.x % kStripes) * npoly * kGrid0Cells, key, 1u);
}

// Hint points in the boundary band: per (polygon, top-level cell), the hint
// points whose first-match descent reaches a cell of the current level that
// is crossed (build time: marked; after the build: refined).  loc[t] holds
// the point's cell `steps` levels above (or -1 once it ended in a leaf above
// the band, where it stays); the kernel descends it and stores the cell.
// `steps` levels of the first-match descent from cell cur (-1 once the
// point ends in a leaf above, or no child holds it).
__device__ __forceinline__ int descend(const GridView& g, V3 q, int cur, int steps) {
  for (int s = 0; s < steps && cur >= 0; s++) {
    const int sc = g.c_subcell[cur];
    if (sc == -1) return -1;
    int next = inside(g, q, sc) ? sc : -1;
    if (next < 0) {
      const int bs = g.c_bstart[cur], bn = g.c_bnum[cur];
      for (int i = 0; i < bn && next < 0; i++) {
        const int tc = g.b_subcell[bs + i];
        if (tc != -1 && inside(g, q, tc)) next = tc;
      }
    }
    cur = next;
  }
  return cur;
}

__global__ void k_hint_band(GridView g, const double* __restrict__ xyz, int32_t* __restrict__ loc, int64_t n,
                            int steps, int marked, int64_t nbins, uint32_t* __restrict__ band) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int key = -1;
  if (t < n) {
    int cur = loc[t];
    if (cur >= 0) {
      const V3 q = load3(xyz + 3 * t);
      cur = descend(g, q, cur, steps);
      loc[t] = cur;
      if (cur >= 0 && (marked ? g.c_rnum[cur] != 0 : g.c_subcell[cur] != -1))
        key = g.c_poly[cur] * kGrid0Cells + g.c_root[cur];
    }
  }
  wave_add<uint32_t>(band + (blockIdx.x % kStripes) * nbins, key, 1u);
}

// out[b] = the sum of the kStripes copies of counter b.
__global__ void k_sum_stripes(const uint32_t* __restrict__ stripes, int64_t nbins, uint32_t* __restrict__ out) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  uint32_t s = 0;
  for (int k = 0; k < kStripes; k++) s += stripes[k * nbins + b];
  out[b] = s;
}

// Expected points per (polygon, top-level cell): the hint histogram times
// the weight per hint point, or the default spread by edge share.
__global__ void k_weights(int npoly, const PolyDev* __restrict__ pd, const uint32_t* __restrict__ hist,
                          double per_point, double q_default, const unsigned long long* __restrict__ estat,
                          float* __restrict__ w) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npoly * kGrid0Cells) return;
  const int p = t / kGrid0Cells;
  if (hist)
    w[t] = (float)(hist[t] * per_point);
  else
    w[t] = (float)(q_default * (double)estat[3 * t] / (double)max(pd[p].N, 1));
}

// Cells and crossed (build time: marked, rnum != 0) or refined (after the
// build: subcell != -1) cells of level range [lo, lo + n) per (polygon,
// top-level cell).  only_refining: count only polygons still refining.
__global__ void k_level_stats(GridView g, const PolyDev* __restrict__ pd, int lo, int n, int marked,
                              int only_refining, int32_t* __restrict__ st) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  int key = -1, crossed = 0;
  if (t < n) {
    const int cid = lo + t;
    const int p = g.c_poly[cid];
    if (!only_refining || pd[p].refining) {
      key = 2 * (p * kGrid0Cells + g.c_root[cid]);
      crossed = (marked ? g.c_rnum[cid] != 0 : g.c_subcell[cid] != -1) ? 1 : 0;
    }
  }
  wave_add<int32_t>(st, key, key >= 0 ? 1 : 0);
  wave_add<int32_t>(st, key >= 0 ? key + 1 : -1, crossed);
}

// The last refined level of polygon P (its bottom level - 1; 0 when it has
// one level): the level whose refined cells stand for the crossed ones.
__device__ __forceinline__ int last_refined(const PolyDev& P) { return max(P.L - 1, 0); }

// k_level_stats after a build, at each polygon's last refined level: every
// cell of the hierarchy is visited and counted when it lies on that level.
__global__ void k_level_stats_last(GridView g, const PolyDev* __restrict__ pd, LevelTab lt,
                                   int32_t* __restrict__ st) {
  const int cid = blockIdx.x * blockDim.x + threadIdx.x;
  int key = -1, crossed = 0;
  if (cid < lt.cstart[lt.n]) {
    int l = 0;
    while (l + 1 < lt.n && lt.cstart[l + 1] <= cid) l++;
    const int p = g.c_poly[cid];
    if (l == last_refined(pd[p])) {
      key = 2 * (p * kGrid0Cells + g.c_root[cid]);
      crossed = g.c_subcell[cid] != -1 ? 1 : 0;
    }
  }
  wave_add<int32_t>(st, key, key >= 0 ? 1 : 0);
  wave_add<int32_t>(st, key >= 0 ? key + 1 : -1, crossed);
}

// descend() on the query layout's node records (the same planes and child
// order, one record fetch per candidate cell).
__device__ __forceinline__ bool node_in(const int32_t* __restrict__ nodes, int cid, V3 q) {
  const int32_t* r = nodes + (int64_t)cid * kNodeInts;
  const int bn = r[1] & 0xff;
  const double* nr = reinterpret_cast<const double*>(r + 10);
  for (int k = 0; k < bn; k++)
    if (lt0(dot(v3(nr[3 * k], nr[3 * k + 1], nr[3 * k + 2]), q))) return false;
  return true;
}
__device__ __forceinline__ int descend_nodes_n(const int32_t* __restrict__ nodes, V3 q, int cur, int steps) {
  for (int s = 0; s < steps && cur >= 0; s++) {
    const int32_t* r = nodes + (int64_t)cur * kNodeInts;
    const int sc = r[0];
    if (sc == -1) return -1;
    int next = node_in(nodes, sc, q) ? sc : -1;
    const int nch = (r[1] >> 8) & 0xff;
    for (int i = 0; i < nch && next < 0; i++) {
      const int tc = r[2 + i];
      if (node_in(nodes, tc, q)) next = tc;
    }
    cur = next;
  }
  return cur;
}

// k_hint_band after a build: each hint point descends from its top-level
// cell (loc) to its polygon's last refined level and counts when the cell
// it reaches there is refined.  nodes: the query layout's node records,
// when packed (else the grid arrays).
__global__ void k_hint_band_last(GridView g, const int32_t* __restrict__ nodes, const PolyDev* __restrict__ pd,
                                 const double* __restrict__ xyz, const int32_t* __restrict__ hpoly,
                                 const int32_t* __restrict__ loc, int64_t n, int npoly, int64_t nbins,
                                 uint32_t* __restrict__ band) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int key = -1;
  if (t < n) {
    const int p = hpoly ? hpoly[t] : 0;
    int cur = loc[t];
    if (cur >= 0 && p >= 0 && p < npoly) {
      const V3 q = load3(xyz + 3 * t);
      cur = nodes ? descend_nodes_n(nodes, q, cur, last_refined(pd[p])) : descend(g, q, cur, last_refined(pd[p]));
      if (cur >= 0 && g.c_subcell[cur] != -1) key = g.c_poly[cur] * kGrid0Cells + g.c_root[cur];
    }
  }
  wave_add<uint32_t>(band + (blockIdx.x % kStripes) * nbins, key, 1u);
}

// Features 0 and 2..11 of (polygon p, top-level cell r) at `level` (feature
// 1 is the caller's: refine-or-stop or the start-table depth).
__device__ void cell_features(float* x, int level, int p, int pr, const PolyDev* __restrict__ pd,
                              const int32_t* __restrict__ st, const unsigned long long* __restrict__ estat,
                              const float* __restrict__ w, const uint32_t* __restrict__ hist,
                              const uint32_t* __restrict__ band) {
  const float Nr = (float)estat[3 * pr];
  const float Ur = (float)st[2 * pr + 1];
  const float Cr = (float)st[2 * pr];
  const double len = Nr > 0.f ? (double)estat[3 * pr + 1] / kFix / Nr : 0.0;
  const double turn = Nr > 0.f ? (double)estat[3 * pr + 2] / kFix / Nr : 0.0;
  float wsum = 0.f;
  for (int r = 0; r < kGrid0Cells; r++) wsum += w[p * kGrid0Cells + r];
  const double cell = 0.55 / (double)(1LL << min(level, 60));
  x[0] = level / 16.0f;
  x[2] = log2f(1.0f + Nr) / 20.0f;
  x[3] = log2f(1.0f + Ur) / 20.0f;
  x[4] = log2f(1.0f + Cr) / 20.0f;
  x[5] = fminf(Ur / fmaxf(Nr, 1.0f), 4.0f);
  x[6] = log2f(1.0f + Nr / fmaxf(Ur, 1.0f)) / 10.0f;
  x[7] = (len > 0.0) ? (float)(log2(len / cell) / 10.0) : 0.0f;
  x[8] = (float)(turn / M_PI);
  // share of r's points in the boundary band (crossed cells of this level):
  // from the hint sample, else the crossed share of r's cells (the same
  // quantity for evenly spread points)
  x[9] = hist ? (float)band[pr] / fmaxf((float)hist[pr], 1.0f) : Ur / fmaxf(Cr, 1.0f);
  x[10] = log2f(1.0f + (float)pd[p].N) / 20.0f;
  x[11] = wsum > 0.f ? w[pr] / wsum : Nr / fmaxf((float)pd[p].N, 1.0f);
}

// Start-table depth feature of candidate c: log2 of the bins per face side
// times the sub-table side (0 without a table), / 12.
__device__ __forceinline__ float table_depth_feature(int c) {
  return log2f(1.0f + (float)(kTableB[c] * kTableS[c])) / 12.0f;
}

// Features of rows (p, r, c) after a build, row = (p * 42 + r) * kTableCand + c,
// at each polygon's last refined level.
__global__ void k_table_features(int npoly, const PolyDev* __restrict__ pd, const int32_t* __restrict__ st,
                                 const unsigned long long* __restrict__ estat, const float* __restrict__ w,
                                 const uint32_t* __restrict__ hist, const uint32_t* __restrict__ band,
                                 float* __restrict__ X) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= npoly * kGrid0Cells * kTableCand) return;
  const int cand = row % kTableCand;
  const int pr = row / kTableCand;
  const int p = pr / kGrid0Cells;
  float* x = X + (int64_t)row * kMlpIn;
  cell_features(x, last_refined(pd[p]), p, pr, pd, st, estat, w, hist, band);
  x[1] = table_depth_feature(cand);
  // expected query density: the points expected in r (in place of the edges
  // per crossed cell, which features 3 and 5 already give)
  x[6] = log2f(1.0f + w[pr]) / 32.0f;
}

// Table build cost per candidate: a[c] + b[c] * N ns for one polygon; in a
// batch, whose launches are shared, kbin ns per table bin (6 B^2 bins).
struct TableCost {
  double a[kTableCand], b[kTableCand];
  double kbin;
};

// Start-table candidate per polygon: the least predicted build + query
// time, query time = query_scale * sum_r (expected points of r) * 2^y.
// Candidates beyond ncand, with more bins than bins_max or a sub-table side
// above sub_max are not considered.
__global__ void k_decide_table(int npoly, const PolyDev* __restrict__ pd, const float* __restrict__ Y,
                               const float* __restrict__ w, int ncand, int bins_max, int sub_max,
                               double query_scale, TableCost cost, int32_t* __restrict__ dec) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npoly) return;
  int best = 0;
  double best_t = 0.0;
  for (int c = 0; c < ncand; c++) {
    if (kTableB[c] > bins_max || kTableS[c] > max(sub_max, 1)) continue;
    double q = 0.0;
    for (int r = 0; r < kGrid0Cells; r++) {
      const int pr = p * kGrid0Cells + r;
      q += (double)w[pr] * exp2((double)Y[(int64_t)pr * kTableCand + c]);
    }
    const double build = npoly == 1 ? cost.a[c] + cost.b[c] * pd[p].N : cost.kbin * 6.0 * kTableB[c] * kTableB[c];
    const double t = query_scale * q + build;
    if (c == 0 || t < best_t) best = c, best_t = t;
  }
  dec[p] = best;
}

// Features of rows (p, r, k), row = (p * 42 + r) * 2 + k.
__global__ void k_features(int npoly, int level, const PolyDev* __restrict__ pd, const int32_t* __restrict__ st,
                           const unsigned long long* __restrict__ estat, const float* __restrict__ w,
                           const uint32_t* __restrict__ hist, const uint32_t* __restrict__ band,
                           float* __restrict__ X) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= npoly * kGrid0Cells * 2) return;
  const int k = row & 1;
  const int pr = row >> 1;
  const int p = pr / kGrid0Cells;
  float* x = X + (int64_t)row * kMlpIn;
  cell_features(x, level, p, pr, pd, st, estat, w, hist, band);
  x[1] = (float)k;
}

// Refine-or-stop per polygon still refining without a fixed depth.
// The per-level cost is shared by the polygons of a batch that still refine
// (n_refining), so each is charged its share.
__global__ void k_decide(int npoly, const PolyDev* __restrict__ pd, const int32_t* __restrict__ st,
                         const float* __restrict__ Y, const float* __restrict__ w, double query_scale,
                         double ns_edge, double ns_cell, double ns_level, int n_refining,
                         int32_t* __restrict__ dec) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npoly) return;
  const PolyDev& P = pd[p];
  if (!P.refining || P.depth != 0) {
    dec[p] = 1;
    return;
  }
  double saving = 0.0;
  for (int r = 0; r < kGrid0Cells; r++) {
    const int pr = p * kGrid0Cells + r;
    if (st[2 * pr + 1] == 0) continue;  // nothing of r is refined either way
    saving += (double)w[pr] * (exp2((double)Y[2 * pr]) - exp2((double)Y[2 * pr + 1]));
  }
  const double cost = ns_edge * P.N + ns_cell * ((double)P.U + P.nbc) + ns_level / max(n_refining, 1);
  // leave the density rule's choice only for a predicted gain of
  // kLearnMargin: before its stop level, stop only if the saving is below
  // cost / margin; at or past it, refine only if the saving exceeds
  // margin * cost
  const float k = P.U * 1.0 / P.N;
  const bool past_rule = k > 0.1 || (k < 0.1 && (0.1 - k) < 0.01);
  const double gain = saving * query_scale;
  const bool refine = past_rule ? gain > kLearnMargin * cost : gain * kLearnMargin > cost;
  dec[p] = (P.U > 0 && refine) ? 1 : 0;
}

int ensure_learn(gahhg_ctx* c) {
  const int64_t np = (int64_t)c->polys.size();
  LearnState& L = c->learn;
  if (L.cap >= np) return GAHHG_OK;
  void* ptrs[] = {L.d_estat, L.d_hist, L.d_band, L.d_w, L.d_stat, L.d_x, L.d_y, L.d_dec};
  for (void* q : ptrs) dfree(q);
  L.d_band = nullptr;
  L.d_estat = nullptr;
  L.d_hist = nullptr;
  L.d_w = nullptr;
  L.d_stat = nullptr;
  L.d_x = nullptr;
  L.d_y = nullptr;
  L.d_dec = nullptr;
  L.cap = 0;
  L.estat_valid = L.hist_valid = false;
  dfree(L.d_stripes);
  L.d_stripes = nullptr;
  const int64_t pr = np * kGrid0Cells;
  if (dalloc(c, (void**)&L.d_estat, pr * 3 * 8) || dalloc(c, (void**)&L.d_hist, pr * 4) ||
      dalloc(c, (void**)&L.d_stripes, pr * kStripes * 4) ||
      dalloc(c, (void**)&L.d_band, pr * 4) ||
      dalloc(c, (void**)&L.d_w, pr * 4) || dalloc(c, (void**)&L.d_stat, pr * 2 * 4) ||
      dalloc(c, (void**)&L.d_x, pr * 2 * kMlpIn * 4) || dalloc(c, (void**)&L.d_y, pr * 2 * 4) ||
      dalloc(c, (void**)&L.d_dec, np * 4))
    return GAHHG_E_NOMEM;
  L.cap = np;
  return GAHHG_OK;
}

unsigned blocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace learn_kernels
using namespace learn_kernels;

// Edge statistics and expected points per (polygon, top-level cell); needs
// the level-0 grid and the polygon records on the device.
int learn_prepare(gahhg_ctx* c) { return learn_prepare_with(c, c->learn.cost[0]); }

int learn_prepare_with(gahhg_ctx* c, double q_total) {
  LearnState& L = c->learn;
  int rc = ensure_learn(c);
  if (rc) return rc;
  const int np = (int)c->polys.size();
  if (!L.estat_valid) {
    rc = edge_stats(c, L.d_estat);
    if (rc) return rc;
    L.estat_valid = true;
  }
  const uint32_t* hist = nullptr;
  if (L.d_hint_xyz && L.hint_n > 0) {
    const int64_t nbins = (int64_t)np * kGrid0Cells;
    if (L.hloc_cap < L.hint_n) {
      dfree(L.d_hloc);
      dfree(L.d_hloc0);
      L.d_hloc = L.d_hloc0 = nullptr;
      L.hloc_cap = 0;
      L.hist_valid = false;
      if (dalloc(c, (void**)&L.d_hloc, L.hint_n * 4) || dalloc(c, (void**)&L.d_hloc0, L.hint_n * 4))
        return GAHHG_E_NOMEM;
      L.hloc_cap = L.hint_n;
    }
    if (!L.hist_valid) {
      rc = fill_i32(c, (int32_t*)L.d_stripes, nbins * kStripes, 0);
      if (rc) return rc;
      hipLaunchKernelGGL(k_hint_hist, dim3(blocks(L.hint_n, 256)), dim3(256), 0, c->stream, view(c),
                         (const PolyDev*)c->d_pd, (const double*)L.d_hint_xyz, (const int32_t*)L.d_hint_poly,
                         L.hint_n, np, L.d_stripes, L.d_hloc0);
      rc = hip_check(c, hipGetLastError(), "hint histogram");
      if (rc) return rc;
      hipLaunchKernelGGL(k_sum_stripes, dim3(blocks(nbins, 256)), dim3(256), 0, c->stream,
                         (const uint32_t*)L.d_stripes, nbins, L.d_hist);
      rc = hip_check(c, hipGetLastError(), "hint histogram sum");
      if (rc) return rc;
      L.hist_valid = true;
    }
    rc = copy_i32(c, L.d_hloc, L.d_hloc0, L.hint_n);  // the band descent starts at the top-level cells
    if (rc) return rc;
    L.hloc_level = 0;
    hist = L.d_hist;
  }
  const double q_default = np > 0 ? q_total / np : 0.0;
  hipLaunchKernelGGL(k_weights, dim3(blocks((int64_t)np * kGrid0Cells, 256)), dim3(256), 0, c->stream, np,
                     (const PolyDev*)c->d_pd, hist, L.hint_weight, q_default,
                     (const unsigned long long*)L.d_estat, L.d_w);
  return hip_check(c, hipGetLastError(), "point weights");
}

// Features of level `level` over the cell range [lo, lo + n) into L.d_x.
static int learn_features(gahhg_ctx* c, int level, int lo, int n, int marked, int only_refining) {
  LearnState& L = c->learn;
  const int np = (int)c->polys.size();
  int rc = fill_i32(c, L.d_stat, (int64_t)np * kGrid0Cells * 2, 0);
  if (rc) return rc;
  if (n > 0) {
    hipLaunchKernelGGL(k_level_stats, dim3(blocks(n, 256)), dim3(256), 0, c->stream, view(c),
                       (const PolyDev*)c->d_pd, lo, n, marked, only_refining, L.d_stat);
    rc = hip_check(c, hipGetLastError(), "level stats");
    if (rc) return rc;
  }
  const bool hint = L.d_hint_xyz && L.hint_n > 0;
  if (hint) {
    const int64_t nbins = (int64_t)np * kGrid0Cells;
    rc = fill_i32(c, (int32_t*)L.d_stripes, nbins * kStripes, 0);
    if (rc) return rc;
    if (level < L.hloc_level) return set_err(c, GAHHG_E_STATE, "features: levels must not decrease");
    hipLaunchKernelGGL(k_hint_band, dim3(blocks(L.hint_n, 256)), dim3(256), 0, c->stream, view(c),
                       (const double*)L.d_hint_xyz, L.d_hloc, L.hint_n, level - L.hloc_level, marked, nbins,
                       L.d_stripes);
    L.hloc_level = level;
    rc = hip_check(c, hipGetLastError(), "hint band");
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum_stripes, dim3(blocks(nbins, 256)), dim3(256), 0, c->stream,
                       (const uint32_t*)L.d_stripes, nbins, L.d_band);
    rc = hip_check(c, hipGetLastError(), "hint band sum");
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_features, dim3(blocks((int64_t)np * kGrid0Cells * 2, 256)), dim3(256), 0, c->stream, np,
                     level, (const PolyDev*)c->d_pd, (const int32_t*)L.d_stat, (const unsigned long long*)L.d_estat,
                     (const float*)L.d_w, hint ? (const uint32_t*)L.d_hist : nullptr,
                     hint ? (const uint32_t*)L.d_band : nullptr, L.d_x);
  return hip_check(c, hipGetLastError(), "features");
}

// Decisions for level `level` (cells [lo, lo + n) just marked and planned);
// queued on the stream into h_dec, valid after the next synchronisation.
int learn_level(gahhg_ctx* c, int level, int lo, int n, int32_t* h_dec) {
  LearnState& L = c->learn;
  const int np = (int)c->polys.size();
  int n_refining = 0;
  for (const PolyDev& P : c->pd) n_refining += P.refining != 0;
  int rc = learn_features(c, level, lo, n, 1, 1);
  if (!rc) rc = mlp_forward(c, L.d_x, (int64_t)np * kGrid0Cells * 2, L.d_weights, L.d_y);
  if (rc) return rc;
  hipLaunchKernelGGL(k_decide, dim3(blocks(np, 64)), dim3(64), 0, c->stream, np, (const PolyDev*)c->d_pd,
                     (const int32_t*)L.d_stat, (const float*)L.d_y, (const float*)L.d_w, L.cost[1], L.cost[2],
                     L.cost[3], L.cost[4], n_refining, L.d_dec);
  rc = hip_check(c, hipGetLastError(), "decide");
  if (rc) return rc;
  return hip_check(c, hipMemcpyAsync(h_dec, L.d_dec, np * 4, hipMemcpyDeviceToHost, c->stream), "decisions");
}

// Features of a built hierarchy at `level` (refined cells stand for the
// cells the build crossed there), for training: X[np * 42 * 2 * 12].
int learn_features_after_build(gahhg_ctx* c, int level, float* h_x) {
  LearnState& L = c->learn;
  const int np = (int)c->polys.size();
  int rc = learn_prepare(c);
  if (rc) return rc;
  const bool in = level >= 0 && level < (int)c->glv.size();
  const LevelDesc G = in ? c->glv[level] : LevelDesc{0, 0, 0, 0, 0, 0, 0};
  rc = learn_features(c, level, G.cstart, G.cnum, 0, 0);
  if (!rc)
    rc = hip_check(c, hipMemcpyAsync(h_x, L.d_x, (int64_t)np * kGrid0Cells * 2 * kMlpIn * 4, hipMemcpyDeviceToHost,
                                     c->stream),
                   "features D2H");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "features sync");
  return rc;
}

// Features of every (polygon, top-level cell, table candidate) of the built
// hierarchy into L.d_tx (and h_x when non-null).
int table_features_after_build(gahhg_ctx* c, float* h_x) {
  LearnState& L = c->learn;
  const int np = (int)c->polys.size();
  int rc = learn_prepare_with(c, L.tcost[0]);
  if (rc) return rc;
  const int64_t pr = (int64_t)np * kGrid0Cells;
  if (L.tcap < np) {
    dfree(L.d_tx);
    dfree(L.d_ty);
    L.d_tx = L.d_ty = nullptr;
    L.tcap = 0;
    if (dalloc(c, (void**)&L.d_tx, pr * kTableCand * kMlpIn * 4) || dalloc(c, (void**)&L.d_ty, pr * kTableCand * 4))
      return GAHHG_E_NOMEM;
    L.tcap = np;
  }
  rc = fill_i32(c, L.d_stat, pr * 2, 0);
  if (rc) return rc;
  const LevelTab lt = level_tab(c);
  if (c->tot_c > 0) {
    hipLaunchKernelGGL(k_level_stats_last, dim3(blocks(c->tot_c, 256)), dim3(256), 0, c->stream, view(c),
                       (const PolyDev*)c->d_pd, lt, L.d_stat);
    rc = hip_check(c, hipGetLastError(), "level stats");
    if (rc) return rc;
  }
  const bool hint = L.d_hint_xyz && L.hint_n > 0;
  if (hint) {
    rc = fill_i32(c, (int32_t*)L.d_stripes, pr * kStripes, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_hint_band_last, dim3(blocks(L.hint_n, 256)), dim3(256), 0, c->stream, view(c),
                       c->prm.query_layout == 1 ? (const int32_t*)nullptr : (const int32_t*)c->q.nodes,
                       (const PolyDev*)c->d_pd, (const double*)L.d_hint_xyz, (const int32_t*)L.d_hint_poly,
                       (const int32_t*)L.d_hloc0, L.hint_n, np, pr, L.d_stripes);
    rc = hip_check(c, hipGetLastError(), "hint band");
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum_stripes, dim3(blocks(pr, 256)), dim3(256), 0, c->stream, (const uint32_t*)L.d_stripes,
                       pr, L.d_band);
    rc = hip_check(c, hipGetLastError(), "hint band sum");
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_table_features, dim3(blocks(pr * kTableCand, 256)), dim3(256), 0, c->stream, np,
                     (const PolyDev*)c->d_pd, (const int32_t*)L.d_stat, (const unsigned long long*)L.d_estat,
                     (const float*)L.d_w, hint ? (const uint32_t*)L.d_hist : nullptr,
                     hint ? (const uint32_t*)L.d_band : nullptr, L.d_tx);
  rc = hip_check(c, hipGetLastError(), "table features");
  if (rc || !h_x) return rc;
  rc = hip_check(c, hipMemcpyAsync(h_x, L.d_tx, pr * kTableCand * kMlpIn * 4, hipMemcpyDeviceToHost, c->stream),
                 "features D2H");
  return rc ? rc : hip_check(c, hipStreamSynchronize(c->stream), "features sync");
}

// The learned start-table candidate of every polygon (one synchronisation).
int table_choose(gahhg_ctx* c, int32_t* h_choice) {
  LearnState& L = c->learn;
  const int np = (int)c->polys.size();
  int rc = table_features_after_build(c, nullptr);
  if (rc) return rc;
  const int64_t rows = (int64_t)np * kGrid0Cells * kTableCand;
  rc = mlp_forward(c, L.d_tx, rows, L.d_tweights, L.d_ty);
  if (rc) return rc;
  TableCost tc;
  for (int k = 0; k < kTableCand; k++) tc.a[k] = L.tcost[2 + 2 * k], tc.b[k] = L.tcost[3 + 2 * k];
  tc.kbin = L.tcost[2 + 2 * kTableCand];
  const int ncand = np == 1 ? kTableCand : kTableCandBatch;
  const int bins_max = np == 1 ? c->prm.query_bins : std::min(c->prm.query_bins, 128);
  const int sub_max = np == 1 ? c->prm.query_sub : 1;
  hipLaunchKernelGGL(k_decide_table, dim3(blocks(np, 64)), dim3(64), 0, c->stream, np, (const PolyDev*)c->d_pd,
                     (const float*)L.d_ty, (const float*)L.d_w, ncand, bins_max, sub_max, L.tcost[1], tc, L.d_dec);
  rc = hip_check(c, hipGetLastError(), "table decision");
  if (!rc) rc = hip_check(c, hipMemcpyAsync(h_choice, L.d_dec, np * 4, hipMemcpyDeviceToHost, c->stream), "decisions");
  return rc ? rc : hip_check(c, hipStreamSynchronize(c->stream), "decision sync");
}

void learn_free(gahhg_ctx* c) {
  LearnState& L = c->learn;
  void* ptrs[] = {L.d_estat, L.d_hist, L.d_band, L.d_w, L.d_stat, L.d_x, L.d_y, L.d_dec, L.d_weights,
                  L.d_hint_xyz, L.d_hint_poly, L.d_hloc, L.d_hloc0, L.d_stripes, L.d_tweights, L.d_tx, L.d_ty};
  for (void* q : ptrs) dfree(q);
  L = LearnState();
}

// Loads this file's code object onto the current device (context set-up).
void load_code_learn() {
  hipFuncAttributes a;
  (void)hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&learn_kernels::k_hint_hist));
}

}  // namespace gahhg
