// gahhg_trace.hip -- build trace: per polygon, per top-level cell (the 42
// cells of level 0) and per level, the cell counts that describe how the
// hierarchy grew there, plus per top-level cell aggregates of the polygon's
// edges.  No counterpart in the reference; these are the inputs of the
// learned refinement model and of the query-cost traces.
#include <string.h>

#include <algorithm>
#include <vector>

#include "gahhg_internal.h"

#pragma clang fp contract(off)

namespace gahhg {

// A named namespace (not an anonymous one) so that profilers show the
// kernels' names.
namespace trace_kernels {

__device__ __forceinline__ int level_of_cell(const LevelTab& lt, int cid) {
  int l = 0;
  while (l + 1 < lt.n && cid >= lt.cstart[l + 1]) l++;
  return l;
}

// out[((p * kMaxLevels + l) * 42 + r) * 4 + k], k = {cells, refined cells,
// bottom cells with refs, refs of those cells}; wave-aggregated adds (every
// lane reaches them, lanes past the end add to key -1).
__global__ void k_cell_trace(GridView g, const PolyDev* __restrict__ pd, LevelTab lt, int64_t C,
                             int32_t* __restrict__ out) {
  const int64_t cid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int key = -1, refined = 0, leaf = 0, refs = 0;
  if (cid < C) {
    const int p = g.c_poly[cid];
    const int l = level_of_cell(lt, (int)cid);
    key = (int)((((int64_t)p * kMaxLevels + l) * kGrid0Cells + g.c_root[cid]) * 4);
    refined = g.c_subcell[cid] != -1 ? 1 : 0;
    const PolyDev& P = pd[p];
    if (cid >= P.bcs && cid < P.bcs + P.bcn && g.c_rnum[cid] > 0) {
      leaf = 1;
      refs = g.c_rnum[cid];
    }
  }
  wave_add<int32_t>(out, key, key >= 0 ? 1 : 0);
  wave_add<int32_t>(out, key >= 0 ? key + 1 : -1, refined);
  wave_add<int32_t>(out, key >= 0 ? key + 2 : -1, leaf);
  wave_add<int32_t>(out, key >= 0 ? key + 3 : -1, refs);
}

}  // namespace trace_kernels
using namespace trace_kernels;

// Edge aggregates per top-level cell of the edge's first endpoint (first
// matching level-0 cell): out[(p * 42 + r) * 3 + k] = {edges, sum of arc
// lengths, sum of turning angles to the next edge of the ring}, the sums in
// fixed point (2^-40 rad) so that they do not depend on summation order.
// Each wave takes kTraceEdges consecutive edges (kTraceEdges / 64 per
// lane, strided by 64 so that the loads stay coalesced) and keeps a running
// sum per lane while its edges stay in one (polygon, top-level cell); the
// wave adds its sums to the totals only when some lane's key changes and at
// its end.  Consecutive edges of a ring mostly share a key, so the totals'
// few addresses take one atomic per counter per kTraceEdges edges instead
// of one per 64 (those atomics, serialised on a handful of addresses, were
// the kernel's time).  The sums are integers, so the order does not matter.
constexpr int kTraceEdges = 1024;
__global__ void k_edge_trace(GridView g, const PolyDev* __restrict__ pd, const double* __restrict__ edge,
                             const int32_t* __restrict__ epoly, const int32_t* __restrict__ pe, int64_t Nt,
                             unsigned long long* __restrict__ out) {
  // every lane reaches the wave-aggregated adds (invalid lanes add to key -1)
  __shared__ Level0Lds l0;  // level-0 geometry is the same for every polygon
  level0_load(g, pd[0].cell0, &l0);
  const int lane = threadIdx.x & 63;
  const int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * kTraceEdges;
  int acc = -1;
  unsigned long long n_acc = 0, len_acc = 0, turn_acc = 0;
  for (int j = 0; j < kTraceEdges; j += 64) {
    const int64_t e = base + j + lane;
    int key = -1;
    unsigned long long len_fx = 0, turn_fx = 0;
    if (e < Nt) {
      const int p = epoly[e];
      const V3 p0 = load3(edge + 9 * e), p1 = load3(edge + 9 * e + 3), ne = load3(edge + 9 * e + 6);
      const int root = level0_cell(&l0, p0);
      if (root >= 0) {
        const double len = acos(fmin(1.0, fmax(-1.0, dot(p0, p1))));
        double turn = 0.0;
        if (e + 1 < Nt && epoly[e + 1] == p && pe[2 * e + 1] == pe[2 * (e + 1)]) {
          const V3 n2 = load3(edge + 9 * (e + 1) + 6);
          turn = acos(fmin(1.0, fmax(-1.0, dot(ne, n2))));
        }
        key = (p * kGrid0Cells + root) * 3;
        len_fx = (unsigned long long)(len * kFix);
        turn_fx = (unsigned long long)(turn * kFix);
      }
    }
    if (__ballot(key >= 0 && acc >= 0 && key != acc)) {  // wave-uniform
      wave_add<unsigned long long>(out, acc, n_acc);
      wave_add<unsigned long long>(out, acc >= 0 ? acc + 1 : -1, len_acc);
      wave_add<unsigned long long>(out, acc >= 0 ? acc + 2 : -1, turn_acc);
      acc = -1;
      n_acc = len_acc = turn_acc = 0;
    }
    if (key >= 0) {
      acc = key;
      n_acc += 1;
      len_acc += len_fx;
      turn_acc += turn_fx;
    }
  }
  wave_add<unsigned long long>(out, acc, n_acc);
  wave_add<unsigned long long>(out, acc >= 0 ? acc + 1 : -1, len_acc);
  wave_add<unsigned long long>(out, acc >= 0 ? acc + 2 : -1, turn_acc);
}

static inline unsigned edge_trace_blocks(int64_t Nt) {
  const int64_t per_block = 4LL * kTraceEdges;  // 256 threads = 4 waves
  return (unsigned)std::max<int64_t>((Nt + per_block - 1) / per_block, 1);
}

int edge_stats(gahhg_ctx* c, unsigned long long* d_out) {
  const int64_t Nt = c->n_edges;
  int rc = fill_i32(c, (int32_t*)d_out, (int64_t)c->polys.size() * kGrid0Cells * 3 * 2, 0);
  if (rc || Nt <= 0) return rc;
  hipLaunchKernelGGL(k_edge_trace, dim3(edge_trace_blocks(Nt)), dim3(256), 0, c->stream, view(c),
                     (const PolyDev*)c->d_pd, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                     (const int32_t*)c->d_pe, Nt, d_out);
  return hip_check(c, hipGetLastError(), "edge stats");
}

int build_trace(gahhg_ctx* c, int32_t* h_cells, uint64_t* h_edges) {
  const int np = (int)c->polys.size();
  const int64_t ncell = (int64_t)np * kMaxLevels * kGrid0Cells * 4;
  const int64_t nedge = (int64_t)np * kGrid0Cells * 3;
  int32_t* d_cells = nullptr;
  unsigned long long* d_edges = nullptr;
  int rc = dalloc(c, (void**)&d_cells, ncell * 4);
  if (!rc) rc = dalloc(c, (void**)&d_edges, nedge * 8);
  if (!rc) rc = fill_i32(c, d_cells, ncell, 0);
  if (!rc) rc = fill_i32(c, (int32_t*)d_edges, nedge * 2, 0);
  if (!rc) {
    GridView g = view(c);
    const int64_t C = c->tot_c;
    hipLaunchKernelGGL(k_cell_trace, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, c->stream, g,
                       (const PolyDev*)c->d_pd, level_tab(c), C, d_cells);
    rc = hip_check(c, hipGetLastError(), "cell trace");
    const int64_t Nt = c->n_edges;
    if (!rc && Nt > 0) {
      hipLaunchKernelGGL(k_edge_trace, dim3(edge_trace_blocks(Nt)), dim3(256), 0, c->stream, g,
                         (const PolyDev*)c->d_pd, (const double*)c->d_edge, (const int32_t*)c->d_epoly,
                         (const int32_t*)c->d_pe, Nt, d_edges);
      rc = hip_check(c, hipGetLastError(), "edge trace");
    }
  }
  if (!rc && h_cells)
    rc = hip_check(c, hipMemcpyAsync(h_cells, d_cells, ncell * 4, hipMemcpyDeviceToHost, c->stream), "trace D2H");
  if (!rc && h_edges)
    rc = hip_check(c, hipMemcpyAsync(h_edges, d_edges, nedge * 8, hipMemcpyDeviceToHost, c->stream), "trace D2H");
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "trace sync");
  dfree(d_cells);
  dfree(d_edges);
  return rc;
}

// Loads this file's code object onto the current device (context set-up).
void load_code_trace() {
  hipFuncAttributes a;
  (void)hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&trace_kernels::k_cell_trace));
}

}  // namespace gahhg
