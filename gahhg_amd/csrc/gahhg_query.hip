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

#include <algorithm>

#include "gahhg_internal.h"

#pragma clang fp contract(off)

namespace gahhg {

constexpr int kQBlock = 256;

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

// Crossing test of the arc c -> p against polygon edge (a, b) with plane
// normals nq = arc_normal(c, p) and ne = arc_normal(a, b): the four dot
// products device_isEEInter evaluates, in its order.
__device__ __forceinline__ bool ee_cross(V3 nq, V3 ne, V3 a, V3 b, V3 c, V3 p) {
  if (same_side(dot(nq, a), dot(nq, b))) return false;
  if (same_side(dot(ne, c), dot(ne, p))) return false;
  return true;
}

// Straightforward query over the build layout.
__global__ void __launch_bounds__(kQBlock) k_query_soa(GridView g, const double* __restrict__ edge,
                                                       const double* __restrict__ xyz, int64_t n,
                                                       int32_t* __restrict__ labels) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const V3 p = load3(xyz + 3 * q);
  int cur = -1;
  for (int i = 0; i < kGrid0Cells; i++) {
    cur = q_inside(g, p, i);
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

__global__ void k_label_counts(const int32_t* __restrict__ labels, int64_t n, unsigned long long* out2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long in = 0, un = 0;
  if (i < n) {
    in = labels[i] == kIn;
    un = labels[i] == kUndefined;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    in += __shfl_xor(in, o, 64);
    un += __shfl_xor(un, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (in) atomicAdd(&out2[0], in);
    if (un) atomicAdd(&out2[1], un);
  }
}

int run_query(gahhg_ctx* c, const double* d_xyz, int64_t n, int32_t* d_labels) {
  if (n <= 0) return GAHHG_OK;
  const unsigned blocks = (unsigned)((n + c->prm.block_query - 1) / c->prm.block_query);
  hipLaunchKernelGGL(k_query_soa, dim3(blocks), dim3(c->prm.block_query), 0, c->stream, view(c),
                     (const double*)c->d_edge, d_xyz, n, d_labels);
  return hip_check(c, hipGetLastError(), "query kernel");
}

int count_labels(gahhg_ctx* c, const int32_t* d_labels, int64_t n, unsigned long long* d_out2) {
  if (n <= 0) return GAHHG_OK;
  hipLaunchKernelGGL(k_label_counts, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, d_labels, n,
                     d_out2);
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

int pack_query_layout(gahhg_ctx* c) {
  (void)c;
  return GAHHG_OK;
}

}  // namespace gahhg
