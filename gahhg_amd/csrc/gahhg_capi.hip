// gahhg_capi.hip -- the extern "C" boundary (include/gahhg.h).
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "gahhg_internal.h"

namespace gahhg {
int count_labels(gahhg_ctx* c, const int32_t* d_labels, int64_t n, unsigned long long* d_out2);
}

using namespace gahhg;

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

#define RCK(x)                       \
  do {                               \
    int _rc = (x);                   \
    if (_rc != GAHHG_OK) return _rc; \
  } while (0)

int ctx_ok(gahhg_ctx* c) { return c ? GAHHG_OK : GAHHG_E_ARG; }

int select_device(gahhg_ctx* c) { return hip_check(c, hipSetDevice(c->device), "hipSetDevice"); }

void free_grid(gahhg_ctx* c) {
  DeviceGrid& g = c->g;
  void* ptrs[] = {g.v, g.b_parent, g.b_partner, g.b_v0, g.b_v1, g.b_subcell, g.b_id, g.b_n, g.c_bstart,
                  g.c_bnum, g.c_rstart, g.c_rnum, g.c_upcell, g.c_subcell, g.c_type, g.c_poly, g.c_root, g.refs};
  for (void* p : ptrs) dfree(p);
  g = DeviceGrid();
  QueryLayout& q = c->q;
  DevBuf* bufs[] = {&q.b_table, &q.b_sub, &q.b_tmp, &q.b_pcnt, &q.b_qc,   &q.b_hdr, &q.b_ib,
                    &q.b_idx,   &q.b_sv,  &q.b_sn,  &q.b_sf,   &q.b_iscr, &q.b_tstart, &q.b_l0, &q.b_nodes};
  for (DevBuf* b : bufs) free_buf(*b);
  q = QueryLayout();
  drop_query_graph(c);
  c->qepoch++;
  c->built = false;
  c->gnum = 0;
}

// Before a rebuild of the same polygons: the grid arrays and the query
// layout's buffers are kept (a build writes every element it reads), only
// the layout's parts are marked absent, so a rebuild allocates nothing new
// unless it needs more room.
void reset_grid(gahhg_ctx* c) {
  QueryLayout& q = c->q;
  QueryLayout kept;
  kept.b_table = q.b_table, kept.b_sub = q.b_sub, kept.b_tmp = q.b_tmp, kept.b_pcnt = q.b_pcnt;
  kept.b_qc = q.b_qc, kept.b_hdr = q.b_hdr, kept.b_ib = q.b_ib, kept.b_idx = q.b_idx;
  kept.b_sv = q.b_sv, kept.b_sn = q.b_sn, kept.b_sf = q.b_sf, kept.b_iscr = q.b_iscr;
  kept.b_tstart = q.b_tstart, kept.b_l0 = q.b_l0, kept.b_nodes = q.b_nodes;
  kept.l0_have = q.l0_have;
  q = kept;
  c->qepoch++;
  c->built = false;
  c->gnum = 0;
}

void free_scratch(gahhg_ctx* c) {
  c->qepoch++;
  void* ptrs[] = {c->d_f0, c->d_f1, c->d_o0, c->d_o1, c->d_f2, c->d_f3, c->d_o2, c->d_o3, c->d_scan,
                  c->d_ecnt, c->d_rcell, c->d_rflag, c->d_redge, c->d_rscan, c->d_loc_cur, c->d_loc_prev};
  for (void* p : ptrs) dfree(p);
  c->d_f0 = c->d_f1 = c->d_o0 = c->d_o1 = c->d_scan = c->d_ecnt = nullptr;
  c->d_f2 = c->d_f3 = c->d_o2 = c->d_o3 = nullptr;
  c->d_rcell = c->d_rflag = c->d_redge = c->d_rscan = c->d_loc_cur = c->d_loc_prev = nullptr;
  dfree(c->d_wrec);
  c->d_wrec = nullptr;
  dfree(c->d_wslot);
  c->d_wslot = nullptr;
  c->fcap = c->scap = c->ecap = c->r2cap = c->wrec_cap = c->wslot_cap = 0;
}

int upload_polygons(gahhg_ctx* c) {
  (void)hipStreamSynchronize(c->stream);
  c->depth_override.clear();  // fixed depths belong to the previous polygons
  c->table_fixed.clear();
  c->learn.estat_valid = c->learn.hist_valid = false;
  free_grid(c);
  free_scratch(c);
  void* ptrs[] = {c->d_pv, c->d_pe, c->d_edge, c->d_epoly, c->d_pd};
  for (void* p : ptrs) dfree(p);
  c->d_pv = nullptr;
  c->d_pe = nullptr;
  c->d_edge = nullptr;
  c->d_epoly = nullptr;
  c->d_pd = nullptr;
  c->have_poly = false;
  int64_t N = 0;
  for (const Polygon& p : c->polys) N += p.N;
  if (N > INT32_MAX / 2) return set_err(c, GAHHG_E_LIMIT, "batch: too many edges");
  c->n_edges = N;
  std::vector<double> pv((size_t)(3 * N));
  std::vector<int32_t> pe((size_t)(2 * N)), ep((size_t)N);
  int64_t off = 0;
  for (size_t k = 0; k < c->polys.size(); k++) {
    const Polygon& p = c->polys[k];
    std::copy(p.pv.begin(), p.pv.end(), pv.begin() + 3 * off);
    for (int64_t i = 0; i < 2LL * p.N; i++) pe[2 * off + i] = (int32_t)(p.pe[i] + off);
    for (int64_t i = 0; i < p.N; i++) ep[off + i] = (int32_t)k;
    off += p.N;
  }
  RCK(dalloc(c, (void**)&c->d_pv, N * 24));
  RCK(dalloc(c, (void**)&c->d_pe, N * 8));
  RCK(dalloc(c, (void**)&c->d_epoly, N * 4));
  RCK(dalloc(c, (void**)&c->d_edge, N * 72));
  RCK(dalloc(c, (void**)&c->d_loc_cur, N * 4));
  RCK(dalloc(c, (void**)&c->d_loc_prev, N * 4));
  RCK(dalloc(c, (void**)&c->d_pd, (int64_t)c->polys.size() * (int64_t)sizeof(PolyDev)));
  RCK(hip_check(c, hipMemcpy(c->d_pv, pv.data(), N * 24, hipMemcpyHostToDevice), "polygon H2D"));
  RCK(hip_check(c, hipMemcpy(c->d_pe, pe.data(), N * 8, hipMemcpyHostToDevice), "polygon H2D"));
  RCK(hip_check(c, hipMemcpy(c->d_epoly, ep.data(), N * 4, hipMemcpyHostToDevice), "polygon H2D"));
  c->have_poly = true;
  return GAHHG_OK;
}

// Per-polygon statistics in the reference's terms (GPISP::getStat).
void poly_stats(gahhg_ctx* c, int p, gahhg_stats* o) {
  memset(o, 0, sizeof(*o));
  const std::vector<LevelDesc>& L = c->plv[p];
  const int depth = c->pd[p].L + 1;
  for (int l = 0; l < depth; l++) {
    o->vnum += L[l].vnum;
    o->bnum += L[l].bnum;
    o->cnum += L[l].cnum;
  }
  o->rnum = c->pd[p].R;
  o->depth = depth;
  o->gamma = c->pgamma[p];
  o->n_edges = c->polys[p].N;
  o->vspace = 24 * o->vnum;
  o->bspace = 48 * o->bnum;
  o->cspace = 28 * o->cnum;
  o->rspace = 4 * o->rnum;
  o->gpu_space_bytes = o->vspace + o->bspace + o->cspace + o->rspace;
  // sizeof(GAHHG) + gnum * sizeof(GAHHGGrid) + sizeof(GPISPParam) +
  // sizeof(GPISPStat) on the reference's 64-bit build (144 + 28 gnum + 552
  // + 96), GPISPAHHG.cu:2100.
  o->cpu_space_bytes = 792 + 28LL * depth;
  o->orientation_flipped = c->polys[p].flips;
  o->mark_passes = p < (int)c->pmarks.size() ? c->pmarks[p] : 0;
}

// Segment map from global to one polygon's local indices.
struct SegMap {
  std::vector<int64_t> g0, g1, l0;
  int32_t map(int32_t x) const {
    if (x < 0) return x;
    for (size_t k = 0; k < g0.size(); k++)
      if (x >= g0[k] && x < g1[k]) return (int32_t)(l0[k] + (x - g0[k]));
    return INT32_MIN;  // outside this polygon (never expected)
  }
};

// Export polygon p's hierarchy in its local (single-polygon reference)
// numbering.
int export_local(gahhg_ctx* c, int p, int64_t* counts, double* v, int32_t* cells, int32_t* borders,
                 double* normals, int32_t* refs, int32_t* levels) {
  RCK(hip_check(c, hipStreamSynchronize(c->stream), "sync"));
  gahhg_stats ps;
  poly_stats(c, p, &ps);
  const int depth = ps.depth;
  if (counts) {
    counts[0] = depth;
    counts[1] = ps.vnum;
    counts[2] = ps.bnum;
    counts[3] = ps.cnum;
    counts[4] = ps.rnum;
  }
  if (!v && !cells && !borders && !normals && !refs && !levels) return GAHHG_OK;
  const std::vector<LevelDesc>& L = c->plv[p];
  SegMap mv, mb, mc;
  int64_t lv = 0, lb = 0, lc = 0;
  for (int l = 0; l < depth; l++) {
    mv.g0.push_back(L[l].vstart), mv.g1.push_back(L[l].vstart + L[l].vnum), mv.l0.push_back(lv);
    mb.g0.push_back(L[l].bstart), mb.g1.push_back(L[l].bstart + L[l].bnum), mb.l0.push_back(lb);
    mc.g0.push_back(L[l].cstart), mc.g1.push_back(L[l].cstart + L[l].cnum), mc.l0.push_back(lc);
    if (levels) {
      int32_t* o = levels + 7 * l;
      o[0] = (int32_t)lv, o[1] = L[l].vnum, o[2] = (int32_t)lb, o[3] = L[l].bnum, o[4] = (int32_t)lc,
      o[5] = L[l].cnum, o[6] = L[l].rnum;
    }
    lv += L[l].vnum, lb += L[l].bnum, lc += L[l].cnum;
  }
  auto d2h = [&](void* h, const void* d, int64_t bytes) {
    return hip_check(c, hipMemcpy(h, d, (size_t)bytes, hipMemcpyDeviceToHost), "export");
  };
  const PolyDev& P = c->pd[p];
  // this polygon's bottom cells are contiguous, so their refs form one run
  // of the cell-major ref array, starting at the first bottom cell's offset
  int32_t rbase = 0;
  RCK(d2h(&rbase, c->g.c_rstart + P.bcs, 4));
  for (int l = 0; l < depth; l++) {
    const LevelDesc& D = L[l];
    const int64_t V = D.vnum, B = D.bnum, C = D.cnum;
    if (v) RCK(d2h(v + 3 * mv.l0[l], c->g.v + 3 * (int64_t)D.vstart, V * 24));
    if (normals) RCK(d2h(normals + 3 * mb.l0[l], c->g.b_n + 3 * (int64_t)D.bstart, B * 24));
    std::vector<int32_t> col;
    if (cells) {
      col.resize((size_t)std::max<int64_t>(C, 1));
      int32_t* cs[7] = {c->g.c_bstart, c->g.c_bnum, c->g.c_rstart, c->g.c_rnum, c->g.c_upcell, c->g.c_subcell,
                        c->g.c_type};
      for (int k = 0; k < 7; k++) {
        RCK(d2h(col.data(), cs[k] + D.cstart, C * 4));
        for (int64_t i = 0; i < C; i++) {
          int32_t x = col[i];
          if (k == 0) x = mb.map(x);
          else if (k == 2 && x >= 0) x -= rbase;
          else if (k == 4 || k == 5) x = mc.map(x);
          cells[7 * (mc.l0[l] + i) + k] = x;
        }
      }
    }
    if (borders) {
      col.resize((size_t)std::max<int64_t>(B, 1));
      int32_t* bs[6] = {c->g.b_parent, c->g.b_partner, c->g.b_v0, c->g.b_v1, c->g.b_subcell, c->g.b_id};
      for (int k = 0; k < 6; k++) {
        RCK(d2h(col.data(), bs[k] + D.bstart, B * 4));
        for (int64_t i = 0; i < B; i++) {
          int32_t x = col[i];
          if (k == 0 || k == 4) x = mc.map(x);
          else if (k == 1) x = mb.map(x);
          else if (k == 2 || k == 3) x = mv.map(x);
          borders[6 * (mb.l0[l] + i) + k] = x;
        }
      }
    }
  }
  if (refs && P.R > 0) {
    RCK(d2h(refs, c->g.refs + rbase, (int64_t)P.R * 4));
    for (int64_t i = 0; i < P.R; i++) refs[i] -= P.e0;
  }
  return GAHHG_OK;
}

}  // namespace

extern "C" {

int gahhg_params_default(gahhg_params* p) {
  if (!p) return GAHHG_E_ARG;
  memset(p, 0, sizeof(*p));
  p->max_depth = 15;
  p->auto_depth = 1;
  p->block_create = 256;
  p->block_classify = 256;
  p->block_query = 256;
  p->device = 0;
  p->verbose = 0;
  p->query_layout = 0;
  p->query_bins = 256;
  p->query_group = 0;
  p->query_sub = 8;
  p->query_sub_budget = 4;
  p->query_sort = 1;
  p->query_sort_bits = 14;
  p->query_index = 1;
  p->query_xcd = 32;
  p->query_table_budget = 8;
  p->query_graph = 0;
  p->build_device = 1;
  p->build_graph = 1;
  p->query_persistent = 0;
  return GAHHG_OK;
}

int gahhg_create(gahhg_ctx** out, const gahhg_params* p) {
  if (!out) return GAHHG_E_ARG;
  *out = nullptr;
  gahhg_ctx* c = new (std::nothrow) gahhg_ctx();
  if (!c) return GAHHG_E_NOMEM;
  if (p)
    c->prm = *p;
  else
    gahhg_params_default(&c->prm);
  if (c->prm.max_depth < 1 || c->prm.max_depth > kMaxLevels - 1 || c->prm.block_query < 64 ||
      c->prm.block_query > 256 || (c->prm.block_query % 64) != 0 || c->prm.query_bins < 0 ||
      c->prm.query_bins > 4096 || c->prm.query_group < 0 || c->prm.query_group > 64 ||
      (c->prm.query_group & (c->prm.query_group - 1)) != 0 || c->prm.query_sub < 0 || c->prm.query_sub > 256 ||
      (c->prm.query_sub & (c->prm.query_sub - 1)) != 0 || c->prm.query_sub_budget < 0 || c->prm.query_sort < 0 ||
      c->prm.query_sort > 2 || c->prm.query_xcd < 0 || c->prm.query_xcd > 4096 ||
      c->prm.query_sort_bits < 1 || c->prm.query_sort_bits > 14 || c->prm.query_index < 0 ||
      c->prm.query_index > 1 || c->prm.query_table_budget < 0 || c->prm.query_graph < 0 ||
      c->prm.query_graph > 1 || c->prm.build_device < 0 || c->prm.build_device > 1 || c->prm.build_graph < 0 ||
      c->prm.build_graph > 1 || c->prm.query_persistent < 0 || c->prm.query_persistent > 1) {
    delete c;
    return GAHHG_E_ARG;
  }
  c->device = c->prm.device;
  memset(&c->st, 0, sizeof(c->st));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c->device || c->device < 0) {
    delete c;
    return GAHHG_E_HIP;
  }
  if (hipSetDevice(c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return GAHHG_E_HIP;
  }
  for (int i = 0; i < 4; i++) (void)hipEventCreate(&c->ev[i]);
  noalloc_init(c->device);
  load_code_build();
  load_code_classify();
  load_code_query();
  load_code_scan();
  load_code_sort();
  load_code_learn();
  load_code_mlp();
  load_code_load();
  load_code_trace();
  if (hipMalloc((void**)&c->d_st, sizeof(DevState)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_st, sizeof(DevState), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_cnt, 64, hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_cnt, hipEventDisableTiming) != hipSuccess) {
    delete c;
    return GAHHG_E_HIP;
  }
  *out = c;
  return GAHHG_OK;
}

void gahhg_destroy(gahhg_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_dev_build(c);
  dfree(c->q_queue);
  c->q_queue = nullptr;
  free_grid(c);
  free_scratch(c);
  dfree(c->qsort.buf);
  dfree(c->d_mlp_part);
  learn_free(c);
  free_text_load(c);
  dfree(c->d_pv);
  dfree(c->d_pe);
  dfree(c->d_edge);
  dfree(c->d_st);
  for (void* p : c->user_allocs) dfree(p);
  if (c->h_st) (void)hipHostFree(c->h_st);
  if (c->h_cnt) (void)hipHostFree(c->h_cnt);
  if (c->ev_cnt) (void)hipEventDestroy(c->ev_cnt);
  for (int i = 0; i < 4; i++)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* gahhg_last_error(const gahhg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int gahhg_load_polygon_txt(gahhg_ctx* c, const char* path) {
  RCK(ctx_ok(c));
  if (!path) return set_err(c, GAHHG_E_ARG, "null path");
  RCK(select_device(c));
  PolygonText t;
  RCK(parse_polygon_txt(path, &t, &c->err));
  // large polygons: the vertices' conversion on the device (same bits)
  std::vector<double> xyz;
  if ((int64_t)t.lo.size() >= kGpuConvertMin) {
    xyz.resize(3 * t.lo.size());
    RCK(lonlat_to_xyz_gpu(c, t.lo.data(), t.la.data(), (int64_t)t.lo.size(), xyz.data()));
  }
  Polygon p;
  RCK(polygon_from_text(t, xyz.empty() ? nullptr : xyz.data(), &p, &c->err));
  c->polys.clear();
  c->polys.push_back(std::move(p));
  return upload_polygons(c);
}

int gahhg_load_polygon_lune(gahhg_ctx* c, const char* path, double* ref_lonlat, int32_t* has_ref) {
  RCK(ctx_ok(c));
  if (!path) return set_err(c, GAHHG_E_ARG, "null path");
  RCK(select_device(c));
  PolygonText t;
  double ref[2] = {0.0, 0.0};
  bool hr = false;
  RCK(parse_polygon_lune(path, &t, ref, &hr, &c->err));
  std::vector<double> xyz;
  if ((int64_t)t.lo.size() >= kGpuConvertMin) {
    xyz.resize(3 * t.lo.size());
    RCK(lonlat_to_xyz_gpu(c, t.lo.data(), t.la.data(), (int64_t)t.lo.size(), xyz.data()));
  }
  Polygon p;
  RCK(polygon_from_text(t, xyz.empty() ? nullptr : xyz.data(), &p, &c->err));
  c->polys.clear();
  c->polys.push_back(std::move(p));
  if (ref_lonlat) {
    ref_lonlat[0] = ref[0];
    ref_lonlat[1] = ref[1];
  }
  if (has_ref) *has_ref = hr ? 1 : 0;
  return upload_polygons(c);
}

int gahhg_load_query_csv_device(gahhg_ctx* c, const char* path, const double** d_xyz, int64_t* n,
                                gahhg_load_info* info) {
  RCK(ctx_ok(c));
  if (!d_xyz || !n) return set_err(c, GAHHG_E_ARG, "null output");
  RCK(select_device(c));
  if (info) memset(info, 0, sizeof(*info));
  RCK(load_query_csv_device(c, path, n, info));
  *d_xyz = c->tl.xyz;
  return GAHHG_OK;
}

int gahhg_parse_query_csv_device(gahhg_ctx* c, const char* text, int64_t bytes, const double** d_xyz,
                                 int64_t* n, gahhg_load_info* info) {
  RCK(ctx_ok(c));
  if (!d_xyz || !n || bytes < 0 || (bytes > 0 && !text)) return set_err(c, GAHHG_E_ARG, "bad text arguments");
  RCK(select_device(c));
  if (info) memset(info, 0, sizeof(*info));
  RCK(parse_query_csv_device(c, text, bytes, n, info));
  *d_xyz = c->tl.xyz;
  return GAHHG_OK;
}

int gahhg_load_polygon_txt_device(gahhg_ctx* c, const char* path, gahhg_load_info* info) {
  RCK(ctx_ok(c));
  if (!path) return set_err(c, GAHHG_E_ARG, "null path");
  RCK(select_device(c));
  if (info) memset(info, 0, sizeof(*info));
  Polygon p;
  int rc = load_polygon_txt_device(c, path, &p, info);
  if (rc == GAHHG_E_FORMAT) {
    // a malformed file: the host loader reads it in file order and names
    // the first fault, as the reference does
    Polygon q;
    const int rc2 = load_polygon_txt(path, &q, &c->err);
    if (rc2) return rc2;
    return set_err(c, GAHHG_E_STATE, "device and host polygon loaders disagree on %s", path);
  }
  if (rc) return rc;
  c->polys.clear();
  c->polys.push_back(std::move(p));
  return upload_polygons(c);
}

int gahhg_load_polygons_txt(gahhg_ctx* c, const char* const* paths, int32_t n) {
  RCK(ctx_ok(c));
  if (n < 1 || !paths) return set_err(c, GAHHG_E_ARG, "batch: need polygon files");
  RCK(select_device(c));
  for (int k = 0; k < n; k++)
    if (!paths[k]) return set_err(c, GAHHG_E_ARG, "batch: null path %d", k);
  // files parsed on up to 16 host threads; the first failing file (in list
  // order) gives the error, as a sequential load would.  All vertices are
  // then converted to xyz in one device pass (the same bits as the host
  // conversion) and the polygons assembled on the threads again.
  std::vector<PolygonText> ts((size_t)n);
  std::vector<Polygon> ps((size_t)n);
  std::vector<int> rcs((size_t)n, GAHHG_OK);
  std::vector<std::string> errs((size_t)n);
  const int threads = std::max(1, std::min({16, n, (int)std::thread::hardware_concurrency()}));
  auto parallel = [&](auto&& fn) {
    std::atomic<int> next(0);
    auto work = [&]() {
      for (int k; (k = next.fetch_add(1)) < n;) fn(k);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) pool.emplace_back(work);
    work();
    for (std::thread& th : pool) th.join();
  };
  auto first_error = [&]() {
    for (int k = 0; k < n; k++)
      if (rcs[(size_t)k]) {
        c->err = errs[(size_t)k];
        return rcs[(size_t)k];
      }
    return (int)GAHHG_OK;
  };
  parallel([&](int k) { rcs[(size_t)k] = parse_polygon_txt(paths[k], &ts[(size_t)k], &errs[(size_t)k]); });
  RCK(first_error());
  std::vector<int64_t> voff((size_t)n + 1, 0);
  for (int k = 0; k < n; k++) voff[(size_t)k + 1] = voff[(size_t)k] + (int64_t)ts[(size_t)k].lo.size();
  const int64_t V = voff[(size_t)n];
  std::vector<double> xyz;
  if (V >= kGpuConvertMin) {
    std::vector<double> lo((size_t)V), la((size_t)V);
    for (int k = 0; k < n; k++) {
      std::copy(ts[(size_t)k].lo.begin(), ts[(size_t)k].lo.end(), lo.begin() + voff[(size_t)k]);
      std::copy(ts[(size_t)k].la.begin(), ts[(size_t)k].la.end(), la.begin() + voff[(size_t)k]);
    }
    xyz.resize((size_t)(3 * V));
    RCK(lonlat_to_xyz_gpu(c, lo.data(), la.data(), V, xyz.data()));
  }
  parallel([&](int k) {
    const double* x = xyz.empty() ? nullptr : xyz.data() + 3 * voff[(size_t)k];
    rcs[(size_t)k] = polygon_from_text(ts[(size_t)k], x, &ps[(size_t)k], &errs[(size_t)k]);
  });
  RCK(first_error());
  c->polys = std::move(ps);
  return upload_polygons(c);
}

int gahhg_set_polygon_lonlat(gahhg_ctx* c, const double* lon, const double* lat, const int64_t* ring_off,
                             int32_t n_rings) {
  RCK(ctx_ok(c));
  RCK(select_device(c));
  Polygon p;
  int rc = polygon_from_lonlat(lon, lat, ring_off, n_rings, &p, &c->err);
  if (rc) return rc;
  c->polys.clear();
  c->polys.push_back(std::move(p));
  return upload_polygons(c);
}

int gahhg_set_polygons_lonlat(gahhg_ctx* c, const double* lon, const double* lat, const int64_t* ring_off,
                              const int32_t* poly_ring_off, int32_t n_polys) {
  RCK(ctx_ok(c));
  if (n_polys < 1 || !poly_ring_off || !ring_off) return set_err(c, GAHHG_E_ARG, "batch: need polygons");
  RCK(select_device(c));
  std::vector<Polygon> ps((size_t)n_polys);
  for (int k = 0; k < n_polys; k++) {
    const int32_t r0 = poly_ring_off[k], r1 = poly_ring_off[k + 1];
    if (r1 <= r0) return set_err(c, GAHHG_E_ARG, "batch: polygon %d has no ring", k);
    int rc = polygon_from_lonlat(lon, lat, ring_off + r0, r1 - r0, &ps[k], &c->err);
    if (rc) return rc;
  }
  c->polys = std::move(ps);
  return upload_polygons(c);
}

int gahhg_build(gahhg_ctx* c) {
  RCK(ctx_ok(c));
  if (!c->have_poly) return set_err(c, GAHHG_E_STATE, "build: no polygon loaded");
  RCK(select_device(c));
  (void)hipStreamSynchronize(c->stream);
  reset_grid(c);
  const double t0 = now_ms();
  double t1 = t0, tl = t0;
  if (c->prm.build_device && !c->learn.on) {
    // level loop, dispatch and classification decided on the device
    bool replayed = false;
    RCK(build_classify_dev(c, &replayed, &c->build_attempts));
    t1 = tl = now_ms();
    c->build_replayed = replayed;
  } else {
    RCK(build_hierarchy(c));
    t1 = now_ms();
    RCK(classify_hierarchy(c));
    RCK(hip_check(c, hipStreamSynchronize(c->stream), "sync"));
    tl = now_ms();
    c->build_replayed = false;
    c->build_attempts = 0;
  }
  RCK(pack_query_layout(c));
  RCK(hip_check(c, hipStreamSynchronize(c->stream), "sync"));
  const double t2 = now_ms();
  c->built = true;
  if (c->polys.size() == 1) {
    poly_stats(c, 0, &c->st);
  } else {
    gahhg_stats tot, ps;
    memset(&tot, 0, sizeof(tot));
    for (int p = 0; p < (int)c->polys.size(); p++) {
      poly_stats(c, p, &ps);
      tot.vnum += ps.vnum, tot.bnum += ps.bnum, tot.cnum += ps.cnum, tot.rnum += ps.rnum;
      tot.vspace += ps.vspace, tot.bspace += ps.bspace, tot.cspace += ps.cspace, tot.rspace += ps.rspace;
      tot.gpu_space_bytes += ps.gpu_space_bytes;
      tot.cpu_space_bytes += ps.cpu_space_bytes;
      tot.depth = std::max(tot.depth, ps.depth);
      tot.mark_passes = std::max(tot.mark_passes, ps.mark_passes);
    }
    tot.n_edges = c->n_edges;
    c->st = tot;
  }
  c->st.create_ms = t1 - t0;
  c->st.classify_ms = t2 - t1;
  c->st.pre_ms = t2 - t0;
  c->st.layout_ms = t2 - tl;
  const DevState& s = *c->h_st;
  c->st.walk_steps_mark = (int64_t)s.steps_mark;
  c->st.walk_steps_dispatch = (int64_t)s.steps_dispatch;
  c->st.walk_partner_missing = (int64_t)s.partner_missing;
  c->st.walk_step_cap_hits = (int64_t)s.cap_hits;
  c->st.walk_unlocated = (int64_t)s.unlocated;
  c->st.undefined_leaves_snapshot = c->und_snap;
  c->st.undefined_leaves_final = c->und_final;
  c->st.build_replayed = c->build_replayed ? 1 : 0;
  c->st.build_attempts = c->build_attempts;
  return GAHHG_OK;
}

// ---- grid serialisation ----------------------------------------------------
// File: "GAHHGGR1", int32 version, polygons, PolyDev size; int64 V, B, C, R;
// int32 global levels + LevelDesc each; per polygon: N, flips, pv[3N], pe[2N],
// PolyDev record, its levels, gamma; then the device arrays in the order of
// grid_arrays(); finally the two undefined-leaf counts.  Native byte order.
}  // extern "C"

namespace {

constexpr char kGridMagic[8] = {'G', 'A', 'H', 'H', 'G', 'G', 'R', '1'};
constexpr int32_t kGridVersion = 2;

struct Arr {
  void* p;
  int64_t bytes;
};

std::vector<Arr> grid_arrays(gahhg_ctx* c, int64_t V, int64_t B, int64_t C, int64_t R) {
  DeviceGrid& g = c->g;
  return {{g.v, 24 * V},         {g.b_parent, 4 * B},  {g.b_partner, 4 * B}, {g.b_v0, 4 * B},
          {g.b_v1, 4 * B},       {g.b_subcell, 4 * B}, {g.b_id, 4 * B},      {g.b_n, 24 * B},
          {g.c_bstart, 4 * C},   {g.c_bnum, 4 * C},    {g.c_rstart, 4 * C},  {g.c_rnum, 4 * C},
          {g.c_upcell, 4 * C},   {g.c_subcell, 4 * C}, {g.c_type, 4 * C},    {g.c_poly, 4 * C},
          {g.c_root, 4 * C},     {g.refs, 4 * R}};
}

// Contents of the arrays of a grid file (in grid_arrays() order) checked
// against every range a kernel indexes with; returns what is wrong, or
// nullptr.  Levels: the global level ranges tile the cells in order, and a
// cell's centre child and border children lie on the next level and its
// parent on the previous one, so every descent ends within the levels.
const char* check_grid_arrays(const std::vector<std::vector<char>>& h, const int64_t* cnt,
                              const std::vector<LevelDesc>& glv, const std::vector<PolyDev>& pd, int64_t n_edges) {
  const int64_t V = cnt[0], B = cnt[1], C = cnt[2], R = cnt[3];
  auto I = [&](int k) { return (const int32_t*)h[k].data(); };
  const int32_t *b_parent = I(1), *b_partner = I(2), *b_v0 = I(3), *b_v1 = I(4), *b_sub = I(5), *b_id = I(6);
  const int32_t *c_bstart = I(8), *c_bnum = I(9), *c_rstart = I(10), *c_rnum = I(11), *c_up = I(12);
  const int32_t *c_sub = I(13), *c_type = I(14), *c_poly = I(15), *c_root = I(16), *refs = I(17);
  std::vector<int32_t> level((size_t)C, -1);
  int64_t at = 0;
  for (size_t l = 0; l < glv.size(); l++) {
    if (glv[l].cstart != at) return "level ranges do not tile the cells";
    for (int64_t k = 0; k < glv[l].cnum; k++) level[(size_t)(at + k)] = (int32_t)l;
    at += glv[l].cnum;
  }
  if (at != C) return "level ranges do not tile the cells";
  const int64_t np = (int64_t)pd.size();
  for (int64_t k = 0; k < C; k++) {
    const int l = level[(size_t)k];
    if (c_bnum[k] < 3 || c_bnum[k] > 6 || c_bstart[k] < 0 || (int64_t)c_bstart[k] + c_bnum[k] > B)
      return "cell border range";
    if (c_type[k] < 0 || c_type[k] > 2 || c_poly[k] < 0 || c_poly[k] >= np || c_root[k] < 0 ||
        c_root[k] >= kGrid0Cells)
      return "cell type, polygon or root";
    if (c_up[k] != -1 && (c_up[k] < 0 || c_up[k] >= C || level[(size_t)c_up[k]] != l - 1)) return "cell parent";
    if (c_sub[k] != -1 && (c_sub[k] < 0 || c_sub[k] >= C || level[(size_t)c_sub[k]] != l + 1))
      return "cell centre child";
    const bool leaf = c_sub[k] == -1;
    if (c_rnum[k] < 0 || (leaf && c_rnum[k] == 999 && c_rstart[k] < 0)) return "cell refs";
    if (c_rnum[k] > 0 && !(c_rnum[k] == 999 && !leaf) &&
        (c_rstart[k] < 0 || (int64_t)c_rstart[k] + c_rnum[k] > R))
      return "cell ref range";
  }
  for (int64_t b = 0; b < B; b++) {
    if (b_parent[b] < 0 || b_parent[b] >= C) return "border parent";
    if (b_partner[b] < -1 || b_partner[b] >= B) return "border partner";
    if (b_v0[b] < 0 || b_v0[b] >= V || b_v1[b] < 0 || b_v1[b] >= V) return "border vertex";
    if (b_id[b] < 0 || b_id[b] >= 6) return "border position";
    if (b_sub[b] != -1 &&
        (b_sub[b] < 0 || b_sub[b] >= C || level[(size_t)b_sub[b]] != level[(size_t)b_parent[b]] + 1))
      return "border child";
  }
  for (int64_t r = 0; r < R; r++)
    if (refs[r] < 0 || refs[r] >= n_edges) return "ref edge";
  return nullptr;
}

bool wr(FILE* f, const void* p, size_t n) { return n == 0 || fwrite(p, 1, n, f) == n; }
bool rd(FILE* f, void* p, size_t n) { return n == 0 || fread(p, 1, n, f) == n; }

void finish_stats(gahhg_ctx* c) {
  if (c->polys.size() == 1) {
    poly_stats(c, 0, &c->st);
  } else {
    gahhg_stats tot, ps;
    memset(&tot, 0, sizeof(tot));
    for (int p = 0; p < (int)c->polys.size(); p++) {
      poly_stats(c, p, &ps);
      tot.vnum += ps.vnum, tot.bnum += ps.bnum, tot.cnum += ps.cnum, tot.rnum += ps.rnum;
      tot.vspace += ps.vspace, tot.bspace += ps.bspace, tot.cspace += ps.cspace, tot.rspace += ps.rspace;
      tot.gpu_space_bytes += ps.gpu_space_bytes;
      tot.cpu_space_bytes += ps.cpu_space_bytes;
      tot.depth = std::max(tot.depth, ps.depth);
    }
    tot.n_edges = c->n_edges;
    c->st = tot;
  }
}

}  // namespace

extern "C" {

int gahhg_save_grid(gahhg_ctx* c, const char* path) {
  RCK(ctx_ok(c));
  if (!path) return set_err(c, GAHHG_E_ARG, "save: null path");
  if (!c->built) return set_err(c, GAHHG_E_STATE, "save before build");
  RCK(select_device(c));
  RCK(hip_check(c, hipStreamSynchronize(c->stream), "sync"));
  FILE* f = fopen(path, "wb");
  if (!f) return set_err(c, GAHHG_E_IO, "save: cannot open %s", path);
  const int32_t np = (int32_t)c->polys.size(), pds = (int32_t)sizeof(PolyDev);
  const int64_t cnt[4] = {c->tot_v, c->tot_b, c->tot_c, c->tot_r};
  const int32_t ng = (int32_t)c->glv.size();
  bool ok = wr(f, kGridMagic, 8) && wr(f, &kGridVersion, 4) && wr(f, &np, 4) && wr(f, &pds, 4) && wr(f, cnt, 32) &&
            wr(f, &ng, 4) && wr(f, c->glv.data(), ng * sizeof(LevelDesc));
  for (int p = 0; ok && p < np; p++) {
    const Polygon& P = c->polys[p];
    const int32_t nl = (int32_t)c->plv[p].size();
    ok = wr(f, &P.N, 4) && wr(f, &P.flips, 4) && wr(f, P.pv.data(), P.pv.size() * 8) &&
         wr(f, P.pe.data(), P.pe.size() * 4) && wr(f, &c->pd[p], sizeof(PolyDev)) && wr(f, &nl, 4) &&
         wr(f, c->plv[p].data(), nl * sizeof(LevelDesc)) && wr(f, &c->pgamma[p], 4);
  }
  std::vector<char> buf;
  for (const Arr& a : grid_arrays(c, cnt[0], cnt[1], cnt[2], cnt[3])) {
    if (!ok) break;
    buf.resize((size_t)a.bytes);
    if (a.bytes && hip_check(c, hipMemcpy(buf.data(), a.p, a.bytes, hipMemcpyDeviceToHost), "save D2H")) {
      fclose(f);
      return GAHHG_E_HIP;
    }
    ok = wr(f, buf.data(), (size_t)a.bytes);
  }
  const int64_t und[2] = {c->und_snap, c->und_final};
  ok = ok && wr(f, und, 16);
  if (fclose(f) != 0) ok = false;
  return ok ? GAHHG_OK : set_err(c, GAHHG_E_IO, "save: write to %s failed", path);
}

int gahhg_load_grid(gahhg_ctx* c, const char* path) {
  RCK(ctx_ok(c));
  if (!path) return set_err(c, GAHHG_E_ARG, "load: null path");
  RCK(select_device(c));
  FILE* f = fopen(path, "rb");
  if (!f) return set_err(c, GAHHG_E_IO, "load: cannot open %s", path);
  char magic[8];
  int32_t ver = 0, np = 0, pds = 0, ng = 0;
  int64_t cnt[4];
  bool ok = rd(f, magic, 8) && memcmp(magic, kGridMagic, 8) == 0 && rd(f, &ver, 4) && ver == kGridVersion &&
            rd(f, &np, 4) && np >= 1 && rd(f, &pds, 4) && pds == (int32_t)sizeof(PolyDev) && rd(f, cnt, 32) &&
            rd(f, &ng, 4) && ng >= 1 && ng <= kMaxLevels;
  for (int k = 0; ok && k < 4; k++) ok = cnt[k] >= 0 && cnt[k] <= INT32_MAX;
  std::vector<LevelDesc> glv(ok ? ng : 0);
  ok = ok && rd(f, glv.data(), ng * sizeof(LevelDesc));
  std::vector<Polygon> polys(ok ? np : 0);
  std::vector<PolyDev> pd(ok ? np : 0);
  std::vector<std::vector<LevelDesc>> plv(ok ? np : 0);
  std::vector<float> gam(ok ? np : 0);
  for (int p = 0; ok && p < np; p++) {
    Polygon& P = polys[p];
    int32_t nl = 0;
    ok = rd(f, &P.N, 4) && P.N >= 3 && P.N <= INT32_MAX / 4 && rd(f, &P.flips, 4);
    if (!ok) break;
    P.pv.resize(3 * (size_t)P.N);
    P.pe.resize(2 * (size_t)P.N);
    ok = rd(f, P.pv.data(), P.pv.size() * 8) && rd(f, P.pe.data(), P.pe.size() * 4) && rd(f, &pd[p], sizeof(PolyDev)) &&
         rd(f, &nl, 4) && nl >= 1 && nl <= kMaxLevels;
    if (!ok) break;
    plv[p].resize(nl);
    ok = rd(f, plv[p].data(), nl * sizeof(LevelDesc)) && rd(f, &gam[p], 4);
    for (int32_t v : P.pe) ok = ok && v >= 0 && v < P.N;
  }
  // ranges the kernels index with must lie inside the arrays
  int64_t e0 = 0;
  for (int p = 0; ok && p < np; p++) {
    const PolyDev& P = pd[p];
    ok = P.e0 == e0 && P.N == polys[p].N && P.cell0 >= 0 && P.cell0 + kGrid0Cells <= cnt[2] && P.bcs >= 0 &&
         P.bcn >= 0 && (int64_t)P.bcs + P.bcn <= cnt[2] && P.roff >= 0 && P.R >= 0 &&
         (int64_t)P.roff + P.R <= cnt[3] && P.L >= 0 && P.L < (int)plv[p].size();
    for (const LevelDesc& L : plv[p])
      ok = ok && L.cstart >= 0 && L.cnum >= 0 && (int64_t)L.cstart + L.cnum <= cnt[2] && L.bstart >= 0 &&
           L.bnum >= 0 && (int64_t)L.bstart + L.bnum <= cnt[1] && L.vstart >= 0 && (int64_t)L.vstart + L.vnum <= cnt[0];
    e0 += P.N;
  }
  for (const LevelDesc& L : glv) ok = ok && L.cstart >= 0 && L.cnum >= 0 && (int64_t)L.cstart + L.cnum <= cnt[2];
  if (!ok) {
    fclose(f);
    return set_err(c, GAHHG_E_FORMAT, "load: %s is not a grid file of this library version", path);
  }
  // every array is read and checked on the host before anything reaches
  // the device: the build, pack and query kernels index through them
  std::vector<std::vector<char>> host(18);
  {
    const int64_t V = cnt[0], B = cnt[1], C = cnt[2], R = cnt[3];
    const int64_t sizes[18] = {24 * V, 4 * B, 4 * B, 4 * B, 4 * B, 4 * B, 4 * B, 24 * B, 4 * C,
                               4 * C,  4 * C, 4 * C, 4 * C, 4 * C, 4 * C, 4 * C, 4 * C, 4 * R};
    for (int k = 0; k < 18 && ok; k++) {
      host[k].resize((size_t)sizes[k]);
      ok = rd(f, host[k].data(), (size_t)sizes[k]);
    }
  }
  int64_t und[2] = {0, 0};
  ok = ok && rd(f, und, 16);
  fclose(f);
  if (!ok) return set_err(c, GAHHG_E_FORMAT, "load: %s is truncated", path);
  const char* bad = check_grid_arrays(host, cnt, glv, pd, e0);
  if (bad) return set_err(c, GAHHG_E_FORMAT, "load: %s: %s", path, bad);
  c->polys = std::move(polys);
  int rc = upload_polygons(c);
  if (!rc) rc = attach_loaded_grid(c, cnt[0], cnt[1], cnt[2], cnt[3]);
  int k = 0;
  for (const Arr& a : grid_arrays(c, cnt[0], cnt[1], cnt[2], cnt[3])) {
    if (rc) break;
    if (a.bytes) rc = hip_check(c, hipMemcpy(a.p, host[k].data(), a.bytes, hipMemcpyHostToDevice), "load H2D");
    k++;
  }
  if (rc) {
    free_grid(c);
    return rc;
  }
  c->pd = std::move(pd);
  c->plv = std::move(plv);
  c->pgamma = std::move(gam);
  c->glv = std::move(glv);
  c->gnum = (int32_t)c->glv.size();
  c->tot_v = cnt[0];
  c->tot_b = cnt[1];
  c->tot_c = cnt[2];
  c->tot_r = cnt[3];
  c->und_snap = und[0];
  c->und_final = und[1];
  for (PolyDev& P : c->pd) P.tbl = -1, P.tbins = 0, P.trel = 0;
  memset(c->h_st, 0, sizeof(DevState));
  RCK(hip_check(c, hipMemcpyAsync(c->d_st, c->h_st, sizeof(DevState), hipMemcpyHostToDevice, c->stream), "state"));
  RCK(hip_check(c, hipMemcpyAsync(c->d_pd, c->pd.data(), c->pd.size() * sizeof(PolyDev), hipMemcpyHostToDevice,
                                  c->stream),
                "polygon upload"));
  const double t0 = now_ms();
  RCK(pack_query_layout(c));
  RCK(hip_check(c, hipStreamSynchronize(c->stream), "sync"));
  c->built = true;
  finish_stats(c);
  c->st.pre_ms = now_ms() - t0;  // the query layout packed after loading
  c->st.undefined_leaves_snapshot = c->und_snap;
  c->st.undefined_leaves_final = c->und_final;
  return GAHHG_OK;
}

int gahhg_query_batch_xyz_device(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n,
                                  int32_t* d_labels) {
  return query_device(c, d_xyz, d_poly, n, d_labels, nullptr);
}

}  // extern "C"

namespace gahhg {
int query_device(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels,
                 const SortScratch* ext) {
  RCK(ctx_ok(c));
  if (!c->built) return set_err(c, GAHHG_E_STATE, "query before build");
  if (n < 0 || (n > 0 && (!d_xyz || !d_labels))) return set_err(c, GAHHG_E_ARG, "query: bad arguments");
  if (!d_poly && c->polys.size() != 1)
    return set_err(c, GAHHG_E_ARG, "query: polygon ids required for a batch of %d", (int)c->polys.size());
  RCK(select_device(c));
  RCK(check_poly_ids(c, d_poly, n));
  RCK(hip_check(c, hipEventRecord(c->ev[0], c->stream), "event"));
  RCK(run_query_with(c, d_xyz, d_poly, n, d_labels, ext));
  RCK(hip_check(c, hipEventRecord(c->ev[1], c->stream), "event"));
  RCK(hip_check(c, hipEventSynchronize(c->ev[1]), "query sync"));
  float ms = 0.f, kms = 0.f, sms = 0.f;
  (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
  if (n > 0) {
    (void)hipEventElapsedTime(&kms, c->ev[2], c->ev[3]);
    (void)hipEventElapsedTime(&sms, c->ev[0], c->ev[2]);
  }
  c->st.query_ms = ms;
  c->st.query_kernel_ms = kms;
  c->st.query_sort_ms = sms;
  c->st.n_queries = n;
  c->st.query_unit_s = n > 0 ? ms * 1e-3 / (double)n : 0.0;
  // label census (outside the timed region)
  RCK(count_labels(c, d_labels, n, &c->d_st->q_in));
  unsigned long long h_two[2] = {0, 0};
  RCK(hip_check(c, hipMemcpyAsync(h_two, &c->d_st->q_in, 16, hipMemcpyDeviceToHost, c->stream), "census"));
  RCK(hip_check(c, hipStreamSynchronize(c->stream), "census sync"));
  c->st.inside = (int64_t)h_two[0];
  c->st.unlocated = (int64_t)h_two[1];
  c->st.outside = n - c->st.inside;
  return GAHHG_OK;
}
}  // namespace gahhg

extern "C" {

int gahhg_autotune_query(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels,
                         int32_t* out3) {
  RCK(ctx_ok(c));
  if (!c->built) return set_err(c, GAHHG_E_STATE, "autotune before build");
  if (n <= 0 || !d_xyz || !d_labels) return set_err(c, GAHHG_E_ARG, "autotune: bad arguments");
  if (!d_poly && c->polys.size() != 1)
    return set_err(c, GAHHG_E_ARG, "autotune: polygon ids required for a batch of %d", (int)c->polys.size());
  if (c->prm.query_layout != 0) return set_err(c, GAHHG_E_STATE, "autotune: the accelerated query is off");
  RCK(select_device(c));
  RCK(check_poly_ids(c, d_poly, n));
  const int groups[] = {0, 4, 8, 16, 32, 64};  // 0: each polygon's own width
  const int blocks[] = {64, 128, 256};  // the query kernels' launch bound is 256
  const gahhg_params keep = c->prm;
  float best = 0.f;
  int bg = keep.query_group, bb = keep.block_query;
  for (int b : blocks)
    for (int G : groups) {
      c->prm.block_query = b;
      c->prm.query_group = G;
      c->qepoch++;  // a recorded query pass holds the old launch shape
      float ms = 0.f;
      for (int rep = 0; rep < 3; rep++) {  // first run warms up; keep the faster of two
        RCK(hip_check(c, hipEventRecord(c->ev[0], c->stream), "event"));
        const int rc = run_query(c, d_xyz, d_poly, n, d_labels);
        if (rc) {
          c->prm = keep;
          c->qepoch++;
          return rc;
        }
        RCK(hip_check(c, hipEventRecord(c->ev[1], c->stream), "event"));
        RCK(hip_check(c, hipEventSynchronize(c->ev[1]), "autotune sync"));
        float t = 0.f;
        (void)hipEventElapsedTime(&t, c->ev[0], c->ev[1]);
        if (rep == 1 || (rep == 2 && t < ms)) ms = t;
      }
      if (best == 0.f || ms < best) {
        best = ms;
        bg = G;
        bb = b;
      }
    }
  c->prm.query_group = bg;
  c->prm.block_query = bb;
  c->qepoch++;
  if (out3) {
    out3[0] = bg;
    out3[1] = bb;
    out3[2] = (int32_t)(best * 1000.0f);  // microseconds
  }
  return GAHHG_OK;
}

int gahhg_autotune_layout(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels,
                          double passes, double* out6) {
  RCK(ctx_ok(c));
  if (!c->built) return set_err(c, GAHHG_E_STATE, "autotune before build");
  if (n <= 0 || !d_xyz || !d_labels || !(passes >= 1.0)) return set_err(c, GAHHG_E_ARG, "autotune: bad arguments");
  if (!d_poly && c->polys.size() != 1)
    return set_err(c, GAHHG_E_ARG, "autotune: polygon ids required for a batch of %d", (int)c->polys.size());
  if (c->prm.query_layout != 0) return set_err(c, GAHHG_E_STATE, "autotune: the accelerated query is off");
  RCK(select_device(c));
  RCK(check_poly_ids(c, d_poly, n));
  struct Cand {
    int bins, sub, sub_budget, table_budget;
  };
  std::vector<Cand> cands;
  const gahhg_params keep = c->prm;
  if (c->polys.size() == 1) {
    cands = {{128, 1, 4, keep.query_table_budget},
             {256, 1, 4, keep.query_table_budget},
             {256, 8, 4, keep.query_table_budget},
             {256, 16, 8, keep.query_table_budget}};
  } else {
    cands = {{keep.query_bins, keep.query_sub, keep.query_sub_budget, 8},
             {keep.query_bins, keep.query_sub, keep.query_sub_budget, 32},
             {keep.query_bins, keep.query_sub, keep.query_sub_budget, 0}};
  }
  int best = -1;
  double best_cost = 0.0, best_pack = 0.0, best_q = 0.0;
  auto pack = [&](const Cand& k, float* ms) {
    c->prm.query_bins = k.bins;
    c->prm.query_sub = k.sub;
    c->prm.query_sub_budget = k.sub_budget;
    c->prm.query_table_budget = k.table_budget;
    int rc = hip_check(c, hipEventRecord(c->ev[0], c->stream), "event");
    if (!rc) rc = pack_query_layout(c);
    if (!rc) rc = hip_check(c, hipEventRecord(c->ev[1], c->stream), "event");
    if (!rc) rc = hip_check(c, hipEventSynchronize(c->ev[1]), "autotune sync");
    if (!rc && ms) (void)hipEventElapsedTime(ms, c->ev[0], c->ev[1]);
    return rc;
  };
  for (size_t i = 0; i < cands.size(); i++) {
    float pms = 0.f;
    int rc = pack(cands[i], &pms);
    float qms = 0.f;
    for (int rep = 0; rep < 2 && !rc; rep++) {  // the first run warms up
      rc = hip_check(c, hipEventRecord(c->ev[0], c->stream), "event");
      if (!rc) rc = run_query(c, d_xyz, d_poly, n, d_labels);
      if (!rc) rc = hip_check(c, hipEventRecord(c->ev[1], c->stream), "event");
      if (!rc) rc = hip_check(c, hipEventSynchronize(c->ev[1]), "autotune sync");
      if (!rc) (void)hipEventElapsedTime(&qms, c->ev[0], c->ev[1]);
    }
    if (rc) {  // restore the context's own layout
      c->prm = keep;
      (void)pack_query_layout(c);
      return rc;
    }
    const double cost = pms / passes + qms;
    if (best < 0 || cost < best_cost) best = (int)i, best_cost = cost, best_pack = pms, best_q = qms;
  }
  RCK(pack(cands[(size_t)best], nullptr));
  if (out6) {
    const Cand& k = cands[(size_t)best];
    out6[0] = k.bins, out6[1] = k.sub, out6[2] = k.sub_budget, out6[3] = k.table_budget;
    out6[4] = best_pack, out6[5] = best_q;
  }
  RCK(run_query(c, d_xyz, d_poly, n, d_labels));
  return hip_check(c, hipStreamSynchronize(c->stream), "autotune sync");
}

int gahhg_trace_query_device(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n,
                             int32_t* d_labels, int32_t* d_trace) {
  RCK(ctx_ok(c));
  if (!c->built) return set_err(c, GAHHG_E_STATE, "trace before build");
  if (n < 0 || (n > 0 && (!d_xyz || !d_labels || !d_trace))) return set_err(c, GAHHG_E_ARG, "trace: bad arguments");
  if (!d_poly && c->polys.size() != 1)
    return set_err(c, GAHHG_E_ARG, "trace: polygon ids required for a batch of %d", (int)c->polys.size());
  RCK(select_device(c));
  RCK(check_poly_ids(c, d_poly, n));
  RCK(run_query_trace(c, d_xyz, d_poly, n, d_labels, d_trace));
  return hip_check(c, hipStreamSynchronize(c->stream), "trace sync");
}

int gahhg_build_trace(gahhg_ctx* c, int32_t* cells, uint64_t* edges) {
  RCK(ctx_ok(c));
  if (!c->built) return set_err(c, GAHHG_E_STATE, "build trace before build");
  RCK(select_device(c));
  return build_trace(c, cells, edges);
}

int gahhg_mlp_forward(gahhg_ctx* c, const float* d_x, int64_t rows, const uint16_t* d_w, float* d_y) {
  RCK(ctx_ok(c));
  if (rows < 0 || (rows > 0 && (!d_x || !d_w || !d_y))) return set_err(c, GAHHG_E_ARG, "mlp: bad arguments");
  RCK(select_device(c));
  RCK(mlp_forward(c, d_x, rows, d_w, d_y));
  return hip_check(c, hipStreamSynchronize(c->stream), "mlp sync");
}

int gahhg_mlp_backward(gahhg_ctx* c, const float* d_x, int64_t rows, const uint16_t* d_w, const float* d_dy,
                       float* d_grad) {
  RCK(ctx_ok(c));
  if (rows < 0 || !d_grad || (rows > 0 && (!d_x || !d_w || !d_dy)))
    return set_err(c, GAHHG_E_ARG, "mlp: bad arguments");
  RCK(select_device(c));
  RCK(mlp_backward(c, d_x, rows, d_w, d_dy, d_grad));
  return hip_check(c, hipStreamSynchronize(c->stream), "mlp sync");
}

int gahhg_mlp_adam(gahhg_ctx* c, float* d_master, const float* d_grad, float* d_m1, float* d_m2, int32_t step,
                   float lr, uint16_t* d_w) {
  RCK(ctx_ok(c));
  if (!d_master || !d_grad || !d_m1 || !d_m2 || !d_w || step < 1 || !(lr > 0.0f))
    return set_err(c, GAHHG_E_ARG, "mlp adam: bad arguments");
  RCK(select_device(c));
  RCK(mlp_adam(c, d_master, d_grad, d_m1, d_m2, step, lr, d_w));
  return hip_check(c, hipStreamSynchronize(c->stream), "mlp sync");
}

int gahhg_set_depth_model(gahhg_ctx* c, const float* params, const double* cost) {
  RCK(ctx_ok(c));
  RCK(select_device(c));
  LearnState& L = c->learn;
  if (!params) {
    L.on = false;
    c->built = false;
    return GAHHG_OK;
  }
  if (!cost) return set_err(c, GAHHG_E_ARG, "depth model: cost constants required");
  for (int i = 0; i < 5; i++)
    if (!(cost[i] >= 0.0)) return set_err(c, GAHHG_E_ARG, "depth model: cost[%d] must be >= 0", i);
  L.params.assign(params, params + kMlpParams);
  std::vector<uint16_t> w((size_t)kMlpParams + 8, 0);
  for (int k = 0; k < kMlpParams; k++) {
    uint32_t u;
    memcpy(&u, &params[k], 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return set_err(c, GAHHG_E_ARG, "depth model: non-finite parameter");
    u += 0x7fffu + ((u >> 16) & 1u);
    w[k] = (uint16_t)(u >> 16);
  }
  if (!L.d_weights) RCK(dalloc(c, (void**)&L.d_weights, (int64_t)w.size() * 2));
  RCK(hip_check(c, hipMemcpy(L.d_weights, w.data(), w.size() * 2, hipMemcpyHostToDevice), "model upload"));
  for (int i = 0; i < 5; i++) L.cost[i] = cost[i];
  L.on = true;
  c->built = false;
  return GAHHG_OK;
}

int gahhg_set_table_model(gahhg_ctx* c, const float* params, const double* cost) {
  RCK(ctx_ok(c));
  RCK(select_device(c));
  LearnState& L = c->learn;
  if (!params) {
    L.table_on = false;
    c->built = false;
    return GAHHG_OK;
  }
  if (!cost) return set_err(c, GAHHG_E_ARG, "table model: cost constants required");
  for (int i = 0; i < 3 + 2 * kTableCand; i++)
    if (!(cost[i] >= 0.0) || !std::isfinite(cost[i]))
      return set_err(c, GAHHG_E_ARG, "table model: cost[%d] must be finite and >= 0", i);
  std::vector<uint16_t> w((size_t)kMlpParams + 8, 0);
  for (int k = 0; k < kMlpParams; k++) {
    uint32_t u;
    memcpy(&u, &params[k], 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return set_err(c, GAHHG_E_ARG, "table model: non-finite parameter");
    u += 0x7fffu + ((u >> 16) & 1u);  // BF16, round to nearest even
    w[k] = (uint16_t)(u >> 16);
  }
  if (!L.d_tweights) RCK(dalloc(c, (void**)&L.d_tweights, (int64_t)w.size() * 2));
  RCK(hip_check(c, hipMemcpy(L.d_tweights, w.data(), w.size() * 2, hipMemcpyHostToDevice), "model upload"));
  L.tparams.assign(params, params + kMlpParams);
  for (int i = 0; i < 3 + 2 * kTableCand; i++) L.tcost[i] = cost[i];
  L.table_on = true;
  c->built = false;
  return GAHHG_OK;
}

int gahhg_set_table_choice(gahhg_ctx* c, const int32_t* cand, int32_t n) {
  RCK(ctx_ok(c));
  if (!cand || n == 0) {
    c->table_fixed.clear();
    c->built = false;
    return GAHHG_OK;
  }
  if (!c->have_poly || n != (int32_t)c->polys.size())
    return set_err(c, GAHHG_E_ARG, "table choice: %d values for %d polygons", n, (int)c->polys.size());
  for (int32_t k = 0; k < n; k++)
    if (cand[k] < -1 || cand[k] >= kTableCand)
      return set_err(c, GAHHG_E_ARG, "table choice: candidate %d of polygon %d outside -1..%d", cand[k], k,
                     kTableCand - 1);
  c->table_fixed.assign(cand, cand + n);
  c->built = false;
  return GAHHG_OK;
}

int gahhg_get_table_layout(gahhg_ctx* c, int32_t* bins_sub) {
  RCK(ctx_ok(c));
  if (!c->built) return set_err(c, GAHHG_E_STATE, "table layout before build");
  if (!bins_sub) return set_err(c, GAHHG_E_ARG, "table layout: null output");
  const size_t np = c->polys.size();
  for (size_t p = 0; p < np; p++) {
    bins_sub[2 * p] = c->table_used.size() == 2 * np ? c->table_used[2 * p] : 0;
    bins_sub[2 * p + 1] = c->table_used.size() == 2 * np ? c->table_used[2 * p + 1] : 1;
  }
  return GAHHG_OK;
}

int gahhg_table_features(gahhg_ctx* c, float* x) {
  RCK(ctx_ok(c));
  if (!c->built) return set_err(c, GAHHG_E_STATE, "table features before build");
  if (!x) return set_err(c, GAHHG_E_ARG, "table features: null output");
  RCK(select_device(c));
  return table_features_after_build(c, x);
}

int gahhg_set_query_hint(gahhg_ctx* c, const double* xyz, const int32_t* poly, int64_t n, double weight) {
  RCK(ctx_ok(c));
  if (n < 0 || (n > 0 && !xyz) || !(weight >= 0.0)) return set_err(c, GAHHG_E_ARG, "query hint: bad arguments");
  RCK(select_device(c));
  LearnState& L = c->learn;
  dfree(L.d_hint_xyz);
  dfree(L.d_hint_poly);
  L.d_hint_xyz = nullptr;
  L.d_hint_poly = nullptr;
  L.hint_n = 0;
  L.hist_valid = false;
  if (n == 0) return GAHHG_OK;
  RCK(dalloc(c, (void**)&L.d_hint_xyz, n * 24));
  RCK(hip_check(c, hipMemcpy(L.d_hint_xyz, xyz, n * 24, hipMemcpyHostToDevice), "hint upload"));
  if (poly) {
    RCK(dalloc(c, (void**)&L.d_hint_poly, n * 4));
    RCK(hip_check(c, hipMemcpy(L.d_hint_poly, poly, n * 4, hipMemcpyHostToDevice), "hint upload"));
  }
  L.hint_n = n;
  L.hint_weight = weight;
  return GAHHG_OK;  // the built hierarchy stays valid: the hint steers the next build
}

int gahhg_depth_features(gahhg_ctx* c, int32_t level, float* x) {
  RCK(ctx_ok(c));
  if (!c->built) return set_err(c, GAHHG_E_STATE, "depth features before build");
  if (!x || level < 0 || level >= kMaxLevels) return set_err(c, GAHHG_E_ARG, "depth features: bad arguments");
  RCK(select_device(c));
  return learn_features_after_build(c, level, x);
}

int gahhg_set_polygon_depths(gahhg_ctx* c, const int32_t* depths, int32_t n) {
  RCK(ctx_ok(c));
  if (n < 0 || (n > 0 && !depths)) return set_err(c, GAHHG_E_ARG, "depths: bad arguments");
  for (int32_t i = 0; i < n; i++)
    if (depths[i] < 0 || depths[i] > kMaxLevels - 1)
      return set_err(c, GAHHG_E_ARG, "depths: polygon %d depth %d outside 0..%d", i, depths[i], kMaxLevels - 1);
  c->depth_override.assign(depths, depths + n);
  c->built = false;
  return GAHHG_OK;
}

int gahhg_query_xyz_device(gahhg_ctx* c, const double* d_xyz, int64_t n, int32_t* d_labels) {
  return gahhg_query_batch_xyz_device(c, d_xyz, nullptr, n, d_labels);
}

int gahhg_query_batch_xyz(gahhg_ctx* c, const double* xyz, const int32_t* poly, int64_t n, int32_t* labels) {
  RCK(ctx_ok(c));
  if (!c->built) return set_err(c, GAHHG_E_STATE, "query before build");
  if (n < 0 || (n > 0 && (!xyz || !labels || !poly))) return set_err(c, GAHHG_E_ARG, "query: bad arguments");
  if (n == 0) return GAHHG_OK;
  for (int64_t i = 0; i < n; i++)
    if (poly[i] < 0 || poly[i] >= (int32_t)c->polys.size()) return set_err(c, GAHHG_E_ARG, "query: bad polygon id");
  RCK(select_device(c));
  double* d_xyz = nullptr;
  int32_t *d_lab = nullptr, *d_poly = nullptr;
  RCK(dalloc(c, (void**)&d_xyz, n * 24));
  int rc = dalloc(c, (void**)&d_lab, n * 4);
  if (!rc) rc = dalloc(c, (void**)&d_poly, n * 4);
  if (!rc) rc = hip_check(c, hipMemcpy(d_xyz, xyz, n * 24, hipMemcpyHostToDevice), "query H2D");
  if (!rc) rc = hip_check(c, hipMemcpy(d_poly, poly, n * 4, hipMemcpyHostToDevice), "query H2D");
  if (!rc) rc = gahhg_query_batch_xyz_device(c, d_xyz, d_poly, n, d_lab);
  if (!rc) rc = hip_check(c, hipMemcpy(labels, d_lab, n * 4, hipMemcpyDeviceToHost), "labels D2H");
  dfree(d_xyz);
  dfree(d_lab);
  dfree(d_poly);
  return rc;
}

int gahhg_query_pairs_xyz(gahhg_ctx* c, const double* xyz, int64_t n, int32_t* labels) {
  RCK(ctx_ok(c));
  if (!c->built) return set_err(c, GAHHG_E_STATE, "query before build");
  if (n < 0 || (n > 0 && (!xyz || !labels))) return set_err(c, GAHHG_E_ARG, "query: bad arguments");
  const int64_t total = n * (int64_t)c->polys.size();
  if (total == 0) return GAHHG_OK;
  if (total > INT32_MAX) return set_err(c, GAHHG_E_LIMIT, "query: %lld pairs exceed 2^31", (long long)total);
  RCK(select_device(c));
  double *d_in = nullptr, *d_xyz = nullptr;
  int32_t *d_lab = nullptr, *d_poly = nullptr;
  RCK(dalloc(c, (void**)&d_in, n * 24));
  int rc = dalloc(c, (void**)&d_xyz, total * 24);
  if (!rc) rc = dalloc(c, (void**)&d_lab, total * 4);
  if (!rc) rc = dalloc(c, (void**)&d_poly, total * 4);
  if (!rc) rc = hip_check(c, hipMemcpy(d_in, xyz, n * 24, hipMemcpyHostToDevice), "query H2D");
  if (!rc) rc = expand_pairs(c, d_in, n, d_xyz, d_poly);
  if (!rc) rc = gahhg_query_batch_xyz_device(c, d_xyz, d_poly, total, d_lab);
  if (!rc) rc = hip_check(c, hipMemcpy(labels, d_lab, total * 4, hipMemcpyDeviceToHost), "labels D2H");
  dfree(d_in);
  dfree(d_xyz);
  dfree(d_lab);
  dfree(d_poly);
  return rc;
}

int gahhg_query_xyz(gahhg_ctx* c, const double* xyz, int64_t n, int32_t* labels) {
  RCK(ctx_ok(c));
  if (!c->built) return set_err(c, GAHHG_E_STATE, "query before build");
  if (n < 0 || (n > 0 && (!xyz || !labels))) return set_err(c, GAHHG_E_ARG, "query: bad arguments");
  if (n == 0) return GAHHG_OK;
  RCK(select_device(c));
  double* d_xyz = nullptr;
  int32_t* d_lab = nullptr;
  RCK(dalloc(c, (void**)&d_xyz, n * 24));
  int rc = dalloc(c, (void**)&d_lab, n * 4);
  if (!rc) rc = hip_check(c, hipMemcpy(d_xyz, xyz, n * 24, hipMemcpyHostToDevice), "query H2D");
  if (!rc) rc = gahhg_query_xyz_device(c, d_xyz, n, d_lab);
  if (!rc) rc = hip_check(c, hipMemcpy(labels, d_lab, n * 4, hipMemcpyDeviceToHost), "labels D2H");
  dfree(d_xyz);
  dfree(d_lab);
  return rc;
}

int gahhg_query_lonlat_to_xyz(const double* lon, const double* lat, int64_t n, double* xyz) {
  if (n < 0 || (n > 0 && (!lon || !lat || !xyz))) return GAHHG_E_ARG;
  query_lonlat_to_xyz(lon, lat, n, xyz);
  return GAHHG_OK;
}

int gahhg_query_lonlat(gahhg_ctx* c, const double* lon, const double* lat, int64_t n, int32_t* labels) {
  RCK(ctx_ok(c));
  if (n < 0 || (n > 0 && (!lon || !lat || !labels))) return set_err(c, GAHHG_E_ARG, "query: bad arguments");
  std::vector<double> xyz((size_t)(3 * n));
  query_lonlat_to_xyz(lon, lat, n, xyz.data());
  return gahhg_query_xyz(c, xyz.data(), n, labels);
}

int gahhg_read_query_csv(const char* path, double* lon, double* lat, int64_t cap, int64_t* n) {
  if (!path || !n) return GAHHG_E_ARG;
  std::vector<double> a, b;
  std::string err;
  int rc = read_query_csv(path, &a, &b, &err);
  if (rc) return rc;
  *n = (int64_t)a.size();
  if (lon && lat) {
    if (cap < (int64_t)a.size()) return GAHHG_E_ARG;
    std::copy(a.begin(), a.end(), lon);
    std::copy(b.begin(), b.end(), lat);
  }
  return GAHHG_OK;
}

int gahhg_read_lune_points(const char* path, double* lon, double* lat, int64_t cap, int64_t* n) {
  if (!path || !n) return GAHHG_E_ARG;
  std::vector<double> a, b;
  std::string err;
  int rc = read_lune_points(path, &a, &b, &err);
  if (rc) return rc;
  *n = (int64_t)a.size();
  if (lon && lat) {
    if (cap < (int64_t)a.size()) return GAHHG_E_ARG;
    std::copy(a.begin(), a.end(), lon);
    std::copy(b.begin(), b.end(), lat);
  }
  return GAHHG_OK;
}

int gahhg_read_lune_polygon(const char* path, double* lon, double* lat, int64_t cap, int64_t* n,
                            double* ref_lonlat, int32_t* has_ref) {
  if (!path || !n) return GAHHG_E_ARG;
  PolygonText t;
  double ref[2] = {0.0, 0.0};
  bool hr = false;
  std::string err;
  int rc = parse_polygon_lune(path, &t, ref, &hr, &err);
  if (rc) return rc;
  const int64_t m = (int64_t)t.lo.size() + 1;  // with the closing vertex
  *n = m;
  if (lon && lat) {
    if (cap < m) return GAHHG_E_ARG;
    std::copy(t.lo.begin(), t.lo.end(), lon);
    std::copy(t.la.begin(), t.la.end(), lat);
    lon[m - 1] = t.lo[0];
    lat[m - 1] = t.la[0];
  }
  if (ref_lonlat) {
    ref_lonlat[0] = ref[0];
    ref_lonlat[1] = ref[1];
  }
  if (has_ref) *has_ref = hr ? 1 : 0;
  return GAHHG_OK;
}

int gahhg_get_stats(gahhg_ctx* c, gahhg_stats* out) {
  RCK(ctx_ok(c));
  if (!out) return GAHHG_E_ARG;
  *out = c->st;
  int64_t bytes = 0;
  bytes += c->g.vcap * 24 + c->g.bcap * (6 * 4 + 24) + c->g.ccap * 7 * 4 + c->g.rcap * 4;
  bytes += c->q.bytes;
  bytes += c->n_edges * (24 + 8 + 72 + 12);
  if (c->qsort.buf) bytes += sort_buffer_bytes(c->qsort.cap);  // grouping buffers (grow-only)
  bytes += c->scap * 4;                                         // scan partial sums (grow-only)
  out->device_bytes = bytes;
  return GAHHG_OK;
}

}  // extern "C"

namespace {

// The report block of GPISP::exportStat2TXT (source/GPISPAHHG.cu:2115-2161).
int write_report_block(gahhg_ctx* c, const char* path, const char* poly_name, const char* query_name,
                       int64_t n_vertices, const gahhg_stats& st) {
  FILE* fp = path ? fopen(path, "a") : stdout;
  if (!fp) return set_err(c, GAHHG_E_IO, "Can not open file %s", path);
  time_t t = time(nullptr);
  struct tm lt;
  localtime_r(&t, &lt);
  fprintf(fp, "\n----------------------------------------%d/%d/%d %d:%d:%d-----------------------------------------\n",
          lt.tm_year + 1900, lt.tm_mon + 1, lt.tm_mday, lt.tm_hour, lt.tm_min, lt.tm_sec);
  fprintf(fp, "\nPolygon file: %s\n", poly_name ? poly_name : "");
  fprintf(fp, "Number of polygon vertice: %lld\n", (long long)n_vertices);
  fprintf(fp, "Query point file: %s\n", query_name ? query_name : "");
  fprintf(fp, "Number of query points: %lld\n", (long long)st.n_queries);
  fprintf(fp, "Depth of GAHHG: %d\n", st.depth);
  fprintf(fp, "Time of preprocessing: %f seconds (creat GAHHG: %f seconds, classify GAHHG: %f seconds)\n",
          st.pre_ms * 1e-3, st.create_ms * 1e-3, st.classify_ms * 1e-3);
  fprintf(fp, "Storage overhead of GAHHG on CPU: %lld Bytes\n", (long long)st.cpu_space_bytes);
  fprintf(fp, "Storage overhead of GAHHG on GPU: %lld KB\n", (long long)(st.gpu_space_bytes / 1000));
  fprintf(fp, "Total test time for all query points: %f seconds\n", st.query_ms * 1e-3);
  fprintf(fp, "Average test time per query point: %.5e seconds\n", st.query_unit_s);
  fprintf(fp, "Test result: inside %lld, outside %lld\n", (long long)st.inside, (long long)st.outside);
  fprintf(fp, "\n---------------------------------------------------------------------------------------------------\n");
  if (path) fclose(fp);
  return GAHHG_OK;
}

}  // namespace

extern "C" {

int gahhg_write_report(gahhg_ctx* c, const char* path, const char* poly_name, const char* query_name) {
  RCK(ctx_ok(c));
  return write_report_block(c, path, poly_name, query_name, c->n_edges, c->st);
}

int gahhg_write_polygon_report(gahhg_ctx* c, int32_t p, const char* path, const char* poly_name,
                               const char* query_name, int64_t n_queries, int64_t inside, double query_ms,
                               double pre_ms) {
  RCK(ctx_ok(c));
  if (!c->built) return set_err(c, GAHHG_E_STATE, "report before build");
  if (p < 0 || p >= (int32_t)c->polys.size() || n_queries < 0 || inside < 0 || inside > n_queries)
    return set_err(c, GAHHG_E_ARG, "polygon report: bad arguments");
  gahhg_stats st;
  poly_stats(c, p, &st);
  st.n_queries = n_queries;
  st.inside = inside;
  st.outside = n_queries - inside;
  st.query_ms = query_ms;
  st.query_unit_s = n_queries > 0 ? query_ms * 1e-3 / (double)n_queries : 0.0;
  st.pre_ms = pre_ms;
  st.create_ms = st.classify_ms = 0.0;
  if (c->polys.size() == 1) {
    st.create_ms = c->st.create_ms;
    st.classify_ms = c->st.classify_ms;
  }
  return write_report_block(c, path, poly_name, query_name, c->polys[p].N, st);
}

int gahhg_write_result_csv(const char* path, const int32_t* labels, int64_t n) {
  if (!path || (n > 0 && !labels)) return GAHHG_E_ARG;
  FILE* fp = fopen(path, "w");
  if (!fp) return GAHHG_E_IO;
  for (int64_t i = 0; i < n; i++) fputs(labels[i] == kIn ? "1\n" : "0\n", fp);
  fclose(fp);
  return GAHHG_OK;
}

int gahhg_device_alloc(gahhg_ctx* c, int64_t bytes, void** d_ptr) {
  RCK(ctx_ok(c));
  if (!d_ptr || bytes < 0) return GAHHG_E_ARG;
  RCK(select_device(c));
  RCK(dalloc(c, d_ptr, bytes));
  c->user_allocs.push_back(*d_ptr);
  return GAHHG_OK;
}

int gahhg_device_free(gahhg_ctx* c, void* p) {
  RCK(ctx_ok(c));
  auto it = std::find(c->user_allocs.begin(), c->user_allocs.end(), p);
  if (it == c->user_allocs.end()) return set_err(c, GAHHG_E_ARG, "device_free: unknown pointer");
  c->user_allocs.erase(it);
  RCK(select_device(c));
  (void)hipStreamSynchronize(c->stream);
  dfree(p);
  return GAHHG_OK;
}

int gahhg_memcpy_h2d(gahhg_ctx* c, void* d, const void* h, int64_t bytes) {
  RCK(ctx_ok(c));
  RCK(select_device(c));
  return hip_check(c, hipMemcpy(d, h, (size_t)bytes, hipMemcpyHostToDevice), "memcpy H2D");
}

int gahhg_memcpy_d2h(gahhg_ctx* c, void* h, const void* d, int64_t bytes) {
  RCK(ctx_ok(c));
  RCK(select_device(c));
  RCK(hip_check(c, hipStreamSynchronize(c->stream), "sync"));
  return hip_check(c, hipMemcpy(h, d, (size_t)bytes, hipMemcpyDeviceToHost), "memcpy D2H");
}

int gahhg_grid_counts(gahhg_ctx* c, int64_t* counts) {
  return gahhg_polygon_grid_counts(c, 0, counts);
}

int gahhg_export_grid(gahhg_ctx* c, double* v, int32_t* cells, int32_t* borders, double* normals, int32_t* refs,
                      int32_t* levels) {
  return gahhg_export_polygon_grid(c, 0, v, cells, borders, normals, refs, levels);
}

int gahhg_num_polygons(gahhg_ctx* c, int32_t* n) {
  RCK(ctx_ok(c));
  if (!n) return GAHHG_E_ARG;
  *n = (int32_t)c->polys.size();
  return GAHHG_OK;
}

int gahhg_polygon_grid_counts(gahhg_ctx* c, int32_t p, int64_t* counts) {
  RCK(ctx_ok(c));
  if (!counts) return GAHHG_E_ARG;
  if (!c->built) return set_err(c, GAHHG_E_STATE, "no grid");
  if (p < 0 || p >= (int32_t)c->polys.size()) return set_err(c, GAHHG_E_ARG, "bad polygon index");
  return export_local(c, p, counts, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
}

int gahhg_export_polygon_grid(gahhg_ctx* c, int32_t p, double* v, int32_t* cells, int32_t* borders,
                              double* normals, int32_t* refs, int32_t* levels) {
  RCK(ctx_ok(c));
  if (!c->built) return set_err(c, GAHHG_E_STATE, "no grid");
  if (p < 0 || p >= (int32_t)c->polys.size()) return set_err(c, GAHHG_E_ARG, "bad polygon index");
  RCK(select_device(c));
  return export_local(c, p, nullptr, v, cells, borders, normals, refs, levels);
}

int gahhg_get_polygon_stats(gahhg_ctx* c, int32_t p, gahhg_stats* out) {
  RCK(ctx_ok(c));
  if (!out) return GAHHG_E_ARG;
  if (!c->built) return set_err(c, GAHHG_E_STATE, "no grid");
  if (p < 0 || p >= (int32_t)c->polys.size()) return set_err(c, GAHHG_E_ARG, "bad polygon index");
  poly_stats(c, p, out);
  return GAHHG_OK;
}

int gahhg_sort_pairs(gahhg_ctx* c, uint32_t* keys, uint32_t* vals, int64_t n, int64_t n_bound, int32_t end_bit) {
  RCK(ctx_ok(c));
  if (n < 0 || n_bound < n || n_bound > INT32_MAX || end_bit < 1 || end_bit > 32 || (n > 0 && (!keys || !vals)))
    return set_err(c, GAHHG_E_ARG, "sort_pairs: bad arguments");
  if (n_bound == 0) return GAHHG_OK;
  RCK(select_device(c));
  void* buf = nullptr;
  const int64_t hist = radix_hist_entries(n_bound);
  RCK(dalloc(c, &buf, 16 * n_bound + 4 * hist + 256));
  uint32_t* k0 = (uint32_t*)buf;
  uint32_t* k1 = k0 + n_bound;
  uint32_t* v0 = k1 + n_bound;
  uint32_t* v1 = v0 + n_bound;
  int32_t* h = (int32_t*)(v1 + n_bound);
  int32_t* d_m = h + hist;
  uint32_t *kr = nullptr, *vr = nullptr;
  int rc = hip_check(c, hipMemcpy(k0, keys, 4 * n, hipMemcpyHostToDevice), "sort H2D");
  if (!rc) rc = hip_check(c, hipMemcpy(v0, vals, 4 * n, hipMemcpyHostToDevice), "sort H2D");
  if (!rc) rc = fill_i32(c, d_m, 1, (int32_t)n);
  if (!rc) rc = radix_sort_pairs(c, k0, v0, k1, v1, d_m, n_bound, end_bit, h, &kr, &vr);
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "sort sync");
  if (!rc) rc = hip_check(c, hipMemcpy(keys, kr, 4 * n, hipMemcpyDeviceToHost), "sort D2H");
  if (!rc) rc = hip_check(c, hipMemcpy(vals, vr, 4 * n, hipMemcpyDeviceToHost), "sort D2H");
  dfree(buf);
  return rc;
}

int gahhg_kat_device(gahhg_ctx* c, int32_t op, const double* in12, int64_t n, double* out3) {
  RCK(ctx_ok(c));
  if (op < 0 || op > 3 || n < 0 || (n > 0 && (!in12 || !out3))) return GAHHG_E_ARG;
  if (n == 0) return GAHHG_OK;
  RCK(select_device(c));
  double *di = nullptr, *dout = nullptr;
  RCK(dalloc(c, (void**)&di, n * 96));
  int rc = dalloc(c, (void**)&dout, n * 24);
  if (!rc) rc = hip_check(c, hipMemcpy(di, in12, n * 96, hipMemcpyHostToDevice), "kat H2D");
  if (!rc) rc = run_kat(c, op, di, n, dout);
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "kat sync");
  if (!rc) rc = hip_check(c, hipMemcpy(out3, dout, n * 24, hipMemcpyDeviceToHost), "kat D2H");
  dfree(di);
  dfree(dout);
  return rc;
}

int gahhg_kat_cp_prop(gahhg_ctx* c, const int32_t* refs, int64_t n_refs, const int32_t* rs_rn,
                      const double* centers, int64_t n, int32_t group, int32_t* out3, double* mindist) {
  RCK(ctx_ok(c));
  if (n < 0 || n > INT32_MAX / 3 || n_refs < 0 || (n > 0 && (!refs || !rs_rn || !centers || !out3 || !mindist)))
    return set_err(c, GAHHG_E_ARG, "kat cp_prop: bad arguments");
  if (group < 1 || group > 64 || (group & (group - 1)) != 0)
    return set_err(c, GAHHG_E_ARG, "kat cp_prop: group must be a power of two up to 64");
  // the edge records (end points, plane normal) are made by the build
  if (!c->built || c->n_edges <= 0 || !c->d_edge) return set_err(c, GAHHG_E_STATE, "kat cp_prop before build");
  for (int64_t k = 0; k < n; k++) {
    const int64_t rs = rs_rn[2 * k], rn = rs_rn[2 * k + 1];
    if (rs < 0 || rn < 1 || rs + rn > n_refs) return set_err(c, GAHHG_E_ARG, "kat cp_prop: case %lld refs out of range", (long long)k);
  }
  for (int64_t i = 0; i < n_refs; i++)
    if (refs[i] < 0 || refs[i] >= c->n_edges) return set_err(c, GAHHG_E_ARG, "kat cp_prop: ref %lld is no edge", (long long)i);
  if (n == 0) return GAHHG_OK;
  RCK(select_device(c));
  int32_t *dr = nullptr, *drn = nullptr, *dout = nullptr;
  double *dc = nullptr, *dm = nullptr;
  int rc = dalloc(c, (void**)&dr, std::max<int64_t>(n_refs, 1) * 4);
  if (!rc) rc = dalloc(c, (void**)&drn, n * 8);
  if (!rc) rc = dalloc(c, (void**)&dc, n * 24);
  if (!rc) rc = dalloc(c, (void**)&dout, n * 12);
  if (!rc) rc = dalloc(c, (void**)&dm, n * 8);
  if (!rc) rc = hip_check(c, hipMemcpy(dr, refs, n_refs * 4, hipMemcpyHostToDevice), "kat H2D");
  if (!rc) rc = hip_check(c, hipMemcpy(drn, rs_rn, n * 8, hipMemcpyHostToDevice), "kat H2D");
  if (!rc) rc = hip_check(c, hipMemcpy(dc, centers, n * 24, hipMemcpyHostToDevice), "kat H2D");
  if (!rc) rc = run_kat_cp_prop(c, dr, drn, dc, n, group, dout, dm);
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "kat sync");
  if (!rc) rc = hip_check(c, hipMemcpy(out3, dout, n * 12, hipMemcpyDeviceToHost), "kat D2H");
  if (!rc) rc = hip_check(c, hipMemcpy(mindist, dm, n * 8, hipMemcpyDeviceToHost), "kat D2H");
  dfree(dr);
  dfree(drn);
  dfree(dc);
  dfree(dout);
  dfree(dm);
  return rc;
}

}  // extern "C"
