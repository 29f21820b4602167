// gahhg_internal.h -- context, device grid layout and cross-file entry points
// of the gfx950 GAHHG library.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/gahhg.h"
#include "gahhg_geom.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else
typedef struct ihipStream_t* hipStream_t;
typedef struct ihipEvent_t* hipEvent_t;
typedef struct ihipGraph* hipGraph_t;
typedef struct hipGraphExec* hipGraphExec_t;
#endif

namespace gahhg {

constexpr int kMaxLevels = 32;
constexpr int kGrid0Cells = 42;
constexpr int kGrid0Verts = 80;
constexpr int kGrid0Borders = 240;

// Per-level index ranges; GAHHGGrid, source/GPISPAHHG.cuh:12-20.
struct LevelDesc {
  int32_t vstart, vnum, bstart, bnum, cstart, cnum, rnum;
};

// One polygon of a batch, shared by host and device.  Cell, border and
// vertex indices are global (all polygons share the arrays); each polygon
// owns one contiguous segment per level, so its local numbering -- the
// single-polygon numbering of the reference -- is the global index minus
// the segment start plus the local start of that level.
struct PolyDev {
  int32_t e0, N;        // edges [e0, e0 + N)
  int32_t cell0;        // first level-0 cell (42 per polygon)
  int32_t refining;     // 1 while this polygon is refined at the current level
  int32_t L;            // bottom level once refinement stopped
  int32_t use_prev;     // dispatch locates edges from the level above the bottom
  int32_t roff, R;      // refs of the bottom cells
  int32_t bcs, bcn;     // bottom-level cell segment
  int32_t cs, cn;       // current level: cell segment
  int32_t bs, bn;       // current level: border segment
  int32_t lbs;          // current level: local start of its borders
  int32_t pbs, plbs;    // previous level: global / local border start
  int32_t U, nv, nbc, nbb;  // refined cells, their corners, border children, their borders
  int32_t nvb, nbb0, ncb;   // next level: first vertex, border, cell
  int32_t tbl;          // first entry of this polygon's start-cell table, -1 none
  int32_t tbins;        // bins per cube-face side of that table (0 = none)
  int32_t trel;         // 1: the table is the shared level-0 lookup (entries + cell0)
  int32_t depth;        // fixed depth for this polygon (0 = density rule)
  int32_t tsub;         // sub-table side S of this polygon's start table (1 = none)
  int32_t qg;           // lanes per edge-test group of this polygon's queries
};

// Adds v to counter[key] for every active lane with one atomic per distinct
// key of the wave (lanes sharing a key are summed first).  Keys of a
// contiguous range of cells or edges are mostly equal, so this removes the
// contention a plain per-lane atomic has on the few counters involved.
#ifdef __HIPCC__
template <typename T>
__device__ __forceinline__ void wave_add(T* counter, int key, T v) {
  unsigned long long todo = __ballot(1);
  const int lane = threadIdx.x & 63;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int k = __shfl(key, leader, 64);
    const unsigned long long same = __ballot(key == k) & todo;
    T s = (same >> lane) & 1ull ? v : (T)0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == leader && k >= 0) atomicAdd(counter + k, s);
    todo &= ~same;
  }
}
#endif

#ifdef __HIPCC__
// MI355X dispatches the workgroups of a launch to its 8 XCDs round-robin
// (consecutive blocks go to consecutive XCDs), and each XCD has its own L2.
// With chunk K > 0 this returns a logical block index under which each XCD
// runs K consecutive logical blocks at a time: within every run of 8K
// blocks, XCD j (the blocks b with b % 8 == j) takes logical blocks
// [j*K, j*K + K); blocks past the last full run keep their index (a
// bijection).  Neighbouring work then shares one L2, while the runs still
// interleave over the launch so that no XCD gets a whole expensive region.
__device__ __forceinline__ int64_t xcd_block(int K) {
  const int64_t b = blockIdx.x;
  if (K <= 0) return b;
  const int64_t super = 8 * (int64_t)K;
  const int64_t full = ((int64_t)gridDim.x / super) * super;
  if (b >= full) return b;
  const int64_t s = b / super, w = b - s * super;
  return s * super + (w & 7) * K + (w >> 3);
}
#endif

// Fixed-point scale of the per top-level cell edge sums (2^-40 rad).
constexpr double kFix = 1099511627776.0;

// Learned refinement model: parameter layout of the 12-64-64-1 perceptron
// (W1 rows padded to 16 inputs).
constexpr int kMlpIn = 12;
constexpr int kMlpW1 = 0, kMlpB1 = 1024, kMlpW2 = 1088, kMlpB2 = 5184, kMlpW3 = 5248, kMlpB3 = 5312;
constexpr int kMlpParams = 5313;

// Start-table entries at or below -kSubCode refer to a sub-table at offset
// -entry - kSubCode of the sub-table array.
constexpr int32_t kSubCode = 16;

// First global cell of every level (cstart[n] = total cells), passed to
// kernels by value to find a cell's level.
struct LevelTab {
  int32_t n;
  int32_t cstart[kMaxLevels + 1];
};

// Device-side counters and the level table, read back by one copy per level.
struct DevState {
  int32_t U;        // crossed cells at the level being refined
  int32_t nv;       // sum of bnum over them (= new vertices = centre borders)
  int32_t e0;       // borders of border children
  int32_t c0;       // border children
  int32_t R;        // bottom-level refs
  int32_t changed;  // classification round changed something
  int32_t nrec;     // records of the query edge stream
  unsigned long long steps_mark, steps_dispatch, partner_missing, cap_hits, unlocated, bad_index;
  unsigned long long zero_ref_edges, undefined_leaves;
  unsigned long long q_in, q_undef;  // label census of the last query
  int32_t bad_poly;                  // a device polygon id was out of range
  int32_t bad_nodes;                 // node records of cells with more than 6 borders
  LevelDesc lv[kMaxLevels];
  // device-driven build (every per-level decision taken on the device):
  int32_t run;        // 1 while the level loop goes on
  int32_t abort;      // the enqueued work stops short: levels left, or an arena array too small
  int32_t overflow;   // an arena array is too small; need_* hold what it must hold
  int32_t limit;      // a count exceeds int32 indexing
  int32_t gnum;       // levels created
  int32_t maxL;       // deepest bottom level
  int32_t tot_c32;    // total cells (the dispatch's scans read it as their length)
  int32_t R_ok;       // R once it fits the ref arrays, else 0 (the length the ref passes read)
  uint32_t rec_levels;  // levels whose cells have walk records (bit l)
  int64_t tot_v, tot_b, tot_c;
  int64_t need_v, need_b, need_c, need_r;
  double extra_next;  // expected walk steps of the next level beyond each edge's first cell
  unsigned long long steps_before;  // steps_mark before this level's mark walk
  unsigned long long und_first;     // leaves unlabelled after each level's first setIECell round
  int32_t ie_n;       // leaves of the level being classified left unlabelled by that round
  int32_t cls_G[kMaxLevels];        // lanes per cell of each level's getCPProp / parity groups
  int32_t nwalk;      // edges the dispatch's write walk walks (the others are copied from walk slots)
};

// Structure-of-arrays hierarchy in HBM.  Vertices and border normals are
// stored as packed xyz triples (24 B) because every consumer gathers whole
// points; topology fields are separate int32 arrays.
struct DeviceGrid {
  int64_t vcap = 0, bcap = 0, ccap = 0;
  double* v = nullptr;  // 3 per vertex
  int32_t* b_parent = nullptr;
  int32_t* b_partner = nullptr;
  int32_t* b_v0 = nullptr;
  int32_t* b_v1 = nullptr;
  int32_t* b_subcell = nullptr;
  int32_t* b_id = nullptr;
  double* b_n = nullptr;  // 3 per border
  int32_t* c_bstart = nullptr;
  int32_t* c_bnum = nullptr;
  int32_t* c_rstart = nullptr;
  int32_t* c_rnum = nullptr;
  int32_t* c_upcell = nullptr;
  int32_t* c_subcell = nullptr;
  int32_t* c_type = nullptr;
  int32_t* c_poly = nullptr;  // polygon of each cell
  int32_t* c_root = nullptr;  // level-0 ancestor (0..41) of each cell
  int64_t rcap = 0;
  int32_t* refs = nullptr;
};

// What one step of the build's edge walk reads about a cell, in one record:
// its border range and, per border, the plane normal, the first end vertex,
// the partner border and the partner's cell -- the same doubles as the grid
// arrays, so the walk's results are unchanged.  A cell's borders form a
// closed loop (border k ends where border k + 1 starts), so border k's second
// vertex is border k + 1's first; a cell whose loop does not close gets
// kWalkRecOpen in its bnum word and is walked on the grid arrays.  A step
// then costs one record fetch instead of a chain of dependent gathers
// through the border and vertex arrays.
struct WalkBorder {
  double n[3], v0[3];
  int32_t partner, pparent;
};
struct alignas(16) WalkRec {
  int32_t bstart, bnum;
  WalkBorder b[6];
  int32_t pad[2];
};
static_assert(sizeof(WalkRec) == 352, "walk record layout");
constexpr int32_t kWalkRecOpen = 0x100;

// Raw pointer bundle passed to kernels by value.
struct GridView {
  const double* v;
  int32_t* b_parent;
  int32_t* b_partner;
  int32_t* b_v0;
  int32_t* b_v1;
  int32_t* b_subcell;
  int32_t* b_id;
  double* b_n;
  int32_t* c_bstart;
  int32_t* c_bnum;
  int32_t* c_rstart;
  int32_t* c_rnum;
  int32_t* c_upcell;
  int32_t* c_subcell;
  int32_t* c_type;
  int32_t* c_poly;
  int32_t* c_root;
  int32_t* refs;
};

#ifdef __HIPCC__
// The 42 level-0 cells (at most 6 borders each: 12 pentagons, 30 hexagons)
// staged in LDS, for kernels that locate many points among them.  Level-0
// geometry is the same for every polygon; the copy is polygon 0's.
struct Level0Lds {
  double n[kGrid0Cells][6][3];
  int bn[kGrid0Cells];
};

// Block-wide load (every thread of the block calls it).
__device__ __forceinline__ void level0_load(const GridView& g, int cell0, Level0Lds* s) {
  for (int t = threadIdx.x; t < kGrid0Cells * 6; t += blockDim.x) {
    const int i = t / 6, k = t % 6;
    const int bs = g.c_bstart[cell0 + i], bn = min(g.c_bnum[cell0 + i], 6);
    if (k == 0) s->bn[i] = bn;
    if (k < bn) {
      const V3 v = load3(g.b_n + 3 * (int64_t)(bs + k));
      s->n[i][k][0] = v.x, s->n[i][k][1] = v.y, s->n[i][k][2] = v.z;
    }
  }
  __syncthreads();
}

// First level-0 cell (index order) whose border planes all pass the
// reference's inside test (n.q >= -eps in border order), or -1.
__device__ __forceinline__ int level0_cell(const Level0Lds* s, V3 q) {
  for (int i = 0; i < kGrid0Cells; i++) {
    bool in = true;
    for (int k = 0; k < s->bn[i] && in; k++)
      in = !lt0(dot(v3(s->n[i][k][0], s->n[i][k][1], s->n[i][k][2]), q));
    if (in) return i;
  }
  return -1;
}
#endif


// Read-only query accelerators built after classification.
//  table: one entry per cube-map bin (6 faces x bins x bins).  For every
//    point inside the bin's bounding cap, the reference's first-match
//    descent provably passes through the same cell; the entry holds that
//    cell (>= 0), -1 when even the level-0 choice is not certain, or
//    -(2 + label) when the cell is an empty leaf (label known outright).
//  qc: centre of every bottom-level cell (mean of its border start
//    vertices, normalised), the value kernel_PISP recomputes per query.
//  Edge stream of every bottom cell with refs: its refs (ascending edge id)
//    as polylines.  Consecutive refs e, e' with end(e) == start(e') share a
//    record, so a chain of k edges is k + 1 records.  Record r holds a
//    vertex (sv, 3 doubles); when bit 2 of sf[r] is set, the edge from
//    record r to record r+1 exists, sn[r] is its plane normal and bits 0/1
//    give the side of the leaf centre (ne.c < -eps / ne.c > eps).  hdr holds
//    (first record, record count) per bottom cell.
// A device allocation kept across builds: grown (contents dropped) only when
// a build needs more.
struct DevBuf {
  void* p = nullptr;
  int64_t cap = 0;  // bytes
};

struct QueryLayout {
  int32_t bins = 0;
  int32_t* table = nullptr;
  int32_t* sub = nullptr;      // sub-tables of the bins whose entry is not final
  int64_t nsub = 0;            // sub-table entries
  int32_t* tmp = nullptr;      // table-build scratch (grow-only)
  int64_t tmpcap = 0;
  double* qc = nullptr;
  int32_t* hdr = nullptr;
  double* sv = nullptr;
  double* sn = nullptr;
  uint8_t* sf = nullptr;
  int64_t nrec = 0;
  int32_t* ib = nullptr;   // per cell: angular index offset in idx, -1 = none
  int32_t* idx = nullptr;  // angular indexes of the leaves' edge streams
  int64_t nidx = 0;
  int64_t leaves = 0;  // bottom cells with refs
  int32_t* pcnt = nullptr;  // per polygon: edge-stream records, leaves with refs
  int64_t bytes = 0;
  int64_t bytes_nodes = 0;  // node records (counted into bytes once the leaf records are sized)
  // allocations behind the pointers above (a pointer is null when its part
  // of the layout is not built)
  DevBuf b_table, b_sub, b_tmp, b_pcnt, b_qc, b_hdr, b_ib, b_idx, b_sv, b_sn, b_sf, b_iscr, b_tstart, b_l0, b_nodes;
  int32_t* tstart = nullptr;  // per polygon: first table entry (np + 1 values, ascending)
  // resolutions whose level-0 table part b_l0 holds (bit k: 16 << k bins per
  // side); the level-0 grid is the same for every polygon and build, so a
  // part is computed once and kept until b_l0 is reallocated or freed
  uint32_t l0_have = 0;
  // per cell, kNodeInts int32 (192 B): the descent's view of the cell in one
  // record -- [0] centre child (-1 = leaf), [1] borders | border children
  // << 8 | type << 16 | (refs != 0) << 24, [2..7] border children in border
  // order (-1 skipped), [10..45] its border normals (6 x 3 doubles)
  int32_t* nodes = nullptr;
};
constexpr int kNodeInts = 48;

// Learned refinement depth: model, cost constants, hint sample and the
// per (polygon, top-level cell) device arrays of the build.
struct LearnState {
  bool on = false;
  std::vector<float> params;
  // expected points (total), query_scale, ns per edge, ns per new cell, ns
  // per level
  double cost[5] = {0, 1, 0, 0, 0};
  uint16_t* d_weights = nullptr;  // BF16 model weights
  double* d_hint_xyz = nullptr;
  int32_t* d_hint_poly = nullptr;
  int64_t hint_n = 0;
  double hint_weight = 1.0;
  unsigned long long* d_estat = nullptr;  // np x 42 x 3 edge statistics
  uint32_t* d_hist = nullptr;             // np x 42 hint points
  uint32_t* d_band = nullptr;             // np x 42 hint points in crossed cells
  int32_t* d_hloc = nullptr;              // per hint point: its cell at hloc_level (-1 = done)
  int32_t* d_hloc0 = nullptr;             // per hint point: its top-level cell
  int64_t hloc_cap = 0;
  int hloc_level = 0;
  // d_estat, and d_hist with d_hloc0, depend only on the polygons (and the
  // hint): kept across rebuilds until those change
  bool estat_valid = false, hist_valid = false;
  uint32_t* d_stripes = nullptr;          // kStripes x np x 42 partial counts
  float* d_w = nullptr;                   // np x 42 expected points
  int32_t* d_stat = nullptr;              // np x 42 x 2 cells, crossed cells
  float* d_x = nullptr;                   // np x 42 x 2 x 12 features
  float* d_y = nullptr;                   // np x 42 x 2 predictions
  int32_t* d_dec = nullptr;               // np decisions
  int64_t cap = 0;                        // polygons the arrays hold
  std::vector<int32_t> h_dec;
  std::vector<int32_t> rule_stop;  // level where the density rule stops, -1 = not yet
  // learned start-table depth (label-neutral): model, cost constants
  // (expected points without a hint, query scale, per candidate the table
  // build ns a_c + b_c * N of one polygon, then the ns per table bin in a
  // batch) and its per (polygon, top-level cell, candidate) arrays
  bool table_on = false;
  std::vector<float> tparams;
  double tcost[3 + 2 * 8] = {0};
  uint16_t* d_tweights = nullptr;
  float* d_tx = nullptr;  // np x 42 x kTableCand x 12 features
  float* d_ty = nullptr;  // np x 42 x kTableCand predictions
  int64_t tcap = 0;
};

// Start-table candidates of the learned table depth: bins per cube-face side
// and sub-table side.  A batch considers the first kTableCandBatch (no
// sub-tables, at most 128 bins); one polygon all of them.
constexpr int kTableCand = 8;
constexpr int kTableCandBatch = 5;
constexpr int kTableB[kTableCand] = {0, 16, 32, 64, 128, 256, 256, 256};
constexpr int kTableS[kTableCand] = {1, 1, 1, 1, 1, 1, 8, 16};

// Window of the learned depth decision (crossed cells / N below which a
// polygon always refines; levels past the density rule's stop it may add).
constexpr float kLearnFloor = 0.02f;
// Copies of the per (polygon, top-level cell) counters that the hint kernels
// add to (block b adds to copy b % kStripes): few blocks add to any one
// address, whose atomics serialise.
constexpr int kStripes = 64;
constexpr int kLearnAbove = 2;
// Predicted gain factor the model needs to depart from the density rule.
constexpr double kLearnMargin = 2.0;

// Grow-only buffers of the query sort: keys and values (two of each, n
// entries) followed by the sort's temporary storage.
struct QuerySort {
  void* buf = nullptr;
  int64_t cap = 0;
};

// A grouped query pass recorded as a HIP graph (query_graph = 1), replayed
// while the points, polygon ids, labels, count and every buffer the pass
// touches stay where they were at capture (epoch unchanged).
struct QueryGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* xyz = nullptr;
  const void* poly = nullptr;
  void* labels = nullptr;
  int64_t n = 0;
  uint64_t epoch = 0;
};

// Queries below this count are not sorted (query_sort = 1).
constexpr int64_t kSortMin = 1 << 20;

struct Polygon {
  int32_t N = 0;
  int32_t flips = 0;
  std::vector<double> pv;  // 3N
  std::vector<int32_t> pe; // 2N
};

// A ring of a parsed polygon file: unique vertices [start, start + n) of
// the file's lon/lat arrays (the closing duplicate is not stored).
struct Ring {
  int64_t start;
  int32_t n;
};

// The text of a polygon file, checked as the reference's loader checks it.
struct PolygonText {
  std::vector<double> lo, la;
  std::vector<Ring> rings;
};

// ---- host (gahhg_host.cpp) ----
int load_polygon_txt(const char* path, Polygon* out, std::string* err);
int parse_polygon_txt(const char* path, PolygonText* out, std::string* err);
// Polygon from parsed text; xyz (parallel to lo/la) or null = convert here.
int polygon_from_text(const PolygonText& t, const double* xyz, Polygon* out, std::string* err);
int polygon_from_lonlat(const double* lon, const double* lat, const int64_t* ring_off,
                        int32_t n_rings, Polygon* out, std::string* err);
// As polygon_from_lonlat with each vertex's point already converted (xyz
// parallel to lon/lat, closing vertices included; null = convert here).
int polygon_from_lonlat_xyz(const double* lon, const double* lat, const double* xyz, const int64_t* ring_off,
                            int32_t n_rings, Polygon* out, std::string* err);
// LLVertex::LL2XYZ of n points (correctly rounded sin and cos).
void lonlat_to_xyz(const double* lon, const double* lat, int64_t n, double* xyz);
void build_grid0(double* v80, int32_t* bi240x6, double* bn240x3, int32_t* ci42x7);
void query_lonlat_to_xyz(const double* lon, const double* lat, int64_t n, double* xyz);
int read_query_csv(const char* path, std::vector<double>* lon, std::vector<double>* lat,
                   std::string* err);
// Lune/Ketzner-format files ("colatitude,longitude" radians per line):
// points as lon/lat degrees in the reference's convention, and a polygon
// file's ring (a PolygonText of one ring) plus the contained point after it.
int read_lune_points(const char* path, std::vector<double>* lon, std::vector<double>* lat, std::string* err);
int parse_polygon_lune(const char* path, PolygonText* out, double* ref_lonlat, bool* has_ref, std::string* err);

}  // namespace gahhg

namespace gahhg {
// Device text loading (gahhg_load.hip): the file's bytes, line offsets,
// parsed output, listed-line indices, per-block counts, pinned staging.
struct TextLoad {
  uint8_t* text = nullptr;
  int64_t text_cap = 0;
  int64_t* starts = nullptr;
  int64_t starts_cap = 0;
  double* xyz = nullptr;
  int64_t xyz_cap = 0;
  int32_t* flags = nullptr;
  int64_t flags_cap = 0;
  int32_t* counts = nullptr;
  int64_t counts_cap = 0;
  int32_t* misc = nullptr;
  int64_t misc_cap = 0;
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
};
int load_query_csv_device(gahhg_ctx* c, const char* path, int64_t* n, gahhg_load_info* info);
int parse_query_csv_device(gahhg_ctx* c, const char* text, int64_t bytes, int64_t* n, gahhg_load_info* info);
int load_polygon_txt_device(gahhg_ctx* c, const char* path, Polygon* out, gahhg_load_info* info);
// LLVertex::LL2XYZ of n host points on the device (the same bits as
// lonlat_to_xyz), result in host xyz.
int lonlat_to_xyz_gpu(gahhg_ctx* c, const double* lon, const double* lat, int64_t n, double* xyz);
// Vertex counts from which the TXT loaders convert on the device.
constexpr int64_t kGpuConvertMin = 16384;
void free_text_load(gahhg_ctx* c);
// Device-driven build (build_hierarchy_dev): the level loop, the dispatch and
// the classification take every per-level decision on the device, so the
// whole build is one stream of launches with no host synchronisation in it
// and can be recorded into a HIP graph and replayed.
struct DevBuildState {
  // level-0 grid template (80 vertices, 240 borders, 42 cells) on the device
  double* g0_v = nullptr;
  double* g0_n = nullptr;
  int32_t* g0_b = nullptr;  // 6 per border
  int32_t* g0_c = nullptr;  // 7 per cell
  PolyDev* pd0 = nullptr;   // polygon records at the start of a build
  int32_t pd0_n = 0;
  LevelDesc* plv = nullptr;  // per polygon kMaxLevels level segments
  int32_t* plvn = nullptr;   // per polygon: levels created
  float* gamma = nullptr;    // per polygon density ratio
  int32_t* pmarks = nullptr; // per polygon levels marked
  int32_t* hist = nullptr;   // radix histograms of the ref sort
  int64_t hist_cap = 0;
  int32_t* csums = nullptr;  // chunk sums of the chunked scans
  int32_t* sscr = nullptr;   // the build's own scan scratch (queries regrow the context's)
  int64_t sscr_cap = 0;
  char* hbuf = nullptr;      // pinned: the build's results read back in one copy each, one wait
  int32_t hcap = 0;          // polygons hbuf holds
  std::vector<PolyDev> pd0_host;  // what pd0 holds
  int32_t pcap = 0;          // polygons the per-polygon arrays hold
  // arena the last build ran with (cells, borders, vertices, refs) and what
  // it used; the next build sizes from these
  int64_t used_v = 0, used_b = 0, used_c = 0, used_r = 0;
  int32_t used_levels = 0;
  int64_t shape_n = -1;      // edges of the polygon set those figures belong to
  int32_t shape_np = -1;
  // the recorded build and what it was recorded against
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uintptr_t> key;
  int32_t graph_levels = 0;
};

}  // namespace gahhg

struct gahhg_ctx {
  gahhg_params prm;
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

  std::vector<gahhg::Polygon> polys;  // a batch; a single polygon is a batch of one
  int64_t n_edges = 0;                // all polygons
  bool have_poly = false;
  double* d_pv = nullptr;             // vertices of all polygons
  int32_t* d_pe = nullptr;            // edge endpoints (global vertex indices)
  int32_t* d_epoly = nullptr;         // polygon of each edge
  double* d_edge = nullptr;           // 9 per edge: a, b, arc_normal(a, b)
  gahhg::PolyDev* d_pd = nullptr;
  std::vector<gahhg::PolyDev> pd;
  std::vector<std::vector<gahhg::LevelDesc>> plv;  // per polygon: global segments per level
  std::vector<float> pgamma;                        // per polygon density ratio
  std::vector<int32_t> depth_override;              // per polygon fixed depth (0 = rule)
  std::vector<int32_t> pmarks;                      // per polygon levels marked in the last build
  std::vector<int32_t> table_fixed;                 // per polygon start-table candidate (-1 = chosen)
  std::vector<int32_t> table_used;                  // per polygon (bins, sub side) of the last layout
  std::vector<gahhg::LevelDesc> glv;                // per level: global ranges over all polygons
  int64_t tot_v = 0, tot_b = 0, tot_c = 0, tot_r = 0;

  gahhg::DeviceGrid g;
  gahhg::LevelDesc lv[gahhg::kMaxLevels];
  int32_t gnum = 0;
  float gamma = 0.f;
  gahhg::DevState* d_st = nullptr;
  gahhg::DevState* h_st = nullptr;  // pinned mirror
  int32_t* d_loc_cur = nullptr;     // per edge: start cell at the last level
  int32_t* d_loc_prev = nullptr;    // per edge: start cell one level up

  // scratch
  int64_t fcap = 0;
  int32_t *d_f0 = nullptr, *d_f1 = nullptr, *d_o0 = nullptr, *d_o1 = nullptr;
  int32_t *d_f2 = nullptr, *d_f3 = nullptr, *d_o2 = nullptr, *d_o3 = nullptr;
  int64_t scap = 0;
  int32_t* d_scan = nullptr;
  int64_t ecap = 0;
  int32_t* d_ecnt = nullptr;  // per edge
  int64_t wslot_cap = 0;      // edges the walk slots hold
  int32_t* d_wslot = nullptr; // per edge: the first cells of its last mark walk (device-driven build)
  int64_t r2cap = 0;
  int32_t *d_rcell = nullptr, *d_rflag = nullptr, *d_redge = nullptr, *d_rscan = nullptr;
  gahhg::WalkRec* d_wrec = nullptr;  // walk record of every cell (build scratch)
  int64_t wrec_cap = 0;
  gahhg::DevBuildState dvb;          // device-driven build: templates, arena sizes, graph

  gahhg::QueryLayout q;
  gahhg::QuerySort qsort;
  gahhg::QueryGraph qgraph;
  uint64_t qepoch = 1;              // bumped when a buffer a query pass reads or writes moves
  unsigned long long* q_queue = nullptr;  // persistent query launch: per-XCD queue counters
  int32_t q_resident = 0;                 // its grid: blocks resident at once
  int32_t* h_cnt = nullptr;         // pinned: points the start table leaves open
  hipEvent_t ev_cnt = nullptr;      // h_cnt written
  gahhg::LearnState learn;
  gahhg::TextLoad tl;
  float* d_mlp_part = nullptr;       // per-wave partial gradients (grow-only)
  int64_t mlp_part_cap = 0;
  int64_t und_snap = 0, und_final = 0;
  bool built = false;
  bool build_replayed = false;  // the last build ran from the recorded graph
  int32_t build_attempts = 0;   // device-driven runs of the last build
  gahhg_stats st;
  double last_query_ms = 0.0;
  std::vector<void*> user_allocs;
};

namespace gahhg {

// ---- scan (gahhg_scan.hip) ----
// Exclusive prefix sum of in[0..n) into out; n is n_bound, or *d_n when
// d_n is non-null (elements past it read as 0 and are not written).  The
// total goes to *d_total when non-null.  inclusive != 0 gives the inclusive
// sum instead.
int scan_i32(gahhg_ctx* c, const int32_t* in, int32_t* out, int64_t n_bound, const int32_t* d_n,
             int32_t* d_total, int inclusive);
// Exclusive (inclusive != 0: inclusive) prefix sums of up to four int32
// columns of the same length *d_n, read on the device: a fixed grid whatever
// the length, two launches for all columns.  total[k] (non-null) receives
// column k's sum; sums holds kScanColsSums int32.
struct ScanCols {
  const int32_t* in[4];
  int32_t* out[4];
  int32_t* total[4];
};
constexpr int kScanColsSums = 4 * 256;
int scan_cols_dev(gahhg_ctx* c, const ScanCols& cols, int ncols, const int32_t* d_n, int inclusive, int32_t* sums);
// Grows the scan scratch for n_bound elements now (a later scan of at most
// that many then neither allocates nor synchronises).
int scan_reserve(gahhg_ctx* c, int64_t n_bound);
// int32 entries of scratch a scan of n_bound elements uses.
int64_t scan_scratch_entries(int64_t n_bound);
// scan_i32 over a caller's scratch of scan_scratch_entries(n_bound) int32
// (nothing allocated); scratch == nullptr uses the context's.
int scan_i32_with(gahhg_ctx* c, const int32_t* in, int32_t* out, int64_t n_bound, const int32_t* d_n,
                  int32_t* d_total, int inclusive, int32_t* scratch);

// ---- build / classify / query (gahhg_build.hip etc.) ----
int build_hierarchy(gahhg_ctx* c);
int classify_hierarchy(gahhg_ctx* c);
// The device-driven build and classification (every per-level decision on
// the device; replayed from a HIP graph when the polygon set, the arena and
// the parameters are those it was recorded with).  *replayed tells whether
// the graph ran.
int build_classify_dev(gahhg_ctx* c, bool* replayed, int32_t* attempts);
// Enqueues the device-driven classification of levels [0, levels) on the
// context's stream (no synchronisation).
int classify_dev_enqueue(gahhg_ctx* c, int levels);
void drop_build_graph(gahhg_ctx* c);
void free_dev_build(gahhg_ctx* c);
int pack_query_layout(gahhg_ctx* c);
int run_query(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels);
// GAHHG_E_ARG unless every d_poly[0..n) is a polygon of the batch (a
// device pass and one synchronisation).
int check_poly_ids(gahhg_ctx* c, const int32_t* d_poly, int64_t n);
// Every polygon against the same n points: writes n_poly * n points and
// polygon ids (point-major within each polygon) to d_xyz_out / d_poly_out.
int expand_pairs(gahhg_ctx* c, const double* d_xyz, int64_t n, double* d_xyz_out, int32_t* d_poly_out);
// Groups the points for the walk; *order = point indices in group order.
// Their number goes to *m, or, with host_count false, stays on the device
// at *d_count (*m = -1) so that nothing waits for it.
// The grouping's scratch: buf holds sort_buffer_bytes(cap) bytes (block
// counts, offsets and total, two key and two value arrays of cap entries,
// the radix histograms); scan = the scans' scratch, or nullptr for the
// context's.
struct SortScratch {
  void* buf = nullptr;
  int64_t cap = 0;
  int32_t* scan = nullptr;
};
int64_t sort_buffer_bytes(int64_t cap);
// ext: the scratch to use (nothing allocated; n <= ext->cap), or nullptr
// for the context's grow-only buffers.
int sort_queries(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels,
                 const int32_t** order, int64_t* m, const int32_t** d_count, bool host_count,
                 const SortScratch* ext = nullptr);
void drop_query_graph(gahhg_ctx* c);
// run_query with the grouping's scratch from ext (nothing allocated, no
// graph); ext == nullptr is run_query.
int run_query_with(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels,
                   const SortScratch* ext);
// The body of gahhg_query_batch_xyz_device: checks, the timed pass (scratch
// from ext when non-null), statistics and the label census.
int query_device(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels,
                 const SortScratch* ext);
// Resolves, once per device, the device address of the allocation-free
// query's scratch array (the device current); called by gahhg_create.
void noalloc_init(int device);

// ---- radix sort (gahhg_sort.hip) ----
// Stable sort of *d_m (<= n_max) pairs (k0, v0) by key bits [0, end_bit);
// (k1, v1) is the second buffer; *kres / *vres = where the result is
// (one of the two).  hist holds radix_hist_entries(n_max) int32.  The
// histogram scans use scan_scratch (scan_scratch_entries(256 * tiles)
// int32), or the context's scan scratch when it is nullptr.
int64_t radix_hist_entries(int64_t n_max);
int radix_sort_pairs(gahhg_ctx* c, uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, const int32_t* d_m,
                     int64_t n_max, int end_bit, int32_t* hist, uint32_t** kres, uint32_t** vres,
                     int32_t* scan_scratch = nullptr);
int run_kat(gahhg_ctx* c, int op, const double* d_in, int64_t n, double* d_out);
// getCPProp's choice at n centres over given ref lists (device arrays),
// with groups of G lanes.
int run_kat_cp_prop(gahhg_ctx* c, const int32_t* d_refs, const int32_t* d_rs_rn, const double* d_centers, int64_t n,
                    int G, int32_t* d_out3, double* d_mindist);
int run_query_trace(gahhg_ctx* c, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels,
                    int32_t* d_trace);
LevelTab level_tab(const gahhg_ctx* c);
int mlp_forward(gahhg_ctx* c, const float* d_x, int64_t rows, const uint16_t* d_w, float* d_y);
int mlp_backward(gahhg_ctx* c, const float* d_x, int64_t rows, const uint16_t* d_w, const float* d_dy,
                 float* d_grad);
int mlp_adam(gahhg_ctx* c, float* d_master, const float* d_grad, float* d_m, float* d_v, int step, float lr,
             uint16_t* d_w);
int build_trace(gahhg_ctx* c, int32_t* h_cells, uint64_t* h_edges);
int edge_stats(gahhg_ctx* c, unsigned long long* d_out);
int learn_prepare(gahhg_ctx* c);
// learn_prepare with q_total expected points when there is no hint.
int learn_prepare_with(gahhg_ctx* c, double q_total);
int attach_loaded_grid(gahhg_ctx* c, int64_t V, int64_t B, int64_t C, int64_t R);
int learn_level(gahhg_ctx* c, int level, int lo, int n, int32_t* h_dec);
int learn_features_after_build(gahhg_ctx* c, int level, float* h_x);
// Learned start-table depth: features of the built hierarchy for every
// candidate (h_x: np x 42 x kTableCand x 12, null = keep on the device) and,
// with a table model set, the chosen candidate per polygon into h_choice.
int table_features_after_build(gahhg_ctx* c, float* h_x);
int table_choose(gahhg_ctx* c, int32_t* h_choice);
void learn_free(gahhg_ctx* c);

GridView view(gahhg_ctx* c);
int set_err(gahhg_ctx* c, int code, const char* fmt, ...);
int hip_check(gahhg_ctx* c, int hip_status, const char* what);
int dalloc(gahhg_ctx* c, void** p, int64_t bytes);
void dfree(void* p);
// *active = a buffer of at least `bytes` from b (grown when too small).
int use_buf(gahhg_ctx* c, DevBuf& b, int64_t bytes, void** active);
void free_buf(DevBuf& b);

// Fills n int32 with v on the context stream (a kernel, so the call stays
// graph-capturable).
int fill_i32(gahhg_ctx* c, int32_t* p, int64_t n, int32_t v);
int copy_i32(gahhg_ctx* c, int32_t* dst, const int32_t* src, int64_t n);

// Each kernel file's code object, loaded onto the current device when a
// context is created (HIP loads a code object at its first use, which would
// otherwise fall into a context's first build).
void load_code_build();
void load_code_classify();
void load_code_query();
void load_code_scan();
void load_code_sort();
void load_code_learn();
void load_code_mlp();
void load_code_load();
void load_code_trace();

}  // namespace gahhg
