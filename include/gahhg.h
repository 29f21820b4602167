/*
 * gahhg.h -- C ABI of the MI355X (gfx950) GAHHG point-in-spherical-polygon
 * library.
 *
 * The reference (superredlark2/GAHHG) has no FFI: its boundary is the
 * in-process call sequence of source/GPISP.cu:47-67
 *   loadPolyFromTXT -> loadTPFromCSV -> createAHHGInit -> createAHHG(maxDepth)
 *   -> classifyAHHG -> test -> getStat -> exportStat2TXT -> exportResult2CSV
 * plus its file formats.  Each entry point below names the reference routine
 * it replaces.  Conventions: every function returns 0 or a negative
 * GAHHG_E_* code and never aborts; gahhg_last_error() gives the message.
 * Callers own host buffers; a context owns its device memory and one HIP
 * stream.  A context is not thread-safe; separate contexts are independent.
 * Labels use the reference's cell codes: 2 = IN, 1 = OUT, 0 = undefined or
 * unlocated (source/GPISPParam.h:9-11).
 */
#ifndef GAHHG_H
#define GAHHG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GAHHG_OK 0
#define GAHHG_E_ARG -1       /* bad argument                                */
#define GAHHG_E_IO -2        /* file could not be opened                    */
#define GAHHG_E_FORMAT -3    /* malformed polygon / query file              */
#define GAHHG_E_STATE -4     /* call out of order (e.g. query before build) */
#define GAHHG_E_HIP -5       /* HIP runtime error / no device               */
#define GAHHG_E_NOMEM -6     /* allocation failed                           */
#define GAHHG_E_LIMIT -7     /* sizes beyond int32 indexing                 */

#define GAHHG_LABEL_UNDEFINED 0
#define GAHHG_LABEL_OUT 1
#define GAHHG_LABEL_IN 2

typedef struct gahhg_ctx gahhg_ctx;

/* Parameters: GPISPParam, source/GPISPParam.h:49-78 (maxDepth=15,
 * autoDepth=true, blocksizes 256). */
typedef struct {
  int32_t max_depth;       /* levels, 1..15 (reference default 15)            */
  int32_t auto_depth;      /* 1: stop refining once crossed cells / N > ~0.09 */
  int32_t block_create;    /* threads per block of build kernels              */
  int32_t block_classify;  /* threads per block of classification kernels     */
  int32_t block_query;     /* threads per block of the query kernel (64..256,
                              a multiple of 64)                              */
  int32_t device;          /* HIP device ordinal                              */
  int32_t verbose;         /* 1: print per-level progress like the reference  */
  int32_t query_layout;    /* 0: table-accelerated query (default), 1: plain
                              descent from level 0 over the build arrays     */
  int32_t query_bins;      /* start-cell table bins per cube face (0 = none)  */
  int32_t query_group;     /* lanes per edge-test group of the accelerated
                              query: 1,2,4,...,64, or 0 = per polygon, from
                              its mean leaf edge-stream length               */
  int32_t query_sub;       /* largest sub-table side S of the start table
                              (power of two <= 256; 0 or 1 = none; default
                              8: fractal_1m's 100M points take 1.3 ms less
                              for 1 ms more build); single polygons only    */
  int32_t query_sub_budget;/* sub-table entries allowed per polygon edge
                              (default 4)                                   */
  int32_t query_sort;      /* accelerated query: answer what the start table
                              resolves, then group the other points by
                              polygon and position before the walk (0 never,
                              1 from 2^20 points, 2 always); labels are
                              unchanged                                     */
  int32_t query_sort_bits; /* grouping key: Z-order of the cube-map face
                              coordinates at up to this many bits per axis
                              (1..14, default 14; fewer when a batch's
                              polygon ids need the key's bits)              */
  int32_t query_index;     /* accelerated query: 1 (default) gives leaves
                              with many edge records an angular index, so a
                              point tests only the records that can cross
                              its arc; 0 tests every record; labels are
                              unchanged                                     */
  int32_t query_xcd;       /* accelerated query: K > 0 runs K consecutive
                              blocks of grouped points on one XCD at a time
                              (workgroups reach the 8 XCDs round-robin, each
                              with its own L2); 0 the plain block order;
                              default 32 (fractal_1m's walk 7.63 -> 7.31
                              ms); labels are unchanged                    */
  int32_t query_table_budget;/* batches: start-table entries allowed per
                              polygon edge (bins per face side a power of
                              two in [16, query_bins]; default 8: least
                              build + one query pass on 256 polygons with
                              64M points and on 624 polygons with 10k);
                              0 = every polygon gets query_bins; single
                              polygons always do                           */
  int32_t query_graph;     /* accelerated query: 1 records the pass (table
                              pass, grouping sort, walk) as a HIP graph on
                              first use and replays it while the points,
                              polygon ids, labels and count are the same
                              buffers and no build came between; the
                              points the table leaves open are counted on
                              the device only.  0 (default): launched
                              kernel by kernel.  Labels are unchanged      */
  int32_t build_device;    /* 1 (default): the build takes every per-level
                              decision on the device (density rule, next
                              level's layout, dispatch sizes, classification
                              fixpoint), one stream of launches with no host
                              synchronisation in it; 0: the host decides
                              between levels (also the path of the learned
                              hierarchy depth).  Grids are the same          */
  int32_t build_graph;     /* with build_device: 1 (default) records the
                              build as a HIP graph and replays it on later
                              builds of the same polygon set (same sizes,
                              parameters and buffers); a replay that runs
                              out of room or levels is redone eagerly and
                              recorded again.  Grids are the same           */
  int32_t query_persistent;/* accelerated query: 1 runs the walk as a
                              persistent, work-stealing launch (a resident
                              grid whose waves take 64-point tiles from
                              per-XCD queues, then from the other XCDs'),
                              0 (default) one thread per point.  Labels
                              are unchanged                                 */
} gahhg_params;

/* Statistics: GPISPStat fields (source/GPISPParam.h:80-100) as filled by
 * GPISP::getStat (source/GPISPAHHG.cu:2088-2113), plus diagnostics. */
typedef struct {
  double pre_ms;          /* create_ms + classify_ms                           */
  double create_ms;       /* createAHHG: level loop + edge dispatch (with
                             build_device also classifyAHHG: one stream)     */
  double classify_ms;     /* classifyAHHG and the query layout (with
                             build_device the query layout alone)            */
  double query_ms;        /* last query: kernel time only                      */
  double query_unit_s;    /* query seconds per point                           */
  int32_t depth;          /* levels built (gnum)                               */
  float gamma;            /* last crossed-cells / N ratio                      */
  int64_t n_edges;        /* polygon edges N                                   */
  int64_t n_queries;      /* points in the last query                          */
  int64_t vnum, bnum, cnum, rnum;
  int64_t vspace, bspace, cspace, rspace; /* 24 V, 48 B, 28 C, 4 R bytes      */
  int64_t gpu_space_bytes;/* reference space formula sum                       */
  int64_t cpu_space_bytes;/* reference host bookkeeping bytes                  */
  int64_t device_bytes;   /* device memory this context holds                  */
  int64_t inside;         /* labels == IN in the last query                    */
  int64_t outside;        /* n_queries - inside                                */
  int64_t unlocated;      /* points that reached no leaf                       */
  int32_t orientation_flipped; /* bit r: ring r reversed on load               */
  int32_t mark_passes;    /* levels at which the polygon's edges were marked
                             (a batch: the largest)                           */
  int64_t walk_steps_mark;
  int64_t walk_steps_dispatch;
  int64_t walk_partner_missing;
  int64_t walk_step_cap_hits;
  int64_t walk_unlocated;
  int64_t zero_ref_edges;
  int64_t undefined_leaves_snapshot;
  int64_t undefined_leaves_final;
  double query_kernel_ms; /* last query: the query kernel alone                */
  double query_sort_ms;   /* last query: grouping of the points before it      */
  double layout_ms;       /* last build: query layout (part of classify_ms)    */
  int32_t build_replayed; /* last build: 1 when it ran from the recorded graph */
  int32_t build_attempts; /* last build: device-driven runs it took (more than
                             one when the arena or the recorded levels were
                             too small and it ran again; 0: host loop)       */
} gahhg_stats;

/* Fills defaults (GPISPParam::GPISPParam, source/GPISPParam.h:66-77). */
int gahhg_params_default(gahhg_params* p);

/* Creates a context on p->device (NULL = defaults). */
int gahhg_create(gahhg_ctx** out, const gahhg_params* p);
void gahhg_destroy(gahhg_ctx* ctx);
const char* gahhg_last_error(const gahhg_ctx* ctx);

/* Replaces GPISPPolygon::loadPolyFromTXT (source/GPISPPolygon.cu:32-278):
 * polygon ring TXT file, outer ring first, orientation normalised. */
int gahhg_load_polygon_txt(gahhg_ctx* ctx, const char* path);

/* Same polygon from arrays: lon/lat degrees; ring_off[n_rings+1] vertex
 * offsets; each ring includes its closing duplicate vertex (as in the file);
 * ring 0 is the outer ring. */
int gahhg_set_polygon_lonlat(gahhg_ctx* ctx, const double* lon, const double* lat,
                             const int64_t* ring_off, int32_t n_rings);

/* Replaces GPISP::createAHHGInit + createAHHG(maxDepth) + classifyAHHG
 * (source/GPISPAHHG.cu:1947-2059): builds and classifies the hierarchy on
 * the GPU. */
int gahhg_build(gahhg_ctx* ctx);

/* Replaces GPISPTP::loadTPFromCSV (source/GPISPTP.cu:13-79) + GPISP::test
 * (source/GPISPAHHG.cu:2062-2085).  Host lon/lat degrees in, labels out. */
int gahhg_query_lonlat(gahhg_ctx* ctx, const double* lon, const double* lat, int64_t n,
                       int32_t* labels);
/* Host XYZ (n x 3 doubles, y up) in, labels out. */
int gahhg_query_xyz(gahhg_ctx* ctx, const double* xyz, int64_t n, int32_t* labels);
/* Device-resident XYZ and labels (pointers from gahhg_device_alloc or any
 * device allocation on the context's device); runs on the context stream
 * and waits for it.  This is GPISP::test's timed region. */
int gahhg_query_xyz_device(gahhg_ctx* ctx, const double* d_xyz, int64_t n, int32_t* d_labels);

/* Text input parsed on the GPU (GPISPTP::loadTPFromCSV, source/GPISPTP.cu:14-77,
 * and GPISPPolygon::loadPolyFromTXT, source/GPISPPolygon.cu:32-278): the
 * file's bytes are copied to device memory once; lines are split, numbers
 * parsed and points converted on the device.  Values equal the host
 * loaders' on every line (lines the device does not convert exactly are
 * parsed on the host with the reference's sscanf formats); a malformed file
 * gives the host loader's error. */
typedef struct gahhg_load_info {
  int64_t lines;       /* lines in the file */
  int64_t bytes;       /* file size */
  int64_t host_lines;  /* lines parsed on the host */
  double read_ms;      /* file read + copy to the device */
  double parse_ms;     /* line split + parse + conversion on the device */
  double total_ms;     /* the whole call */
} gahhg_load_info;

/* Query CSV ("%lf,%lf" = longitude, latitude per line) -> *d_xyz, n points
 * (FakeLL2XYZ) in device memory owned by the context, valid until the next
 * device load or gahhg_destroy; pass it to gahhg_query_xyz_device.  info may
 * be null. */
int gahhg_load_query_csv_device(gahhg_ctx* ctx, const char* path, const double** d_xyz, int64_t* n,
                                gahhg_load_info* info);
/* The same from text already in host memory (bytes long). */
int gahhg_parse_query_csv_device(gahhg_ctx* ctx, const char* text, int64_t bytes, const double** d_xyz,
                                 int64_t* n, gahhg_load_info* info);
/* gahhg_load_polygon_txt with the vertex lines parsed on the device. */
int gahhg_load_polygon_txt_device(gahhg_ctx* ctx, const char* path, gahhg_load_info* info);

/* Batches (no counterpart in the reference, which handles one polygon per
 * process): n_polys polygons whose rings are ring_off[poly_ring_off[k] ..
 * poly_ring_off[k+1]] (each ring as in gahhg_set_polygon_lonlat).  All are
 * built by gahhg_build with shared launches; each polygon's hierarchy and
 * labels equal what a single-polygon context builds for it. */
int gahhg_set_polygons_lonlat(gahhg_ctx* ctx, const double* lon, const double* lat, const int64_t* ring_off,
                              const int32_t* poly_ring_off, int32_t n_polys);
/* A batch read from n polygon files, each as gahhg_load_polygon_txt
 * (GPISPPolygon::loadPolyFromTXT, source/GPISPPolygon.cu:32); polygon k is
 * paths[k].  Replaces the reference's one process per file of
 * test/explist.bat with one batch. */
int gahhg_load_polygons_txt(gahhg_ctx* ctx, const char* const* paths, int32_t n);
int gahhg_num_polygons(gahhg_ctx* ctx, int32_t* n);
/* Queries against a batch: poly[i] selects the polygon of point i (the
 * device form checks every id on the device first: GAHHG_E_ARG if one is
 * outside the batch). */
int gahhg_query_batch_xyz(gahhg_ctx* ctx, const double* xyz, const int32_t* poly, int64_t n, int32_t* labels);
int gahhg_query_batch_xyz_device(gahhg_ctx* ctx, const double* d_xyz, const int32_t* d_poly, int64_t n,
                                 int32_t* d_labels);
/* Allocation-free form of gahhg_query_batch_xyz_device (d_poly may be null
 * for a single polygon): the same labels and statistics, but the grouping
 * pass's scratch (block counts, sort keys and point indices, radix
 * histograms, scan partial sums) comes from a fixed device array of the
 * library instead of buffers the context allocates or grows, and the pass is
 * never recorded as a graph, so the call makes no hipMalloc / hipFree.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (serving loops and harnesses that forbid device allocation).  At most
 * GAHHG_NOALLOC_MAX_POINTS points per call (GAHHG_E_LIMIT beyond).  Each GPU
 * holds one copy of the array (about 68 MB) until the process ends, shared
 * by every context on that GPU: calls that overlap on one GPU race.  The
 * call still waits for its stream, as the allocating form does. */
#define GAHHG_NOALLOC_MAX_POINTS 4194304
int gahhg_query_batch_xyz_device_noalloc(gahhg_ctx* ctx, const double* d_xyz, const int32_t* d_poly, int64_t n,
                                         int32_t* d_labels);
/* Every polygon of the batch against the same n points, as a series of
 * GPISP.exe runs with one query file (test/explist.bat): labels[p * n + i]
 * is point i in polygon p.  The points are sent once and repeated on the
 * device; n x polygons must stay below 2^31. */
int gahhg_query_pairs_xyz(gahhg_ctx* ctx, const double* xyz, int64_t n, int32_t* labels);
/* Per-polygon statistics and grid export in that polygon's own numbering. */
int gahhg_get_polygon_stats(gahhg_ctx* ctx, int32_t p, gahhg_stats* out);
int gahhg_polygon_grid_counts(gahhg_ctx* ctx, int32_t p, int64_t* counts5);
int gahhg_export_polygon_grid(gahhg_ctx* ctx, int32_t p, double* v, int32_t* cells, int32_t* borders,
                              double* normals, int32_t* refs, int32_t* levels);

/* Fixed depth per polygon: AHHG.exe's /maxdepth mode (test/testAHHG.bat:6,
 * GPISPParam.maxDepth with autoDepth off, source/GPISPAHHG.cu:1976-2015)
 * chosen per polygon of a batch.  depths[k] = levels for polygon k (capped
 * at max_depth), 0 = the context's own rule; polygons k >= n use the rule.
 * Call after setting the polygons (setting polygons clears it); takes
 * effect at the next gahhg_build. */
int gahhg_set_polygon_depths(gahhg_ctx* ctx, const int32_t* depths, int32_t n);

/* Autotuner (no counterpart in the reference): times the accelerated query
 * on the given device points (a representative sample) for every lane-group
 * width {per polygon, 4, 8, 16, 32, 64} (edge records tested together per
 * point; reported as 0 for per polygon) and
 * block size {64, 128, 256}, keeps the fastest in the context's parameters
 * and writes {group, block, microseconds} to out3 (may be NULL).  d_labels
 * receives the labels (the same for every setting). */
int gahhg_autotune_query(gahhg_ctx* ctx, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels,
                         int32_t* out3);
/* Query-layout autotuner (no counterpart in the reference): re-packs the
 * built hierarchy's query layout for each candidate (one polygon: start
 * table of 128 / 256 bins per side with no sub-tables, sub-tables of 8 at
 * 4 entries per edge, or 16 at 8; a batch: start-table budget 8, 32 or 0
 * entries per edge), times the layout build and the query on the given
 * device points, keeps the candidate with the least layout ms / passes +
 * query ms (passes >= 1: query passes expected per build) and writes
 * {query_bins, query_sub, query_sub_budget, query_table_budget, layout ms,
 * query ms} of it to out6 (may be NULL).  Labels are the same for every
 * candidate; d_labels receives them. */
int gahhg_autotune_layout(gahhg_ctx* ctx, const double* d_xyz, const int32_t* d_poly, int64_t n, int32_t* d_labels,
                          double passes, double* out6);

/* Trace recorder (no counterpart in the reference).  Query trace: the
 * accelerated query of gahhg_query_batch_xyz_device (same labels) that also
 * writes 6 int32 per point to d_trace: {leaf cell of the first-match
 * descent (global cell id, -1 = none), its top-level cell 0..41, its level,
 * level of the start cell taken from the table (-1 = level-0 scan, -2 =
 * resolved by the table alone), border planes evaluated, leaf edge records
 * scanned}.  d_poly may be NULL for a single polygon. */
int gahhg_trace_query_device(gahhg_ctx* ctx, const double* d_xyz, const int32_t* d_poly, int64_t n,
                             int32_t* d_labels, int32_t* d_trace);
/* Build trace of the last build.  cells[((p*32 + l)*42 + r)*4 + k] for
 * polygon p, level l (0..31), top-level cell r: k = {cells, refined cells,
 * bottom cells with refs, refs}.  edges[(p*42 + r)*3 + k] for edges whose
 * first endpoint lies in top-level cell r: k = {edges, sum of arc lengths,
 * sum of turning angles to the next edge of the ring}, sums in units of
 * 2^-40 rad.  Either pointer may be NULL. */
int gahhg_build_trace(gahhg_ctx* ctx, int32_t* cells, uint64_t* edges);

/* Learned refinement model (no counterpart in the reference, whose only
 * refinement rule is divGrid's density stop, source/GPISPAHHG.cu:1595-1605):
 * a 12-64-64-1 perceptron, ReLU hidden layers, BF16 weights, FP32
 * accumulation.  Parameter vector of GAHHG_MLP_PARAMS values, in order:
 * W1[64][16] (input columns 12..15 are padding and stay zero), b1[64],
 * W2[64][64], b2[64], w3[64], b3[1].  Device pointers; each call runs on the
 * context stream and waits for it.
 * forward:  y[r] = model(x[r*12 .. r*12+11]) for r < rows.
 * backward: grad = d/dparams of sum_r dy[r] * y[r] (FP32, deterministic).
 * adam:     one Adam step (beta 0.9 / 0.999, eps 1e-8) on the FP32 master
 *           parameters with moments m1, m2; writes the BF16 weights w. */
#define GAHHG_MLP_INPUTS 12
#define GAHHG_MLP_PARAMS 5313
int gahhg_mlp_forward(gahhg_ctx* ctx, const float* d_x, int64_t rows, const uint16_t* d_w, float* d_y);
int gahhg_mlp_backward(gahhg_ctx* ctx, const float* d_x, int64_t rows, const uint16_t* d_w, const float* d_dy,
                       float* d_grad);
int gahhg_mlp_adam(gahhg_ctx* ctx, float* d_master, const float* d_grad, float* d_m1, float* d_m2, int32_t step,
                   float lr, uint16_t* d_w);

/* Learned start-table depth (label-neutral; no counterpart in the
 * reference).  The hierarchy -- and with it every label -- stays the density
 * rule's (source/GPISPAHHG.cu:1595-1605).  What the model chooses per
 * polygon is the depth of the query's start structure: candidate c of
 *   bins per cube-face side {0, 16, 32, 64, 128, 256, 256, 256}
 *   sub-table side          {1,  1,  1,  1,   1,   1,   8,  16}
 * (a batch considers c <= 4; no candidate above query_bins / query_sub).
 * A table entry is a cell every point of its bin reaches under the
 * reference's first-match descent, so every candidate gives the same labels.
 * For each top-level cell r of the polygon the model predicts y_r(c), the
 * log2 query time per point (ns) of points in r with candidate c (feature 1
 * = log2(1 + bins x side) / 12, feature 6 = log2(1 + w_r) / 32; the other 10
 * as for the depth model, at the polygon's last refined level), and the
 * build takes
 *   argmin_c  cost[1] * sum_r w_r 2^y_r(c) + build_c
 * with w_r the points expected in r (query hint, else cost[0] per polygon
 * spread by edge share) and build_c = cost[2 + 2c] + cost[3 + 2c] * N ns
 * for one polygon, cost[18] * 6 * bins_c^2 ns in a batch (whose table
 * launches are shared).  params: GAHHG_MLP_PARAMS values (rounded to BF16
 * here); NULL switches the model off (the size rule of query_table_budget
 * applies).  cost: 19 values.  Takes effect at the next build. */
int gahhg_set_table_model(gahhg_ctx* ctx, const float* params, const double* cost);
/* Fixed start-table candidate per polygon (-1: the model's or the size
 * rule's choice), n = the polygons loaded; NULL or n = 0 clears.  Takes
 * effect at the next build. */
int gahhg_set_table_choice(gahhg_ctx* ctx, const int32_t* cand, int32_t n);
/* The start table of each polygon in the last build: bins_sub[2p] = bins per
 * face side (0 = none), bins_sub[2p + 1] = sub-table side (1 = none). */
int gahhg_get_table_layout(gahhg_ctx* ctx, int32_t* bins_sub);
/* The table model's features of the last build for every (polygon,
 * top-level cell, candidate): x[((p * 42 + r) * 8 + c) * 12 + f]. */
int gahhg_table_features(gahhg_ctx* ctx, float* x);

/* Learned refinement depth (changes labels: the reference's labels depend
 * on the depth of its hierarchy, so a depth other than the density rule's
 * gives other labels for some points; not used by the default paths).
 * With a model set, every polygon without a
 * fixed depth is refined level by level while the model's predicted query
 * time saved exceeds the predicted build time of the next level:
 *   sum_r w_r (2^y_r(stop) - 2^y_r(refine)) * cost[1]
 *       > cost[2] * N + cost[3] * (cells of the next level) + cost[4] / m   (ns)
 * over the top-level cells r of the polygon, with w_r the points expected
 * in r (from the query hint, else cost[0] points per polygon spread by edge
 * share), y_r(k) the model's log2 query time per point (ns) with
 * refinement stopping at this level (k = 0) or after one more (k = 1), and
 * m the polygons of the batch still refining (they share a level's cost).
 * The model decides only inside the window it is trained on: a polygon
 * always refines while its crossed cells are below 0.02 N, and refines at
 * most 2 levels past the level where the density rule stops.  It departs
 * from the density rule only for a predicted factor of 2: before the
 * rule's stop a polygon stops when 2 x saving < cost, at or past it a
 * polygon refines when saving > 2 x cost.
 * The hierarchy keeps the reference's structure (one bottom level per
 * polygon); max_depth still caps it.  params: GAHHG_MLP_PARAMS values as
 * for gahhg_mlp_*, rounded to BF16 here; NULL switches the model off.
 * cost: 5 values {expected points, query-time scale, ns per edge, ns per
 * next-level cell, ns per level}.  Takes effect at the next build. */
int gahhg_set_depth_model(gahhg_ctx* ctx, const float* params, const double* cost);
/* Sample of the points the hierarchy will answer (host XYZ, polygon ids
 * or NULL for polygon 0), each standing for `weight` points: gives the
 * expected points per polygon and top-level cell.  n = 0 clears it. */
int gahhg_set_query_hint(gahhg_ctx* ctx, const double* xyz, const int32_t* poly, int64_t n, double weight);
/* The model's 12 features for every (polygon, top-level cell, stop/refine)
 * row at `level` of the last build (cells refined there stand for the
 * cells crossed): x[((p * 42 + r) * 2 + k) * 12 + f]. */
int gahhg_depth_features(gahhg_ctx* ctx, int32_t level, float* x);

/* Grid serialisation (no counterpart in the reference, which rebuilds per
 * run): save the built hierarchies (polygons, every level, refs, labels) to
 * a file; load them into a context, which is then built (the query layout
 * is packed on load; pre_ms reports that time).  Files are tied to this
 * library version and byte order. */
int gahhg_save_grid(gahhg_ctx* ctx, const char* path);
int gahhg_load_grid(gahhg_ctx* ctx, const char* path);

/* Reads a query CSV ("lon,lat" per line) into host arrays.  With lon/lat
 * NULL it returns the count in *n. */
int gahhg_read_query_csv(const char* path, double* lon, double* lat, int64_t cap, int64_t* n);
/* Lune/Ketzner-format files: the reference ships copies of its polygons and
 * query points in the format of the comparators it was measured against
 * (README.md:5, test/data_for_Lune) and no reader of them; these read them
 * into the reference's own convention (lon/lat degrees) so that they take
 * the reference's path from there.  A line is "colatitude,longitude" in
 * radians, longitude counted from 180 deg West: lat = 90 - deg(colat),
 * lon = (deg(lon) mod 360) - 180, mod taking the divisor's sign.
 * A polygon file is one ring closed by repeating its first vertex, then a
 * point known to be inside (ref_lonlat, *has_ref = 1 when present).
 *
 * gahhg_read_lune_points: as gahhg_read_query_csv (NULL lon/lat = count).
 * gahhg_read_lune_polygon: the ring with its closing vertex (*n of them;
 *   NULL lon/lat = count), for gahhg_set_polygon_lonlat with one ring.
 * gahhg_load_polygon_lune: as gahhg_load_polygon_txt on the ring. */
int gahhg_read_lune_points(const char* path, double* lon, double* lat, int64_t cap, int64_t* n);
int gahhg_read_lune_polygon(const char* path, double* lon, double* lat, int64_t cap, int64_t* n,
                            double* ref_lonlat, int32_t* has_ref);
int gahhg_load_polygon_lune(gahhg_ctx* ctx, const char* path, double* ref_lonlat, int32_t* has_ref);

/* Query lon/lat -> XYZ exactly as the reference converts test points
 * (FakeLLVertex round trip, source/GPISPPolygon.cuh:72,82-92). */
int gahhg_query_lonlat_to_xyz(const double* lon, const double* lat, int64_t n, double* xyz);

/* GPISP::getStat (source/GPISPAHHG.cu:2088-2113). */
int gahhg_get_stats(gahhg_ctx* ctx, gahhg_stats* out);
/* GPISP::exportStat2TXT (source/GPISPAHHG.cu:2115-2161): appends the
 * report block; names are the polygon / query file names to print. */
int gahhg_write_report(gahhg_ctx* ctx, const char* path, const char* poly_name,
                       const char* query_name);
/* The same report block for polygon p of a batch, with the caller's query
 * figures for it (points, points inside, query and build milliseconds; a
 * batch shares its build and query launches, so callers pass their share). */
int gahhg_write_polygon_report(gahhg_ctx* ctx, int32_t p, const char* path, const char* poly_name,
                               const char* query_name, int64_t n_queries, int64_t inside, double query_ms,
                               double pre_ms);
/* GPISPTP::exportResult2CSV (source/GPISPTP.cu:81-99): "1" for IN else "0". */
int gahhg_write_result_csv(const char* path, const int32_t* labels, int64_t n);

/* Device memory helpers for callers without their own allocator. */
int gahhg_device_alloc(gahhg_ctx* ctx, int64_t bytes, void** d_ptr);
int gahhg_device_free(gahhg_ctx* ctx, void* d_ptr);
int gahhg_memcpy_h2d(gahhg_ctx* ctx, void* d_dst, const void* h_src, int64_t bytes);
int gahhg_memcpy_d2h(gahhg_ctx* ctx, void* h_dst, const void* d_src, int64_t bytes);

/* Grid export (for parity checks and serialisation): counts[5] =
 * {depth, V, B, C, R}; then host arrays sized from them: v[3V],
 * cells[7C] = (bstart,bnum,rstart,rnum,upcell,subcell,type),
 * borders[6B] = (parent,partner,v0,v1,subcell,id), normals[3B], refs[R]
 * (each cell's refs ascending), levels[7*depth] (vstart,vnum,bstart,bnum,
 * cstart,cnum,rnum).  Any output pointer may be NULL. */
int gahhg_grid_counts(gahhg_ctx* ctx, int64_t* counts5);
int gahhg_export_grid(gahhg_ctx* ctx, double* v, int32_t* cells, int32_t* borders,
                      double* normals, int32_t* refs, int32_t* levels);

/* Known-answer hooks: run the device predicates on n inputs.
 * op 0: normalize(a) -> out[3]; op 1: cross(a, b) -> out[3];
 * op 2: isEEInter(a, b, c, d) -> out[0] in {0,1};
 * op 3: isEdgeBorderInter(a, b, c, d) -> out[0] = class, out[1] = reserved. */
int gahhg_kat_device(gahhg_ctx* ctx, int32_t op, const double* in12, int64_t n, double* out3);
/* Known-answer hook for device_getCPProp (source/GPISPAHHG.cu:61-171) as
 * setBottomNECell runs it (a group of `group` lanes per cell, group a power
 * of two up to 64), over the built context's polygon edges (the edge
 * records the build makes): case k takes the refs
 * refs[rs_rn[2k], + rs_rn[2k+1]) (edge indices, at least one) and the
 * centre centers[3k..3k+2]; out3[3k..] = type (1 out, 2 in), the slot in the
 * case's refs of the chosen edge (-1: none) and its nearest vertex (-1: a
 * point inside the edge); mindist[k] = that feature's squared distance. */
int gahhg_kat_cp_prop(gahhg_ctx* ctx, const int32_t* refs, int64_t n_refs, const int32_t* rs_rn,
                      const double* centers, int64_t n, int32_t group, int32_t* out3, double* mindist);

/* Parity hook of the grouping sort (no reference counterpart): sorts the n
 * host pairs (keys[i], vals[i]) in place, stably, by key bits
 * [0, end_bit) (1..32), on the device, with launches sized for n_bound >= n
 * points and the count n read from device memory, as the grouped query
 * runs it. */
int gahhg_sort_pairs(gahhg_ctx* ctx, uint32_t* keys, uint32_t* vals, int64_t n, int64_t n_bound, int32_t end_bit);

#ifdef __cplusplus
}
#endif
#endif
