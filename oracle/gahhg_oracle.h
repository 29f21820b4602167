/*
 * gahhg_oracle.h -- CPU restatement of the GAHHG point-in-spherical-polygon
 * algorithm (hierarchy build, leaf classification, query).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against, and the timed CPU baseline.  Nothing in
 * the product (gahhg_amd/, include/) links, loads or calls it.
 *
 * Every routine follows the reference source with the same FP64 operation
 * order (no FMA contraction: built with -ffp-contract=off).  Citations are to
 * the reference repository (source/...).  Where the reference is racy or
 * reads out of bounds, the deterministic semantics chosen here are stated at
 * the routine.
 */
#ifndef GAHHG_ORACLE_H
#define GAHHG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct og_ctx og_ctx;

/* Level ranges: source/GPISPAHHG.cuh:12-20 (GAHHGGrid). */
typedef struct {
    int32_t vstart, vnum, bstart, bnum, cstart, cnum, rnum;
} og_level;

/* Diagnostics for behaviour the reference leaves undefined, plus roofline
 * counters.  All counts are exact. */
typedef struct {
    int64_t walk_steps_mark;       /* cells visited by all marking walks      */
    int64_t walk_steps_dispatch;   /* cells visited by the counting walk       */
    int64_t walk_partner_missing;  /* X/T1/OVERLAP exit through partner -1    */
    int64_t walk_step_cap_hits;    /* walks stopped by the step cap            */
    int64_t walk_unlocated;        /* edges whose p0 was not located           */
    int64_t overlap_bad_index;     /* marking OVERLAP index outside the array  */
    int64_t zero_ref_edges;        /* edges with 0 bottom-level refs           */
    int64_t undefined_leaves_snapshot; /* leaves a single classification pass
                                          leaves UNDEFINED                     */
    int64_t undefined_leaves_final;    /* leaves UNDEFINED after fixpoint      */
    int64_t partner_conflicts;     /* setSCPartner writes that disagree        */
    int32_t mark_passes;           /* marking passes run (M)                   */
    int32_t pad;
} og_diag;

/* Per-query work counters (roofline numerators). */
typedef struct {
    int64_t plane_tests;     /* T: n.p plane evaluations in the descent      */
    int64_t boundary_hits;   /* queries ending in a leaf with rnum > 0       */
    int64_t ref_tests;       /* isEEInter evaluations                        */
    int64_t bnum_sum;        /* sum of bnum over boundary leaves hit         */
    int64_t unlocated;       /* queries that reached no leaf                 */
    int64_t inside;          /* labels == IN                                 */
} og_qcount;

og_ctx* og_new(void);
void og_free(og_ctx* c);
const char* og_last_error(og_ctx* c);

/* Polygon input.  TXT format: source/GPISPPolygon.cu:32-278. */
int og_load_polygon_txt(og_ctx* c, const char* path);
/* lon/lat rings, each ring including its closing duplicate vertex;
 * ring_off has n_rings+1 entries (vertex offsets), ring 0 is the outer ring. */
int og_set_polygon_lonlat(og_ctx* c, const double* lon, const double* lat,
                          const int64_t* ring_off, int32_t n_rings);
int32_t og_num_edges(og_ctx* c);
int32_t og_orientation_flips(og_ctx* c);  /* bit r set: ring r was reversed */
void og_get_polygon(og_ctx* c, double* pv /*3N*/, int32_t* pe /*2N*/);

/* Point conversions: source/GPISPPolygon.cuh:21-30 and :72,82-92. */
/* sin and cos of x[i], correctly rounded (certified with long double and,
 * near a rounding midpoint, quad precision). */
void og_sincos_cr(const double* x, int64_t n, double* s, double* c);
void og_ll2xyz(const double* lon, const double* lat, int64_t n, double* xyz);
void og_query_lonlat_to_xyz(const double* lon, const double* lat, int64_t n,
                            double* xyz);
/* Query CSV "lon,lat" lines: source/GPISPTP.cu:13-79.  Returns count or -1;
 * call with lon==NULL to count only. */
int64_t og_load_query_csv(og_ctx* c, const char* path, double* lon, double* lat,
                          int64_t cap);

/* Level-0 grid: 42 cells / 80 vertices / 240 borders (restated from the
 * geometry the reference's closed grid-0 library implies). */
int og_grid0(double* v /*80*3*/, int32_t* border_i /*240*6*/,
             double* border_n /*240*3*/, int32_t* cell_i /*42*7*/);

/* Build + classify (source/GPISPAHHG.cu:1947-2059).  max_depth as in
 * GPISPParam (default 15); auto_depth 1 = the U/N density stop rule. */
int og_build(og_ctx* c, int32_t max_depth, int32_t auto_depth);
/* Build only up to (and including) dispatch, no classification. */
int og_build_create_only(og_ctx* c, int32_t max_depth, int32_t auto_depth);
int og_classify(og_ctx* c);

/* Counts: [0]=gnum [1]=V [2]=B [3]=C [4]=R; gamma via og_gamma. */
void og_counts(og_ctx* c, int64_t* out5);
float og_gamma(og_ctx* c);
void og_levels(og_ctx* c, og_level* out /*gnum*/);
void og_diagnostics(og_ctx* c, og_diag* out);
/* Grid export.  cells: 7 ints per cell (bstart,bnum,rstart,rnum,upcell,
 * subcell,type); borders: 6 ints (parent,partner,v0,v1,subcell,id) + 3
 * doubles (n); refs: R ints with each cell's run sorted ascending. */
void og_export(og_ctx* c, double* v, int32_t* cells, int32_t* borders_i,
               double* borders_n, int32_t* refs);

/* Query (source/GPISPAHHG.cu:1422-1504).  Labels: 2 IN, 1 OUT, 0 undefined
 * or unlocated.  nthreads <= 0 uses all OpenMP threads. */
void og_query_xyz(og_ctx* c, const double* xyz, int64_t n, int32_t* labels,
                  int32_t nthreads, og_qcount* counts /*nullable*/);

/* Leaf cell (index into the exported cells, -1 = unlocated) that the
 * first-match descent of kernel_PISP reaches for each point. */
void og_query_leaf(og_ctx* c, const double* xyz, int64_t n, int32_t* leaf);
/* Provenance of every cell's type after og_classify: rule[i] = 0 none,
 * 1 setBottomNECell (getCPProp), 2 setBottomDECell (parity from src across
 * border), 3 setIECell (adopted from src), 4 setDECell copy of src, 5
 * setDECell parity from src across border, 6 setTopECell (adopted from
 * src). */
void og_provenance(og_ctx* c, int32_t* rule, int32_t* src, int32_t* border);
/* Test hook: overwrites a cell's type (plants a mistake). */
void og_debug_set_type(og_ctx* c, int32_t cell, int32_t type);

/* Exact FP64 primitives exposed for known-answer tests. */
void og_kat_normalize(const double* in3, double* out3);
void og_kat_cross(const double* a3, const double* b3, double* out3);
int32_t og_kat_is_ee_inter(const double* p0, const double* p1,
                           const double* v0, const double* v1);
int32_t og_kat_edge_border_inter(const double* p0, const double* p1,
                                 const double* v0, const double* v1,
                                 int32_t* reserved);
/* device_getCPProp (source/GPISPAHHG.cu:61-171) over refs[0, rn) of the
 * polygon pv (3 doubles per vertex) / pe (2 vertex indices per edge):
 * the type (1 out, 2 in); slot = the chosen edge's position in refs (-1:
 * none chosen), nv = its nearest vertex (-1: a point inside the edge),
 * mindist = that feature's squared distance. */
int32_t og_kat_cp_prop(const double* pv, const int32_t* pe, const int32_t* refs, int32_t rn,
                       const double* center, int32_t* slot, int32_t* nv, double* mindist);

#ifdef __cplusplus
}
#endif
#endif
