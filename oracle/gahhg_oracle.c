/*
 * gahhg_oracle.c -- CPU oracle for the GAHHG point-in-spherical-polygon path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).  The
 * product library never links or calls this file.
 *
 * The reference (superredlark2/GAHHG) cannot be built here: it is a Windows
 * CUDA project and its level-0 grid lives only in a compiled library.  This
 * file restates its algorithm in plain C99 with the same FP64 operation order
 * (build with -ffp-contract=off).  Each routine cites the reference lines it
 * follows.  Deterministic choices where the reference is racy or undefined:
 *   - refs inside a cell are in ascending edge order (the reference's
 *     atomicAdd order is unspecified, source/GPISPAHHG.cu:1011-1027);
 *   - classification reads neighbour types from a snapshot taken before each
 *     pass and propagation passes repeat until nothing changes (the
 *     reference's single racy passes, :1074-1124,1366-1419);
 *   - a query that reaches no leaf gets label 0 (the reference leaves the
 *     output unwritten, :1446-1447,1470-1471);
 *   - a walk that would step through a border with no partner (a read of
 *     dev_b[-1] in the reference, :748-749,1222-1223) stops there; an edge
 *     whose start point is not located is dropped from deeper levels;
 *   - walks are capped at OG_STEP_CAP steps (the reference's while(1) loops
 *     are unbounded, :736,872,1210).
 * The diagnostics structure counts every occurrence of these cases.
 */
#include "gahhg_oracle.h"

#include <float.h>
#include <math.h>
#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Constants: source/GPISPParam.h:6-17, source/GPISPDS.cuh:7. */
#define OG_PI 3.14159265
#define OG_TRAVEL_EPS 1.0e-5
#define OG_EPS 1.0e-8
#define CELL_UNDEFINED 0
#define CELL_OUT 1
#define CELL_IN 2
#define INTER_NO 0
#define INTER_X 1
#define INTER_T1 2
#define INTER_T2 3
#define INTER_OVERLAP 4
#define EQU_ZERO(a) ((a) < OG_EPS && (a) > -OG_EPS)
#define LESS_ZERO(a) ((a) < -OG_EPS)
#define GRE_ZERO(a) ((a) > OG_EPS)
#define LESS_EQU_ZERO(a) ((a) < OG_EPS)
#define GRE_EQU_ZERO(a) ((a) > -OG_EPS)
#define OG_MAX_LEVELS 32
#define OG_STEP_CAP (1 << 24)

/* ---------------------------------------------------------------------- */
/* FP64 3-vector, operation order of source/GPISPDS.cuh:30-106.            */
/* ---------------------------------------------------------------------- */
typedef struct { double x, y, z; } v3;

static inline v3 mk(double x, double y, double z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vneg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline v3 vmul(v3 a, double k) { return mk(a.x * k, a.y * k, a.z * k); }
/* operator/ multiplies by the reciprocal: GPISPDS.cuh:66-74 */
static inline v3 vdiv(v3 a, double k) { double s = (double)1.0 / k; return mk(a.x * s, a.y * s, a.z * s); }
/* normalize multiplies by 1/sqrt: GPISPDS.cuh:76-82 */
static inline v3 vnorm(v3 a) {
    double t = 1.0 / sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return mk(a.x * t, a.y * t, a.z * t);
}
static inline double vsq(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline v3 vcross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* ---------------------------------------------------------------------- */
/* Grid records: source/common.h:169-186.                                  */
/* ---------------------------------------------------------------------- */
typedef struct { int32_t parent, partner, v[2], subcell, id; double n[3]; } og_border;
typedef struct { int32_t bstart, bnum, rstart, rnum, upcell, subcell, type; } og_cell;

struct og_ctx {
    char err[512];
    /* polygon */
    int32_t N;
    v3* pv;
    int32_t* pe;
    int32_t flips;
    /* grid */
    v3* v; int64_t vcap;
    og_border* b; int64_t bcap;
    og_cell* c; int64_t ccap;
    int32_t* refs; int64_t nref;
    og_level g[OG_MAX_LEVELS];
    int32_t gnum;
    float gamma;
    int32_t* loc_cur;  /* dev_r[2]: start cell of each edge at the last level */
    int32_t* loc_prev; /* dev_r[1]: start cell at the level before            */
    og_diag diag;
    int built;
    /* provenance of every cell's type (test hook): rule (OG_PROV_*), the
     * cell it came from and the border crossed (-1 = none) */
    int32_t* prov; int32_t* prov_src; int32_t* prov_b; int64_t provcap;
};

static void set_err(og_ctx* c, const char* msg) {
    snprintf(c->err, sizeof(c->err), "%s", msg);
}

og_ctx* og_new(void) {
    og_ctx* c = (og_ctx*)calloc(1, sizeof(og_ctx));
    return c;
}

static void free_grid(og_ctx* c) {
    free(c->v); free(c->b); free(c->c); free(c->refs);
    free(c->loc_cur); free(c->loc_prev);
    free(c->prov); free(c->prov_src); free(c->prov_b);
    c->v = NULL; c->b = NULL; c->c = NULL; c->refs = NULL;
    c->loc_cur = c->loc_prev = NULL;
    c->prov = c->prov_src = c->prov_b = NULL;
    c->provcap = 0;
    c->vcap = c->bcap = c->ccap = 0;
    c->nref = 0;
    c->gnum = 0;
    c->built = 0;
    memset(&c->diag, 0, sizeof(c->diag));
    memset(c->g, 0, sizeof(c->g));
}

void og_free(og_ctx* c) {
    if (!c) return;
    free_grid(c);
    free(c->pv); free(c->pe);
    free(c);
}

const char* og_last_error(og_ctx* c) { return c ? c->err : "null context"; }

/* ---------------------------------------------------------------------- */
/* Lon/lat conversions.                                                    */
/* ---------------------------------------------------------------------- */
/* sin and cos correctly rounded to double.  The reference calls the
 * platform's sin and cos, whose last bit depends on the C library; this
 * oracle (and the product) use the correctly rounded values, a function of
 * the argument alone.  Here: sinl/cosl (64-bit significand; within 1.3 units
 * of their last place on |x| <= 2 pi, measured against the quad routines)
 * decide the rounding unless their value lies within 8 such units of a
 * midpoint between two doubles; then sinq/cosq (113-bit significand) decide. */
static int near_double_midpoint(long double v) {
    double d = (double)v;
    if ((long double)d == v) return 0;
    double n = nextafter(d, v > (long double)d ? INFINITY : -INFINITY);
    long double mid = ((long double)d + (long double)n) * 0.5L;  /* exact */
    long double unit = ldexpl(1.0L, ilogbl(v) - 63);
    return fabsl(v - mid) <= 8.0L * unit;
}

static double cr_sin(double x) {
    long double v = sinl((long double)x);
    return near_double_midpoint(v) ? (double)sinq((__float128)x) : (double)v;
}

static double cr_cos(double x) {
    long double v = cosl((long double)x);
    return near_double_midpoint(v) ? (double)cosq((__float128)x) : (double)v;
}

void og_sincos_cr(const double* x, int64_t n, double* s, double* c) {
    for (int64_t i = 0; i < n; i++) {
        s[i] = cr_sin(x[i]);
        c[i] = cr_cos(x[i]);
    }
}

/* LLVertex::LL2XYZ, source/GPISPPolygon.cuh:21-30 (y is up). */
static inline v3 ll2xyz(double lo, double la) {
    double a = la * OG_PI / 180.0;
    double o = lo * OG_PI / 180.0;
    double ca = cr_cos(a);
    v3 p;
    p.x = ca * cr_sin(o);
    p.y = cr_sin(a);
    p.z = ca * cr_cos(o);
    return p;
}

/* FakeLLVertex(ll) + FakeLL2XYZ, source/GPISPPolygon.cuh:72,82-92. */
static inline v3 fakell2xyz(double lo, double la) {
    double fakela = la + 90;
    double realla = fakela > 180 ? 270 - fakela : fakela - 90;
    return ll2xyz(lo, realla);
}

void og_ll2xyz(const double* lon, const double* lat, int64_t n, double* xyz) {
    for (int64_t i = 0; i < n; i++) {
        v3 p = ll2xyz(lon[i], lat[i]);
        xyz[3 * i] = p.x; xyz[3 * i + 1] = p.y; xyz[3 * i + 2] = p.z;
    }
}

void og_query_lonlat_to_xyz(const double* lon, const double* lat, int64_t n, double* xyz) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        v3 p = fakell2xyz(lon[i], lat[i]);
        xyz[3 * i] = p.x; xyz[3 * i + 1] = p.y; xyz[3 * i + 2] = p.z;
    }
}

/* ---------------------------------------------------------------------- */
/* Polygon assembly: orientation + edge list.                              */
/* ---------------------------------------------------------------------- */
/* isPolyClockwise, source/GPISPPolygon.cu:280-313: at the first vertex of
 * maximum latitude a1 (with neighbours a0, b), CCW iff
 * normalize(a0 x (a1 - a0)) . b > 0. */
static int is_clockwise(const double* lo, const double* la, int vnum) {
    double maxvalue = -FLT_MAX;
    int maxindex = 0;
    for (int i = 0; i < vnum; i++) {
        if (la[i] > maxvalue) { maxvalue = la[i]; maxindex = i; }
    }
    v3 a0, a1, b;
    if (maxindex == 0) a0 = ll2xyz(lo[vnum - 1], la[vnum - 1]);
    else a0 = ll2xyz(lo[maxindex - 1], la[maxindex - 1]);
    a1 = ll2xyz(lo[maxindex], la[maxindex]);
    if (maxindex == vnum - 1) b = ll2xyz(lo[0], la[0]);
    else b = ll2xyz(lo[maxindex + 1], la[maxindex + 1]);
    v3 e = vsub(a1, a0);
    v3 n = vnorm(vcross(a0, e));
    double temp = vdot(n, b);
    return temp > 0 ? 0 : 1;
}

/* Builds pv/pe from rings given WITHOUT their closing duplicate.
 * ring_n[r] = unique vertex count of ring r.  Edge order and reversal:
 * source/GPISPPolygon.cu:209-264. */
static int assemble_polygon(og_ctx* c, const double* lo, const double* la,
                            const int64_t* ring_start, const int32_t* ring_n,
                            int32_t n_rings) {
    int64_t total = 0;
    for (int r = 0; r < n_rings; r++) total += ring_n[r];
    if (total <= 0 || total > 0x3fffffff) { set_err(c, "polygon: bad vertex count"); return -1; }
    free(c->pv); free(c->pe);
    c->pv = (v3*)malloc(sizeof(v3) * (size_t)total);
    c->pe = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)total);
    if (!c->pv || !c->pe) { set_err(c, "polygon: out of memory"); return -1; }
    c->N = (int32_t)total;
    c->flips = 0;
    int64_t k = 0;
    for (int r = 0; r < n_rings; r++) {
        for (int j = 0; j < ring_n[r]; j++) {
            c->pv[k++] = ll2xyz(lo[ring_start[r] + j], la[ring_start[r] + j]);
        }
    }
    int32_t index = 0;
    for (int r = 0; r < n_rings; r++) {
        int n = ring_n[r];
        int cw = is_clockwise(lo + ring_start[r], la + ring_start[r], n);
        /* outer ring reversed if clockwise, holes reversed if counter-clockwise */
        int change = (r == 0) ? cw : !cw;
        int startindex = index;
        if (change) c->flips |= (r < 31) ? (1 << r) : 0;
        if (!change) {
            for (int j = 0; j < n; j++) {
                c->pe[index * 2] = index;
                c->pe[index * 2 + 1] = index + 1;
                index++;
            }
            c->pe[index * 2 - 1] = startindex;
        } else {
            for (int j = n - 1; j >= 0; j--) {
                c->pe[index * 2] = startindex + j;
                c->pe[index * 2 + 1] = startindex + j - 1;
                index++;
            }
            c->pe[index * 2 - 1] = startindex + n - 1;
        }
    }
    free_grid(c);
    return 0;
}

int og_set_polygon_lonlat(og_ctx* c, const double* lon, const double* lat,
                          const int64_t* ring_off, int32_t n_rings) {
    if (n_rings < 1) { set_err(c, "polygon: need at least one ring"); return -1; }
    int64_t* starts = (int64_t*)malloc(sizeof(int64_t) * n_rings);
    int32_t* ns = (int32_t*)malloc(sizeof(int32_t) * n_rings);
    for (int r = 0; r < n_rings; r++) {
        int64_t cnt = ring_off[r + 1] - ring_off[r];
        /* count includes the closing duplicate; the reference rejects < 4
         * (source/GPISPPolygon.cu:87-92,125-130) */
        if (cnt < 4) { free(starts); free(ns); set_err(c, "polygon: ring with fewer than 4 vertices"); return -1; }
        starts[r] = ring_off[r];
        ns[r] = (int32_t)(cnt - 1);
    }
    int ret = assemble_polygon(c, lon, lat, starts, ns, n_rings);
    free(starts); free(ns);
    return ret;
}

/* Polygon TXT reader, source/GPISPPolygon.cu:32-278: header "POLYGON 1 %d",
 * outer ring header "0 %d", hole headers "%d %d", vertex lines "%lf %lf"
 * (lon lat); every ring repeats its first vertex last and that line is
 * skipped. */
static char* read_line(FILE* f, char** buf, size_t* cap) {
    size_t len = 0;
    int ch;
    if (*cap == 0) { *cap = 256; *buf = (char*)malloc(*cap); }
    while ((ch = fgetc(f)) != EOF) {
        if (len + 1 >= *cap) { *cap *= 2; *buf = (char*)realloc(*buf, *cap); }
        if (ch == '\n') break;
        (*buf)[len++] = (char)ch;
    }
    if (ch == EOF && len == 0) return NULL;
    (*buf)[len] = 0;
    return *buf;
}

int og_load_polygon_txt(og_ctx* c, const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) { set_err(c, "polygon: cannot open file"); return -1; }
    char* buf = NULL; size_t cap = 0;
    int inNum = 0, tmp = 0;
    char* line = read_line(f, &buf, &cap);
    if (!line || sscanf(line, "POLYGON 1 %d", &inNum) != 1 || inNum < 0) {
        fclose(f); free(buf); set_err(c, "polygon: line 1: wrong format"); return -1;
    }
    int n_rings = inNum + 1;
    int64_t* starts = (int64_t*)malloc(sizeof(int64_t) * n_rings);
    int32_t* ns = (int32_t*)malloc(sizeof(int32_t) * n_rings);
    int64_t vcap = 1024, vn = 0;
    double* lo = (double*)malloc(sizeof(double) * vcap);
    double* la = (double*)malloc(sizeof(double) * vcap);
    int ok = 1;
    for (int r = 0; r < n_rings && ok; r++) {
        line = read_line(f, &buf, &cap);
        int cnt = 0;
        if (r == 0) {
            if (!line || sscanf(line, "0 %d", &cnt) != 1) { ok = 0; set_err(c, "polygon: outer ring header: wrong format"); break; }
        } else {
            if (!line || sscanf(line, "%d %d", &tmp, &cnt) != 2) { ok = 0; set_err(c, "polygon: hole header: wrong format"); break; }
        }
        if (cnt < 4) { ok = 0; set_err(c, "polygon: number of edges should be greater than 3"); break; }
        starts[r] = vn;
        ns[r] = cnt - 1;
        for (int j = 0; j < cnt; j++) {
            line = read_line(f, &buf, &cap);
            double x, y;
            if (!line || sscanf(line, "%lf %lf", &x, &y) != 2) { ok = 0; set_err(c, "polygon: vertex line: wrong format"); break; }
            if (j == cnt - 1) break; /* closing duplicate */
            if (vn == vcap) {
                vcap *= 2;
                lo = (double*)realloc(lo, sizeof(double) * vcap);
                la = (double*)realloc(la, sizeof(double) * vcap);
            }
            lo[vn] = x; la[vn] = y; vn++;
        }
    }
    fclose(f);
    free(buf);
    int ret = -1;
    if (ok) ret = assemble_polygon(c, lo, la, starts, ns, n_rings);
    free(lo); free(la); free(starts); free(ns);
    return ret;
}

int32_t og_num_edges(og_ctx* c) { return c->N; }
int32_t og_orientation_flips(og_ctx* c) { return c->flips; }

void og_get_polygon(og_ctx* c, double* pv, int32_t* pe) {
    for (int i = 0; i < c->N; i++) {
        pv[3 * i] = c->pv[i].x; pv[3 * i + 1] = c->pv[i].y; pv[3 * i + 2] = c->pv[i].z;
        pe[2 * i] = c->pe[2 * i]; pe[2 * i + 1] = c->pe[2 * i + 1];
    }
}

/* Query CSV, source/GPISPTP.cu:13-79: one "lon,lat" per line ("%lf,%lf");
 * reading stops at end of file; any other line is an error. */
int64_t og_load_query_csv(og_ctx* c, const char* path, double* lon, double* lat, int64_t cap) {
    FILE* f = fopen(path, "rb");
    if (!f) { set_err(c, "query: cannot open file"); return -1; }
    char* buf = NULL; size_t bc = 0;
    int64_t n = 0;
    char* line;
    while ((line = read_line(f, &buf, &bc)) != NULL) {
        double x, y;
        if (sscanf(line, "%lf,%lf", &x, &y) != 2) {
            fclose(f); free(buf); set_err(c, "query: wrong format"); return -1;
        }
        if (lon) {
            if (n >= cap) { fclose(f); free(buf); set_err(c, "query: capacity exceeded"); return -1; }
            lon[n] = x; lat[n] = y;
        }
        n++;
    }
    fclose(f);
    free(buf);
    return n;
}

/* ---------------------------------------------------------------------- */
/* Level-0 grid.                                                           */
/* The reference builds it in a closed library (initgrid0, source/common.h */
/* :162-201, called at GPISPAHHG.cu:1962); only its counts are known:      */
/* 42 cells (12 pentagons, 30 hexagons), 80 vertices, 240 borders, every  */
/* border CCW seen from outside with n = normalize(v0 x (v1 - v0)).        */
/* Construction: dual of the frequency-2 geodesic icosahedron.  Cell      */
/* centres are the 12 icosahedron vertices and 30 edge midpoints; cell    */
/* corners are the normalised centroids of the 80 sub-triangles.          */
/* ---------------------------------------------------------------------- */
static void build_grid0(v3* V80, og_border* B240, og_cell* C42) {
    const double pi = 3.14159265358979323846;
    v3 ico[12];
    double lat = atan(0.5);
    ico[0] = mk(0.0, 1.0, 0.0);
    for (int k = 0; k < 5; k++) {
        double lon = (double)k * (2.0 * pi / 5.0);
        ico[1 + k] = vnorm(mk(cos(lat) * sin(lon), sin(lat), cos(lat) * cos(lon)));
        double lon2 = ((double)k + 0.5) * (2.0 * pi / 5.0);
        ico[6 + k] = vnorm(mk(cos(-lat) * sin(lon2), sin(-lat), cos(-lat) * cos(lon2)));
    }
    ico[11] = mk(0.0, -1.0, 0.0);
    int face[20][3];
    for (int k = 0; k < 5; k++) {
        int k1 = (k + 1) % 5;
        face[k][0] = 0; face[k][1] = 1 + k; face[k][2] = 1 + k1;
        face[5 + k][0] = 1 + k; face[5 + k][1] = 6 + k; face[5 + k][2] = 1 + k1;
        face[10 + k][0] = 1 + k1; face[10 + k][1] = 6 + k; face[10 + k][2] = 6 + k1;
        face[15 + k][0] = 6 + k; face[15 + k][1] = 11; face[15 + k][2] = 6 + k1;
    }
    /* orient every face counter-clockwise seen from outside */
    for (int f = 0; f < 20; f++) {
        v3 a = ico[face[f][0]], b = ico[face[f][1]], cc = ico[face[f][2]];
        v3 nrm = vcross(vsub(b, a), vsub(cc, a));
        if (vdot(nrm, vadd(vadd(a, b), cc)) < 0) {
            int t = face[f][1]; face[f][1] = face[f][2]; face[f][2] = t;
        }
    }
    /* 30 edges in order of first appearance; centre points */
    int edge[30][2];
    int ne = 0;
    v3 P[42];
    for (int i = 0; i < 12; i++) P[i] = ico[i];
    int fm[20][3]; /* centre index of midpoint of face edge (j, j+1) */
    for (int f = 0; f < 20; f++) {
        for (int j = 0; j < 3; j++) {
            int a = face[f][j], b = face[f][(j + 1) % 3];
            int lo = a < b ? a : b, hi = a < b ? b : a;
            int found = -1;
            for (int e = 0; e < ne; e++) if (edge[e][0] == lo && edge[e][1] == hi) { found = e; break; }
            if (found < 0) {
                edge[ne][0] = lo; edge[ne][1] = hi;
                P[12 + ne] = vnorm(vadd(ico[lo], ico[hi]));
                found = ne++;
            }
            fm[f][j] = 12 + found;
        }
    }
    /* 80 sub-triangles (CCW), corner = normalised centroid */
    int tri[80][3];
    for (int f = 0; f < 20; f++) {
        int a = face[f][0], b = face[f][1], cc = face[f][2];
        int mab = fm[f][0], mbc = fm[f][1], mca = fm[f][2];
        int* t0 = tri[4 * f + 0]; t0[0] = a; t0[1] = mab; t0[2] = mca;
        int* t1 = tri[4 * f + 1]; t1[0] = mab; t1[1] = b; t1[2] = mbc;
        int* t2 = tri[4 * f + 2]; t2[0] = mca; t2[1] = mbc; t2[2] = cc;
        int* t3 = tri[4 * f + 3]; t3[0] = mab; t3[1] = mbc; t3[2] = mca;
    }
    for (int t = 0; t < 80; t++) {
        V80[t] = vnorm(vadd(vadd(P[tri[t][0]], P[tri[t][1]]), P[tri[t][2]]));
    }
    /* cells: sub-triangles around each centre in CCW order */
    int bpos = 0;
    for (int x = 0; x < 42; x++) {
        int order[6];
        int cnt = 0;
        int first = -1, fy = -1, fz = -1;
        for (int t = 0; t < 80 && first < 0; t++) {
            for (int j = 0; j < 3; j++) if (tri[t][j] == x) {
                first = t; fy = tri[t][(j + 1) % 3]; fz = tri[t][(j + 2) % 3]; break;
            }
        }
        (void)fy;
        order[cnt++] = first;
        int z = fz;
        for (;;) {
            int nxt = -1, nz = -1;
            for (int t = 0; t < 80 && nxt < 0; t++) {
                for (int j = 0; j < 3; j++) {
                    if (tri[t][j] == x && tri[t][(j + 1) % 3] == z) { nxt = t; nz = tri[t][(j + 2) % 3]; break; }
                }
            }
            if (nxt == first || nxt < 0) break;
            order[cnt++] = nxt;
            z = nz;
        }
        og_cell* cell = &C42[x];
        cell->bstart = bpos; cell->bnum = cnt; cell->rstart = -1; cell->rnum = 0;
        cell->upcell = -1; cell->subcell = -1; cell->type = CELL_UNDEFINED;
        for (int i = 0; i < cnt; i++) {
            og_border* bd = &B240[bpos + i];
            bd->parent = x; bd->partner = -1; bd->subcell = -1; bd->id = i;
            bd->v[0] = order[i]; bd->v[1] = order[(i + 1) % cnt];
            /* getENormal, source/GPISPAHHG.cu:18-25 */
            v3 e0 = V80[bd->v[0]];
            v3 e1 = vsub(V80[bd->v[1]], V80[bd->v[0]]);
            v3 n = vnorm(vcross(e0, e1));
            bd->n[0] = n.x; bd->n[1] = n.y; bd->n[2] = n.z;
        }
        bpos += cnt;
    }
    for (int i = 0; i < 240; i++) {
        for (int j = 0; j < 240; j++) {
            if (B240[i].v[0] == B240[j].v[1] && B240[i].v[1] == B240[j].v[0]) { B240[i].partner = j; break; }
        }
    }
}

int og_grid0(double* v, int32_t* bi, double* bn, int32_t* ci) {
    v3 V[80]; og_border B[240]; og_cell C[42];
    build_grid0(V, B, C);
    for (int i = 0; i < 80; i++) { v[3 * i] = V[i].x; v[3 * i + 1] = V[i].y; v[3 * i + 2] = V[i].z; }
    for (int i = 0; i < 240; i++) {
        bi[6 * i] = B[i].parent; bi[6 * i + 1] = B[i].partner; bi[6 * i + 2] = B[i].v[0];
        bi[6 * i + 3] = B[i].v[1]; bi[6 * i + 4] = B[i].subcell; bi[6 * i + 5] = B[i].id;
        bn[3 * i] = B[i].n[0]; bn[3 * i + 1] = B[i].n[1]; bn[3 * i + 2] = B[i].n[2];
    }
    for (int i = 0; i < 42; i++) {
        ci[7 * i] = C[i].bstart; ci[7 * i + 1] = C[i].bnum; ci[7 * i + 2] = C[i].rstart;
        ci[7 * i + 3] = C[i].rnum; ci[7 * i + 4] = C[i].upcell; ci[7 * i + 5] = C[i].subcell;
        ci[7 * i + 6] = C[i].type;
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Storage growth (checkStorage, source/GPISPAHHG.cu:1820-1910): arrays   */
/* grow by doubling; contents are preserved.                              */
/* ---------------------------------------------------------------------- */
static int ensure(void** p, int64_t* cap, int64_t need, size_t elem) {
    if (need <= *cap) return 0;
    int64_t nc = *cap > 0 ? *cap * 2 : 1024;
    while (nc < need) nc *= 2;
    void* q = realloc(*p, (size_t)nc * elem);
    if (!q) return -1;
    *p = q; *cap = nc;
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Predicates.                                                             */
/* ---------------------------------------------------------------------- */
/* device_isPointInsideCell, source/GPISPAHHG.cu:208-224 */
static inline int inside_cell(const og_ctx* c, v3 p, int cid, int64_t* tests) {
    const og_cell* cc = &c->c[cid];
    for (int i = 0; i < cc->bnum; i++) {
        const double* n = c->b[cc->bstart + i].n;
        double ret = n[0] * p.x + n[1] * p.y + n[2] * p.z;
        if (tests) (*tests)++;
        if (ret < -OG_EPS) return -1;
    }
    return cid;
}

/* device_isEdgeBorderInter, source/GPISPAHHG.cu:226-284 */
static inline int edge_border_inter(v3 p0, v3 p1, v3 v0, v3 v1, int* reserved) {
    v3 n0 = vnorm(vcross(v0, vsub(v1, v0)));
    double r0 = vdot(n0, p0);
    double r1 = vdot(n0, p1);
    v3 n1 = vnorm(vcross(p0, vsub(p1, p0)));
    double r2 = vdot(n1, v0);
    double r3 = vdot(n1, v1);
    if (EQU_ZERO(r0) && EQU_ZERO(r1)) return INTER_OVERLAP;
    int same23 = (LESS_ZERO(r2) && LESS_ZERO(r3)) || (GRE_ZERO(r2) && GRE_ZERO(r3));
    int same01 = (LESS_ZERO(r0) && LESS_ZERO(r1)) || (GRE_ZERO(r0) && GRE_ZERO(r1));
    if (!same23 && EQU_ZERO(r0)) { *reserved = 0; return INTER_T1; }
    if (!same23 && EQU_ZERO(r1)) { *reserved = 1; return INTER_T1; }
    if (EQU_ZERO(r2) && EQU_ZERO(r3)) return INTER_OVERLAP;
    if (!same01 && EQU_ZERO(r2)) { *reserved = 0; return INTER_T2; }
    if (!same01 && EQU_ZERO(r3)) { *reserved = 1; return INTER_T2; }
    if (same01 || same23) return INTER_NO;
    return INTER_X;
}

/* device_isEEInter, source/GPISPAHHG.cu:351-378 */
static inline int is_ee_inter(v3 p0, v3 p1, v3 v0, v3 v1) {
    v3 n0 = vnorm(vcross(v0, vsub(v1, v0)));
    double r0 = vdot(n0, p0);
    double r1 = vdot(n0, p1);
    v3 n1 = vnorm(vcross(p0, vsub(p1, p0)));
    double r2 = vdot(n1, v0);
    double r3 = vdot(n1, v1);
    if ((LESS_ZERO(r0) && LESS_ZERO(r1)) || (GRE_ZERO(r0) && GRE_ZERO(r1)) ||
        (LESS_ZERO(r2) && LESS_ZERO(r3)) || (GRE_ZERO(r2) && GRE_ZERO(r3)))
        return 0;
    return 1;
}

/* dev_isInterSeg, source/GPISPAHHG.cu:174-205 */
static inline int is_inter_seg(v3 v0, v3 v1, v3 p0, v3 p1) {
    v3 n0 = vnorm(vcross(v0, vsub(v1, v0)));
    double r0 = vdot(n0, p0);
    double r1 = vdot(n0, p1);
    if ((LESS_ZERO(r0) && LESS_ZERO(r1)) || (GRE_ZERO(r0) && GRE_ZERO(r1))) return 0;
    v3 n1 = vnorm(vcross(p0, vsub(p1, p0)));
    r0 = vdot(n1, v0);
    r1 = vdot(n1, v1);
    if ((LESS_ZERO(r0) && LESS_ZERO(r1)) || (GRE_ZERO(r0) && GRE_ZERO(r1))) return 0;
    return 1;
}

static inline v3 vtx(const og_ctx* c, int i) { return c->v[i]; }

/* device_getSegExitBorder, source/GPISPAHHG.cu:286-348 */
static int seg_exit_border(const og_ctx* c, v3 p0, v3 p1, int cid, int in, int* intertype, int* info) {
    const og_cell* cc = &c->c[cid];
    for (int i = cc->bstart; i < cc->bstart + cc->bnum; i++) {
        if (i == in) continue;
        int temp = 0;
        int inter = edge_border_inter(p0, p1, vtx(c, c->b[i].v[0]), vtx(c, c->b[i].v[1]), &temp);
        if (inter == INTER_X) { *intertype = INTER_X; return i; }
        if (inter == INTER_T1) {
            if (temp == 0) { *intertype = INTER_T1; *info = 0; return i; }
            *intertype = INTER_T1; *info = 1; return -1;
        }
        if (inter == INTER_T2) {
            if (in != -1) {
                if (c->b[i].v[temp] == c->b[in].v[0] || c->b[i].v[temp] == c->b[in].v[1]) continue;
            }
            *intertype = INTER_T2; *info = temp; return i;
        }
        if (inter == INTER_OVERLAP) { *intertype = INTER_OVERLAP; return i; }
    }
    return -1;
}

/* ---------------------------------------------------------------------- */
/* Edge walk (markNECell :1153-1289, calcRefNum :685-818, writeRef        */
/* :820-955).  mode 0 marks crossed cells with rnum=999, mode 1 counts    */
/* visited cells, mode 2 writes visited cell ids.                          */
/* ---------------------------------------------------------------------- */
enum { WALK_MARK = 0, WALK_COUNT = 1, WALK_WRITE = 2 };

typedef struct { int64_t partner_missing, cap_hits, unlocated, bad_index, steps; } walk_diag;

static int64_t walk_edge(og_ctx* c, int level, int e, int mode, int32_t* buf,
                         int32_t* out_cells, int64_t nb_valid, walk_diag* d) {
    v3 p0 = c->pv[c->pe[e * 2]];
    v3 p1 = c->pv[c->pe[e * 2 + 1]];
    int cur = -1;
    if (level == 0) {
        int lim = c->g[0].cnum; /* 42 */
        for (int i = 0; i < lim; i++) {
            cur = inside_cell(c, p0, i, NULL);
            if (cur != -1) break;
        }
        if (cur == -1) { d->unlocated++; if (mode == WALK_MARK) buf[e] = -1; return 0; }
        if (mode == WALK_MARK) buf[e] = cur;
    } else {
        int par = buf[e];
        if (par < 0) { d->unlocated++; return 0; }
        int sc = c->c[par].subcell;
        cur = sc >= 0 ? inside_cell(c, p0, sc, NULL) : -1;
        if (cur == -1) {
            const og_cell* pc = &c->c[par];
            for (int i = 0; i < pc->bnum; i++) {
                cur = c->b[pc->bstart + i].subcell;
                if (cur == -1) continue;
                cur = inside_cell(c, p0, cur, NULL);
                if (cur != -1) break;
            }
        }
        if (cur == -1) { d->unlocated++; if (mode == WALK_MARK) buf[e] = -1; return 0; }
        if (mode == WALK_MARK || mode == WALK_WRITE) buf[e] = cur;
    }

    int in = -1, out;
    int64_t steps = 0;
    for (;;) {
        if (steps >= OG_STEP_CAP) { d->cap_hits++; break; }
        if (mode == WALK_MARK) c->c[cur].rnum = 999;
        else if (mode == WALK_WRITE) out_cells[steps] = cur;
        steps++;
        int intertype = 0, info = 0;
        out = seg_exit_border(c, p0, p1, cur, in, &intertype, &info);
        if (out == -1) break;
        if (intertype == INTER_X || intertype == INTER_T1) {
            int pt = c->b[out].partner;
            if (pt == -1) { d->partner_missing++; break; }
            in = pt;
            cur = c->b[in].parent;
        } else if (intertype == INTER_T2) {
            v3 curpoint = vtx(c, c->b[out].v[info]);
            v3 ed = vnorm(vsub(p1, p0));
            curpoint = vadd(curpoint, vmul(ed, OG_TRAVEL_EPS));
            int u = c->c[cur].bnum;
            int pt = c->b[out].partner;
            if (pt != -1) {
                int ret = inside_cell(c, curpoint, c->b[pt].parent, NULL);
                if (ret != -1) { cur = ret; in = pt; continue; }
            }
            int other;
            if (info == 0) other = c->c[cur].bstart + (c->b[out].id + u - 1) % u;
            else other = c->c[cur].bstart + (c->b[out].id + 1) % u;
            int po = c->b[other].partner;
            if (po != -1) {
                int ret = inside_cell(c, curpoint, c->b[po].parent, NULL);
                if (ret != -1) { cur = ret; in = po; continue; }
                break;
            }
            break;
        } else { /* INTER_OVERLAP */
            v3 e0 = vnorm(vsub(p1, p0));
            v3 e1 = vnorm(vsub(vtx(c, c->b[out].v[1]), vtx(c, c->b[out].v[0])));
            double temp = vdot(e0, e1);
            int u = c->c[cur].bnum;
            int bs = c->c[cur].bstart;
            int k = temp > 0 ? (out - bs + 1) % u : (out - bs + u - 1) % u;
            int tgt = bs + k;
            int pt = c->b[tgt].partner;
            if (pt == -1) { d->partner_missing++; break; }
            int nc = c->b[pt].parent;
            if (mode == WALK_MARK) {
                /* source/GPISPAHHG.cu:1279-1285 computes `in` after curcell
                 * has changed, with the new cell's bstart; kept as is. */
                int nbs = c->c[nc].bstart;
                int kk = temp > 0 ? (out - nbs + 1) % u : (out - nbs + u - 1) % u;
                int idx = nbs + kk;
                if (idx < 0 || idx >= nb_valid) { d->bad_index++; in = -1; }
                else in = c->b[idx].partner;
                cur = nc;
            } else {
                cur = nc;
                in = pt;
            }
        }
    }
    d->steps += steps;
    return steps;
}

static void walk_all(og_ctx* c, int level, int mode, int32_t* buf, const int64_t* off,
                     int32_t* counts, int32_t* refs_cell) {
    int64_t nb_valid = (int64_t)c->g[level].bstart + c->g[level].bnum;
    int64_t pm = 0, ch = 0, un = 0, bi = 0, st = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : pm, ch, un, bi, st)
    for (int e = 0; e < c->N; e++) {
        walk_diag d = {0, 0, 0, 0, 0};
        int64_t s = walk_edge(c, level, e, mode, buf, refs_cell ? refs_cell + off[e] : NULL, nb_valid, &d);
        if (counts) counts[e] = (int32_t)s;
        pm += d.partner_missing; ch += d.cap_hits; un += d.unlocated; bi += d.bad_index; st += d.steps;
    }
    c->diag.walk_partner_missing += pm;
    c->diag.walk_step_cap_hits += ch;
    c->diag.walk_unlocated += un;
    c->diag.overlap_bad_index += bi;
    if (mode == WALK_MARK) c->diag.walk_steps_mark += st;
    if (mode == WALK_COUNT) c->diag.walk_steps_dispatch += st;
}

/* ---------------------------------------------------------------------- */
/* Build.                                                                  */
/* ---------------------------------------------------------------------- */
static int grow_v(og_ctx* c, int64_t n) { return ensure((void**)&c->v, &c->vcap, n, sizeof(v3)); }
static int grow_b(og_ctx* c, int64_t n) { return ensure((void**)&c->b, &c->bcap, n, sizeof(og_border)); }
static int grow_c(og_ctx* c, int64_t n) { return ensure((void**)&c->c, &c->ccap, n, sizeof(og_cell)); }

static v3 cell_center_sum(const og_ctx* c, int cid) {
    const og_cell* cc = &c->c[cid];
    v3 center = mk(0.0, 0.0, 0.0);
    for (int i = 0; i < cc->bnum; i++) center = vadd(center, c->v[c->b[cc->bstart + i].v[0]]);
    center = vdiv(center, cc->bnum);
    return vnorm(center);
}

static void set_border_normal(og_ctx* c, int bi) {
    v3 e0 = c->v[c->b[bi].v[0]];
    v3 e1 = vsub(c->v[c->b[bi].v[1]], c->v[c->b[bi].v[0]]);
    v3 n = vnorm(vcross(e0, e1));
    c->b[bi].n[0] = n.x; c->b[bi].n[1] = n.y; c->b[bi].n[2] = n.z;
}

/* device_setBorderPartner, source/GPISPAHHG.cu:42-58 */
static void set_border_partner(og_ctx* c, int cellindex, int edgeindex, int32_t* wr) {
    og_cell* cc = &c->c[cellindex];
    for (int i = 0; i < cc->bnum; i++) {
        int j = cc->bstart + i;
        if (c->b[edgeindex].v[0] == c->b[j].v[1] && c->b[edgeindex].v[1] == c->b[j].v[0] && edgeindex < j) {
            if (wr[edgeindex] != INT32_MIN && wr[edgeindex] != j) c->diag.partner_conflicts++;
            if (wr[j] != INT32_MIN && wr[j] != edgeindex) c->diag.partner_conflicts++;
            c->b[edgeindex].partner = j; wr[edgeindex] = j;
            c->b[j].partner = edgeindex; wr[j] = edgeindex;
        }
    }
}

static void set_partner_w(og_ctx* c, int bi, int val, int32_t* wr) {
    if (wr[bi] != INT32_MIN && wr[bi] != val) c->diag.partner_conflicts++;
    c->b[bi].partner = val;
    wr[bi] = val;
}

/* GAHHG::divGrid, source/GPISPAHHG.cu:1558-1697.  Returns 0 = stop
 * (auto-depth), 1 = refined, -1 = error. */
static int div_grid(og_ctx* c, int lv, int auto_depth) {
    og_level* G = &c->g[lv];
    /* kernel_DivGridFast_zeroRefCount :1127-1150 */
    for (int t = 0; t < G->cnum; t++) {
        og_cell* cc = &c->c[G->cstart + t];
        cc->rnum = 0;
        if (G->cstart == 0) {
            cc->rstart = -1; cc->subcell = -1; cc->upcell = -1; cc->type = CELL_UNDEFINED;
            for (int i = 0; i < cc->bnum; i++) c->b[cc->bstart + i].subcell = -1;
        }
    }
    /* kernel_DivGridFast_markNECell :1153-1289; buffers dev_r[1], dev_r[2] */
    memcpy(c->loc_prev, c->loc_cur, sizeof(int32_t) * c->N);
    walk_all(c, lv, WALK_MARK, c->loc_cur, NULL, NULL, NULL);
    c->diag.mark_passes++;
    /* kernel_DivGrid_calcCenterOffsets :381-404 */
    int U = 0;
    for (int t = 0; t < G->cnum; t++) if (c->c[G->cstart + t].rnum != 0) U++;
    if (auto_depth) {
        /* :1597-1604 */
        float k = U * 1.0 / c->N;
        c->gamma = k;
        if (k > 0.1 || (k < 0.1 && (0.1 - k) < 0.01)) return 0;
    }
    int64_t vbase = (int64_t)G->vstart + G->vnum;
    int64_t bbase = (int64_t)G->bstart + G->bnum;
    int64_t cbase = (int64_t)G->cstart + G->cnum;
    /* exclusive scans of (bnum, 1) over refined cells */
    int64_t nv = 0;
    for (int t = 0; t < G->cnum; t++) if (c->c[G->cstart + t].rnum != 0) nv += c->c[G->cstart + t].bnum;
    if (vbase + nv > 0x7fffffff || bbase + 7 * nv > 0x7fffffff || cbase + U + nv > 0x7fffffff) {
        set_err(c, "build: grid exceeds int32 indexing"); return -1;
    }
    if (grow_v(c, vbase + nv) || grow_b(c, bbase + nv + 6 * nv) || grow_c(c, cbase + U + nv)) {
        set_err(c, "build: out of memory"); return -1;
    }
    /* kernel_DivGrid_createCenter :406-470 */
    int64_t o0 = 0, o1 = 0;
    for (int t = 0; t < G->cnum; t++) {
        int pid = G->cstart + t;
        if (c->c[pid].rnum == 0) continue;
        og_cell P = c->c[pid];
        v3 center;
        if (P.bnum == 4) {
            center = vmul(vadd(c->v[c->b[P.bstart].v[0]], c->v[c->b[P.bstart].v[1]]), 0.5);
        } else {
            center = mk(0.0, 0.0, 0.0);
            for (int i = 0; i < P.bnum; i++) center = vadd(center, c->v[c->b[P.bstart + i].v[0]]);
            center = vdiv(center, P.bnum);
        }
        center = vnorm(center);
        for (int i = 0; i < P.bnum; i++)
            c->v[vbase + o0 + i] = vnorm(vmul(vadd(c->v[c->b[P.bstart + i].v[0]], center), 0.5));
        int32_t child = (int32_t)(cbase + o1);
        for (int i = 0; i < P.bnum; i++) {
            int64_t bi = bbase + o0 + i;
            og_border* bd = &c->b[bi];
            bd->id = i; bd->parent = child; bd->partner = -1; bd->subcell = -1;
            bd->v[0] = (int32_t)(vbase + o0 + i);
            bd->v[1] = (int32_t)(vbase + o0 + (i + 1) % P.bnum);
            set_border_normal(c, (int)bi);
        }
        og_cell* cc = &c->c[child];
        cc->bstart = (int32_t)(bbase + o0); cc->bnum = P.bnum; cc->rstart = -1; cc->rnum = -1;
        cc->upcell = pid; cc->subcell = -1; cc->type = CELL_UNDEFINED;
        c->c[pid].subcell = child;
        o0 += P.bnum; o1 += 1;
    }
    og_level* H = &c->g[lv + 1];
    H->vstart = (int32_t)vbase; H->vnum = (int32_t)o0;
    H->bstart = (int32_t)bbase; H->bnum = (int32_t)o0;
    H->cstart = (int32_t)cbase; H->cnum = (int32_t)o1;
    G->rnum = (int32_t)o1;
    /* kernel_DivGrid_calcBorderSCOffsets :473-509 + createBorderSC :512-607 */
    int64_t eoff = (int64_t)H->bstart + H->bnum, coff = (int64_t)H->cstart + H->cnum;
    int64_t e0 = 0, c0 = 0;
    int nbc = H->bnum;
    for (int t = 0; t < nbc; t++) {
        int bi = H->bstart + t;
        og_border B = c->b[bi];
        int Pid = c->c[B.parent].upcell;
        og_cell P = c->c[Pid];
        int ue = P.bstart + B.id;
        if (P.bnum == 4 && B.id == 0) continue;
        og_border up = c->b[ue];
        int eo = (int)(eoff + e0), co = (int)(coff + c0);
        if (up.partner != -1 && c->c[c->b[up.partner].parent].subcell != -1 &&
            up.parent < c->b[up.partner].parent) {
            og_cell psub = c->c[c->c[c->b[up.partner].parent].subcell];
            og_border pse = c->b[psub.bstart + c->b[up.partner].id];
            int vv[6] = {up.v[0], pse.v[1], pse.v[0], up.v[1], B.v[1], B.v[0]};
            for (int i = 0; i < 6; i++) {
                og_border* nb = &c->b[eo + i];
                nb->id = i; nb->v[0] = vv[i]; nb->v[1] = vv[(i + 1) % 6];
                nb->parent = co; nb->partner = -1; nb->subcell = -1;
                set_border_normal(c, eo + i);
            }
            og_cell* cc = &c->c[co];
            cc->bnum = 6; cc->bstart = eo; cc->rstart = -1; cc->rnum = -1; cc->subcell = -1;
            cc->upcell = up.parent; cc->type = CELL_UNDEFINED;
            c->b[ue].subcell = co;
            c->b[up.partner].subcell = co;
            e0 += 6; c0 += 1;
        } else if (up.partner == -1 || c->c[c->b[up.partner].parent].subcell == -1) {
            int vv[4] = {up.v[0], up.v[1], B.v[1], B.v[0]};
            for (int i = 0; i < 4; i++) {
                og_border* nb = &c->b[eo + i];
                nb->id = i; nb->v[0] = vv[i]; nb->v[1] = vv[(i + 1) % 4];
                nb->parent = co; nb->partner = -1; nb->subcell = -1;
                set_border_normal(c, eo + i);
            }
            og_cell* cc = &c->c[co];
            cc->bnum = 4; cc->bstart = eo; cc->rstart = -1; cc->rnum = -1; cc->subcell = -1;
            cc->upcell = up.parent; cc->type = CELL_UNDEFINED;
            c->b[ue].subcell = co;
            e0 += 4; c0 += 1;
        }
    }
    int bsc_start = H->cstart + H->cnum;
    int bsc_num = (int)c0;
    H->cnum += (int32_t)c0;
    H->bnum += (int32_t)e0;
    /* kernel_DivGrid_setSCPartner :610-682 */
    int64_t nbtot = (int64_t)H->bstart + H->bnum;
    int32_t* wr = (int32_t*)malloc(sizeof(int32_t) * (size_t)nbtot);
    for (int64_t i = 0; i < nbtot; i++) wr[i] = INT32_MIN;
    for (int t = 0; t < bsc_num; t++) {
        int X = bsc_start + t;
        og_cell cur = c->c[X];
        int Pu = cur.upcell;
        int upedge = -1;
        for (int i = 0; i < c->c[Pu].bnum; i++) {
            if (c->b[c->c[Pu].bstart + i].subcell == X) { upedge = c->c[Pu].bstart + i; break; }
        }
        if (upedge == -1) continue;
        int ub = c->c[Pu].bnum;
        int upid = c->b[upedge].id;
        int te;
        if (cur.bnum == 6) {
            int pe_ = c->b[upedge].partner;
            int pc = c->b[pe_].parent;
            int u = c->c[pc].bnum;
            int pid = c->b[pe_].id;
            te = c->c[pc].bstart + (pid + 1) % u;
            if (c->b[te].subcell != -1) set_border_partner(c, c->b[te].subcell, cur.bstart, wr);
            else set_partner_w(c, cur.bstart, -1, wr);
            te = c->c[c->c[pc].subcell].bstart + pid;
            set_partner_w(c, cur.bstart + 1, te, wr);
            set_partner_w(c, te, cur.bstart + 1, wr);
            te = c->c[pc].bstart + (pid + u - 1) % u;
            if (c->b[te].subcell != -1) set_border_partner(c, c->b[te].subcell, cur.bstart + 2, wr);
            else set_partner_w(c, cur.bstart + 2, -1, wr);
            te = c->c[Pu].bstart + (upid + 1) % ub;
            if (c->b[te].subcell != -1) set_border_partner(c, c->b[te].subcell, cur.bstart + 3, wr);
            else set_partner_w(c, cur.bstart + 3, -1, wr);
            te = c->c[c->c[Pu].subcell].bstart + upid;
            set_partner_w(c, cur.bstart + 4, te, wr);
            set_partner_w(c, te, cur.bstart + 4, wr);
            te = c->c[Pu].bstart + (upid + ub - 1) % ub;
            if (c->b[te].subcell != -1) set_border_partner(c, c->b[te].subcell, cur.bstart + 5, wr);
            else set_partner_w(c, cur.bstart + 5, -1, wr);
        } else if (cur.bnum == 4) {
            te = c->c[Pu].bstart + (upid + 1) % ub;
            if (c->b[te].subcell != -1) set_border_partner(c, c->b[te].subcell, cur.bstart + 1, wr);
            else set_partner_w(c, cur.bstart + 1, -1, wr);
            te = c->c[c->c[Pu].subcell].bstart + upid;
            set_partner_w(c, cur.bstart + 2, te, wr);
            set_partner_w(c, te, cur.bstart + 2, wr);
            te = c->c[Pu].bstart + (upid + ub - 1) % ub;
            if (c->b[te].subcell != -1) set_border_partner(c, c->b[te].subcell, cur.bstart + 3, wr);
            else set_partner_w(c, cur.bstart + 3, -1, wr);
        }
    }
    free(wr);
    return 1;
}

static int cmp_i32(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* GAHHG::dispatchSELast, source/GPISPAHHG.cu:1700-1784, with its
 * parent-cell buffer `par` (dev_r[2] after the swaps at :1999-2004). */
static int dispatch_last(og_ctx* c, int32_t* par) {
    int L = c->gnum - 1;
    int N = c->N;
    int32_t* cnt = (int32_t*)malloc(sizeof(int32_t) * N);
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (N + 1));
    walk_all(c, L, WALK_COUNT, par, NULL, cnt, NULL);
    off[0] = 0;
    for (int e = 0; e < N; e++) off[e + 1] = off[e] + cnt[e];
    int64_t R = off[N];
    if (R > 0x7fffffff) { free(cnt); free(off); set_err(c, "dispatch: too many refs"); return -1; }
    for (int e = 0; e < N; e++) if (cnt[e] == 0) c->diag.zero_ref_edges++;
    int32_t* rcell = (int32_t*)malloc(sizeof(int32_t) * (size_t)(R > 0 ? R : 1));
    walk_all(c, L, WALK_WRITE, par, off, NULL, rcell);
    /* ref -> edge id by head flags + inclusive scan (:958-994): flag at the
     * first ref of every edge but edge 0; edges without refs share a slot. */
    int32_t* flag = (int32_t*)calloc((size_t)(R > 0 ? R : 1), sizeof(int32_t));
    for (int e = 1; e < N; e++) if (off[e] < R) flag[off[e]] = 1;
    int32_t* redge = (int32_t*)malloc(sizeof(int32_t) * (size_t)(R > 0 ? R : 1));
    int32_t run = 0;
    for (int64_t i = 0; i < R; i++) { run += flag[i]; redge[i] = run; }
    /* counting sort into cell-major order (:997-1043) */
    og_level* G = &c->g[L];
    int32_t* ccnt = (int32_t*)calloc((size_t)G->cnum + 1, sizeof(int32_t));
    for (int64_t i = 0; i < R; i++) ccnt[rcell[i] - G->cstart]++;
    int64_t* coff = (int64_t*)malloc(sizeof(int64_t) * ((size_t)G->cnum + 1));
    coff[0] = 0;
    for (int t = 0; t < G->cnum; t++) coff[t + 1] = coff[t] + ccnt[t];
    free(c->refs);
    c->refs = (int32_t*)malloc(sizeof(int32_t) * (size_t)(R > 0 ? R : 1));
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * ((size_t)G->cnum + 1));
    memcpy(fill, coff, sizeof(int64_t) * ((size_t)G->cnum + 1));
    for (int64_t i = 0; i < R; i++) c->refs[fill[rcell[i] - G->cstart]++] = redge[i];
    for (int t = 0; t < G->cnum; t++) {
        /* ascending edge id within each cell */
        if (ccnt[t] > 1) qsort(c->refs + coff[t], (size_t)ccnt[t], sizeof(int32_t), cmp_i32);
        c->c[G->cstart + t].rnum = ccnt[t];
        c->c[G->cstart + t].rstart = (int32_t)coff[t];
    }
    c->nref = R;
    G->rnum = (int32_t)R;
    free(cnt); free(off); free(rcell); free(flag); free(redge); free(ccnt); free(coff); free(fill);
    return 0;
}

int og_build_create_only(og_ctx* c, int32_t max_depth, int32_t auto_depth) {
    if (!c->pv || c->N <= 0) { set_err(c, "build: no polygon"); return -1; }
    if (max_depth < 1 || max_depth > OG_MAX_LEVELS - 1) { set_err(c, "build: max_depth out of range"); return -1; }
    free_grid(c);
    /* createAHHGInit :1947-1974 */
    int N = c->N;
    if (grow_v(c, N > 160 ? N : 160) || grow_b(c, (int64_t)N * 4 > 480 ? (int64_t)N * 4 : 480) ||
        grow_c(c, N > 84 ? N : 84)) { set_err(c, "build: out of memory"); return -1; }
    build_grid0(c->v, c->b, c->c);
    c->g[0].vstart = 0; c->g[0].vnum = 80;
    c->g[0].bstart = 0; c->g[0].bnum = 240;
    c->g[0].cstart = 0; c->g[0].cnum = 42;
    c->loc_cur = (int32_t*)calloc((size_t)N, sizeof(int32_t));
    c->loc_prev = (int32_t*)calloc((size_t)N, sizeof(int32_t));
    /* createAHHG :1976-2015 */
    c->gnum = max_depth;
    int early = 0;
    for (int i = 0; i < c->gnum - 1; i++) {
        int r = div_grid(c, i, auto_depth);
        if (r < 0) return -1;
        if (r == 0) { early = 1; c->gnum = i + 1; break; }
    }
    if (max_depth == 1) c->gamma = (float)(42 * 1.0 / c->N);
    int32_t* par = c->loc_cur;
    if (early && c->gnum != 1) par = c->loc_prev;
    if (dispatch_last(c, par) < 0) return -1;
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Classification (source/GPISPAHHG.cu:1046-1124,1292-1419,1787-1817,      */
/* 2018-2059).                                                             */
/* ---------------------------------------------------------------------- */
/* device_getCPProp, source/GPISPAHHG.cu:61-171, over refs[0, rn); the
 * chosen feature goes to slot (its position in refs, -1 if none was
 * chosen), nvo (its vertex, -1 for a point inside the edge) and md (its
 * squared distance) when those are not NULL. */
static int cp_prop_refs(const v3* pv, const int32_t* pe, const int32_t* refs, int rn, v3 center,
                        int* slot, int* nvo, double* md) {
    double mindist = FLT_MAX;
    int ne = refs[0];
    v3 ne_normal = mk(0.0, 0.0, 0.0);
    int nv_index = -1;
    int sl = -1;
    for (int i = 0; i < rn; i++) {
        int se = refs[i];
        double curdist = FLT_MAX;
        v3 v0 = pv[pe[se * 2]];
        v3 v1 = pv[pe[se * 2 + 1]];
        v3 n0 = vnorm(vcross(v0, vsub(v1, v0)));
        v3 n1 = vnorm(vcross(center, n0));
        v3 inter = vnorm(vcross(n0, n1));
        double temp = vdot(inter, v0);
        if (LESS_ZERO(temp)) inter = vneg(inter);
        v3 t0 = vnorm(vsub(v0, inter));
        v3 t1 = vnorm(vsub(v1, inter));
        double temp3 = vdot(t0, t1);
        int cur_nv = -1;
        if (LESS_ZERO(temp3)) {
            curdist = vsq(vsub(inter, center));
            cur_nv = -1;
        } else {
            curdist = vsq(vsub(center, v0));
            double vdist = vsq(vsub(center, v1));
            if (vdist < curdist) { curdist = vdist; cur_nv = pe[se * 2 + 1]; }
            else cur_nv = pe[se * 2];
        }
        if (nv_index != -1 && nv_index == cur_nv) {
            double a1d = vdot(ne_normal, center);
            double a2d = vdot(n0, center);
            if ((LESS_EQU_ZERO(a1d) && GRE_EQU_ZERO(a2d)) || (GRE_EQU_ZERO(a1d) && LESS_EQU_ZERO(a2d))) {
                int another;
                if (pe[ne * 2] == nv_index) another = pe[ne * 2 + 1];
                else another = pe[ne * 2];
                v3 a0 = vsub(pv[another], pv[nv_index]);
                if (pe[se * 2] == nv_index) another = pe[se * 2 + 1];
                else another = pe[se * 2];
                v3 a1 = vsub(pv[another], pv[nv_index]);
                v3 bb = vsub(center, pv[nv_index]);
                a0 = vnorm(a0);
                a1 = vnorm(a1);
                a1d = vdot(a0, bb);
                a2d = vdot(a1, bb);
                if (a2d > a1d) {
                    mindist = curdist; ne = se; ne_normal = n0; nv_index = cur_nv; sl = i;
                    continue;
                }
            }
        }
        if (curdist < mindist) {
            mindist = curdist; ne = se; ne_normal = n0; nv_index = cur_nv; sl = i;
        }
    }
    if (slot) *slot = sl;
    if (nvo) *nvo = nv_index;
    if (md) *md = mindist;
    double temp = vdot(center, ne_normal);
    return GRE_ZERO(temp) ? CELL_IN : CELL_OUT;
}

static int cp_prop(const og_ctx* c, int cell, v3 center) {
    const og_cell* cc = &c->c[cell];
    return cp_prop_refs(c->pv, c->pe, c->refs + cc->rstart, cc->rnum, center, NULL, NULL, NULL);
}

/* Provenance of a cell's type (test hook; no counterpart in the
 * reference): which classification rule set it, from which cell, across
 * which border. */
enum { OG_PROV_NONE = 0, OG_PROV_NE = 1, OG_PROV_BOTTOM_DE = 2, OG_PROV_IE = 3, OG_PROV_DE_COPY = 4,
       OG_PROV_DE_PARITY = 5, OG_PROV_TOP = 6 };
static inline void prov_set(og_ctx* c, int cell, int rule, int src, int border) {
    c->prov[cell] = rule; c->prov_src[cell] = src; c->prov_b[cell] = border;
}

/* Parity of crossings from the centre of `from` to the midpoint of border
 * `mb` over the refs of `from` (:1094-1119, :1332-1357). */
static int parity_type(const og_ctx* c, int from, int mb) {
    v3 m = vnorm(vmul(vadd(c->v[c->b[mb].v[0]], c->v[c->b[mb].v[1]]), 0.5));
    v3 center = cell_center_sum(c, from);
    int cnt = 0;
    const og_cell* fc = &c->c[from];
    for (int i = 0; i < fc->rnum; i++) {
        int se = c->refs[fc->rstart + i];
        if (is_inter_seg(c->pv[c->pe[se * 2]], c->pv[c->pe[se * 2 + 1]], center, m)) cnt++;
    }
    return (cnt % 2 == 0) ? c->c[from].type : 3 - c->c[from].type;
}

/* kernel_Classify_setIECell (:1366-1392) repeated to a fixpoint; each pass
 * reads a snapshot of the types.  Returns the UNDEFINED leaves left after
 * the first pass. */
static int64_t ie_fixpoint(og_ctx* c, int lv, int32_t* snap) {
    og_level* G = &c->g[lv];
    int64_t after_first = -1;
    for (int round = 0;; round++) {
        for (int t = 0; t < G->cnum; t++) snap[t] = c->c[G->cstart + t].type;
        int changed = 0;
#pragma omp parallel for schedule(static) reduction(+ : changed)
        for (int t = 0; t < G->cnum; t++) {
            og_cell* cc = &c->c[G->cstart + t];
            if (snap[t] != CELL_UNDEFINED || cc->subcell != -1) continue;
            for (int i = 0; i < cc->bnum; i++) {
                int pe_ = c->b[cc->bstart + i].partner;
                if (pe_ == -1) continue;
                int pc = c->b[pe_].parent;
                int pt = snap[pc - G->cstart];
                if (c->c[pc].rnum == 0 && pt != CELL_UNDEFINED) {
                    cc->type = pt; prov_set(c, G->cstart + t, OG_PROV_IE, pc, pe_); changed++; break;
                }
            }
        }
        if (round == 0) {
            after_first = 0;
            for (int t = 0; t < G->cnum; t++)
                if (c->c[G->cstart + t].subcell == -1 && c->c[G->cstart + t].type == CELL_UNDEFINED) after_first++;
        }
        if (!changed) break;
    }
    return after_first;
}

int og_classify(og_ctx* c) {
    int L = c->gnum - 1;
    og_level* GL = &c->g[L];
    int64_t maxc = 0;
    for (int l = 0; l < c->gnum; l++) if (c->g[l].cnum > maxc) maxc = c->g[l].cnum;
    int32_t* snap = (int32_t*)malloc(sizeof(int32_t) * (size_t)(maxc + 1));
    int64_t und_snapshot = 0;
    {
        const int64_t C = (int64_t)GL->cstart + GL->cnum;
        if (c->provcap < C) {
            free(c->prov); free(c->prov_src); free(c->prov_b);
            c->prov = (int32_t*)malloc(sizeof(int32_t) * (size_t)C);
            c->prov_src = (int32_t*)malloc(sizeof(int32_t) * (size_t)C);
            c->prov_b = (int32_t*)malloc(sizeof(int32_t) * (size_t)C);
            c->provcap = C;
        }
        for (int64_t k = 0; k < C; k++) c->prov[k] = OG_PROV_NONE, c->prov_src[k] = -1, c->prov_b[k] = -1;
    }
    /* kernel_Classify_setBottomNECell :1046-1071 */
#pragma omp parallel for schedule(dynamic, 64)
    for (int t = 0; t < GL->cnum; t++) {
        int cid = GL->cstart + t;
        if (c->c[cid].rnum != 0) {
            c->c[cid].type = cp_prop(c, cid, cell_center_sum(c, cid));
            prov_set(c, cid, OG_PROV_NE, -1, -1);
        }
    }
    /* kernel_Classify_setBottomDECell :1074-1124, snapshot of types */
    for (int t = 0; t < GL->cnum; t++) snap[t] = c->c[GL->cstart + t].type;
#pragma omp parallel for schedule(dynamic, 64)
    for (int t = 0; t < GL->cnum; t++) {
        int cid = GL->cstart + t;
        if (snap[t] != CELL_UNDEFINED) continue;
        og_cell* cc = &c->c[cid];
        for (int i = 0; i < cc->bnum; i++) {
            int pe_ = c->b[cc->bstart + i].partner;
            if (pe_ == -1) continue;
            int nb = c->b[pe_].parent;
            if (snap[nb - GL->cstart] == CELL_UNDEFINED) continue;
            cc->type = parity_type(c, nb, pe_);
            prov_set(c, cid, OG_PROV_BOTTOM_DE, nb, pe_);
        }
    }
    und_snapshot += ie_fixpoint(c, L, snap);
    /* upper levels :2025-2048 */
    for (int lv = L - 1; lv >= 0; lv--) {
        og_level* G = &c->g[lv];
        /* kernel_Classify_setDECell :1292-1363 */
#pragma omp parallel for schedule(dynamic, 64)
        for (int t = 0; t < G->cnum; t++) {
            int cid = G->cstart + t;
            og_cell* cc = &c->c[cid];
            if (!(cc->rnum == 0 && cc->type == CELL_UNDEFINED)) continue;
            int pe_ = -1, pc = -1;
            for (int i = 0; i < cc->bnum; i++) {
                pe_ = c->b[cc->bstart + i].partner;
                if (pe_ != -1 && c->c[c->b[pe_].parent].subcell != -1) { pc = c->b[pe_].parent; break; }
            }
            if (pc == -1) continue;
            int target = c->b[pe_].subcell;
            if (target < 0) continue;
            while (c->c[target].subcell != -1) target = c->c[target].subcell;
            if (c->c[target].rnum == 0) {
                cc->type = c->c[target].type;
                prov_set(c, cid, OG_PROV_DE_COPY, target, pe_);
            } else {
                cc->type = parity_type(c, target, pe_);
                prov_set(c, cid, OG_PROV_DE_PARITY, target, pe_);
            }
        }
        und_snapshot += ie_fixpoint(c, lv, snap);
    }
    /* kernel_Classify_setTopECell :1395-1419 as a fixpoint over snapshots */
    og_level* G0 = &c->g[0];
    for (int round = 0;; round++) {
        for (int t = 0; t < G0->cnum; t++) snap[t] = c->c[t].type;
        int changed = 0;
        for (int t = 0; t < G0->cnum; t++) {
            og_cell* cc = &c->c[t];
            if (!(cc->rnum == 0 && snap[t] == CELL_UNDEFINED)) continue;
            for (int i = 0; i < cc->bnum; i++) {
                int pe_ = c->b[cc->bstart + i].partner;
                if (pe_ != -1 && snap[c->b[pe_].parent] != CELL_UNDEFINED) {
                    cc->type = snap[c->b[pe_].parent];
                    prov_set(c, t, OG_PROV_TOP, c->b[pe_].parent, pe_);
                    changed++; break;
                }
            }
        }
        if (!changed) break;
    }
    int64_t und_final = 0;
    for (int l = 0; l < c->gnum; l++) {
        for (int t = 0; t < c->g[l].cnum; t++) {
            og_cell* cc = &c->c[c->g[l].cstart + t];
            if (cc->subcell == -1 && cc->type == CELL_UNDEFINED) und_final++;
        }
    }
    c->diag.undefined_leaves_snapshot = und_snapshot;
    c->diag.undefined_leaves_final = und_final;
    free(snap);
    c->built = 1;
    return 0;
}

int og_build(og_ctx* c, int32_t max_depth, int32_t auto_depth) {
    if (og_build_create_only(c, max_depth, auto_depth) < 0) return -1;
    return og_classify(c);
}

/* ---------------------------------------------------------------------- */
/* Query: kernel_PISP, source/GPISPAHHG.cu:1422-1504.                      */
/* ---------------------------------------------------------------------- */
/* First-match descent to the leaf holding p (-1 = no leaf), counting the
 * border planes evaluated. */
static int locate_leaf(const og_ctx* c, v3 p, int64_t* tests) {
    int cur = -1;
    for (int i = 0; i < 42; i++) {
        cur = inside_cell(c, p, i, tests);
        if (cur != -1) break;
    }
    if (cur == -1) return -1;
    while (c->c[cur].subcell != -1) {
        int found = 0;
        int ret = inside_cell(c, p, c->c[cur].subcell, tests);
        if (ret != -1) { cur = ret; continue; }
        const og_cell* cc = &c->c[cur];
        for (int i = 0; i < cc->bnum; i++) {
            int tc = c->b[cc->bstart + i].subcell;
            if (tc == -1) continue;
            ret = inside_cell(c, p, tc, tests);
            if (ret != -1) { cur = ret; found = 1; break; }
        }
        if (!found) return -1;
    }
    return cur;
}

static int32_t query_one(const og_ctx* c, v3 p, og_qcount* q) {
    int64_t tests = 0;
    int cur = locate_leaf(c, p, &tests);
    if (cur == -1) { q->plane_tests += tests; q->unlocated++; return 0; }
    q->plane_tests += tests;
    const og_cell* lc = &c->c[cur];
    if (lc->rnum == 0) return lc->type;
    v3 center = cell_center_sum(c, cur);
    int cnt = 0;
    for (int i = 0; i < lc->rnum; i++) {
        int se = c->refs[lc->rstart + i];
        if (is_ee_inter(c->pv[c->pe[se * 2]], c->pv[c->pe[se * 2 + 1]], center, p)) cnt++;
    }
    q->boundary_hits++;
    q->ref_tests += lc->rnum;
    q->bnum_sum += lc->bnum;
    return (cnt % 2 == 0) ? lc->type : 3 - lc->type;
}

void og_query_xyz(og_ctx* c, const double* xyz, int64_t n, int32_t* labels, int32_t nthreads,
                  og_qcount* counts) {
    og_qcount tot = {0, 0, 0, 0, 0, 0};
    int64_t pt = 0, bh = 0, rt = 0, bs = 0, un = 0, ins = 0;
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nt) reduction(+ : pt, bh, rt, bs, un, ins)
#endif
    for (int64_t i = 0; i < n; i++) {
        og_qcount q = {0, 0, 0, 0, 0, 0};
        int32_t lab = query_one(c, mk(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]), &q);
        labels[i] = lab;
        pt += q.plane_tests; bh += q.boundary_hits; rt += q.ref_tests; bs += q.bnum_sum; un += q.unlocated;
        ins += (lab == CELL_IN);
    }
    (void)nthreads;
    tot.plane_tests = pt; tot.boundary_hits = bh; tot.ref_tests = rt; tot.bnum_sum = bs;
    tot.unlocated = un; tot.inside = ins;
    if (counts) *counts = tot;
}

void og_query_leaf(og_ctx* c, const double* xyz, int64_t n, int32_t* leaf) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4096)
#endif
    for (int64_t i = 0; i < n; i++) {
        int64_t t = 0;
        leaf[i] = locate_leaf(c, mk(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]), &t);
    }
}

/* ---------------------------------------------------------------------- */
/* Export.                                                                 */
/* ---------------------------------------------------------------------- */
void og_counts(og_ctx* c, int64_t* out) {
    int L = c->gnum - 1;
    if (L < 0) { out[0] = out[1] = out[2] = out[3] = out[4] = 0; return; }
    out[0] = c->gnum;
    out[1] = (int64_t)c->g[L].vstart + c->g[L].vnum;
    out[2] = (int64_t)c->g[L].bstart + c->g[L].bnum;
    out[3] = (int64_t)c->g[L].cstart + c->g[L].cnum;
    out[4] = c->nref;
}

float og_gamma(og_ctx* c) { return c->gamma; }

void og_levels(og_ctx* c, og_level* out) {
    for (int l = 0; l < c->gnum; l++) out[l] = c->g[l];
}

void og_diagnostics(og_ctx* c, og_diag* out) { *out = c->diag; }

void og_export(og_ctx* c, double* v, int32_t* cells, int32_t* bi, double* bn, int32_t* refs) {
    int64_t cnt[5];
    og_counts(c, cnt);
    if (v) for (int64_t i = 0; i < cnt[1]; i++) { v[3 * i] = c->v[i].x; v[3 * i + 1] = c->v[i].y; v[3 * i + 2] = c->v[i].z; }
    if (cells) for (int64_t i = 0; i < cnt[3]; i++) {
        og_cell* cc = &c->c[i];
        int32_t* o = cells + 7 * i;
        o[0] = cc->bstart; o[1] = cc->bnum; o[2] = cc->rstart; o[3] = cc->rnum;
        o[4] = cc->upcell; o[5] = cc->subcell; o[6] = cc->type;
    }
    if (bi) for (int64_t i = 0; i < cnt[2]; i++) {
        og_border* b = &c->b[i];
        int32_t* o = bi + 6 * i;
        o[0] = b->parent; o[1] = b->partner; o[2] = b->v[0]; o[3] = b->v[1]; o[4] = b->subcell; o[5] = b->id;
        if (bn) { bn[3 * i] = b->n[0]; bn[3 * i + 1] = b->n[1]; bn[3 * i + 2] = b->n[2]; }
    }
    if (refs) memcpy(refs, c->refs, sizeof(int32_t) * (size_t)cnt[4]);
}

/* Provenance of every cell's type (rule, source cell, border crossed; -1 =
 * none) after classification. */
void og_provenance(og_ctx* c, int32_t* rule, int32_t* src, int32_t* border) {
    int64_t cnt[5];
    og_counts(c, cnt);
    for (int64_t i = 0; i < cnt[3]; i++) {
        rule[i] = c->prov ? c->prov[i] : OG_PROV_NONE;
        src[i] = c->prov ? c->prov_src[i] : -1;
        border[i] = c->prov ? c->prov_b[i] : -1;
    }
}

/* Overwrites one cell's type: plants a mistake for tests of the checkers. */
void og_debug_set_type(og_ctx* c, int32_t cell, int32_t type) {
    int64_t cnt[5];
    og_counts(c, cnt);
    if (cell >= 0 && cell < cnt[3]) c->c[cell].type = type;
}

/* ---------------------------------------------------------------------- */
/* Known-answer test hooks.                                                */
/* ---------------------------------------------------------------------- */
void og_kat_normalize(const double* in, double* out) {
    v3 r = vnorm(mk(in[0], in[1], in[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void og_kat_cross(const double* a, const double* b, double* out) {
    v3 r = vcross(mk(a[0], a[1], a[2]), mk(b[0], b[1], b[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
int32_t og_kat_is_ee_inter(const double* p0, const double* p1, const double* v0, const double* v1) {
    return is_ee_inter(mk(p0[0], p0[1], p0[2]), mk(p1[0], p1[1], p1[2]), mk(v0[0], v0[1], v0[2]), mk(v1[0], v1[1], v1[2]));
}
int32_t og_kat_edge_border_inter(const double* p0, const double* p1, const double* v0, const double* v1, int32_t* reserved) {
    int r = -1;
    int t = edge_border_inter(mk(p0[0], p0[1], p0[2]), mk(p1[0], p1[1], p1[2]), mk(v0[0], v0[1], v0[2]), mk(v1[0], v1[1], v1[2]), &r);
    *reserved = r;
    return t;
}

int32_t og_kat_cp_prop(const double* pv, const int32_t* pe, const int32_t* refs, int32_t rn, const double* center,
                       int32_t* slot, int32_t* nv, double* mindist) {
    int sl = -1, v = -1;
    const int t = cp_prop_refs((const v3*)pv, pe, refs, rn, mk(center[0], center[1], center[2]), &sl, &v, mindist);
    *slot = sl;
    *nv = v;
    return t;
}
