// gahhg_host.cpp -- host-side input handling and the level-0 grid.
//
// Polygon TXT format and ring handling: source/GPISPPolygon.cu:32-278.
// Orientation rule: source/GPISPPolygon.cu:280-313.
// LL -> XYZ (y up): source/GPISPPolygon.cuh:21-30; query points go through
// the FakeLLVertex round trip of source/GPISPPolygon.cuh:72,82-92.
// Lon/lat conversions use correctly rounded sin and cos (gahhg_trig.h, shared
// with the device loader); the level-0 grid's constants use the C library.
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <string>
#include <vector>

#include "gahhg_internal.h"
#include "gahhg_decimal.h"
#include "gahhg_trig.h"

#pragma clang fp contract(off)

namespace gahhg {

// Correctly rounded sin and cos (gahhg_trig.h), the same function as the
// device conversion, so both give identical bits.
static V3 ll_to_xyz(double lo, double la) {
  double a = la * kPi / 180.0;
  double o = lo * kPi / 180.0;
  double sa, ca, so, co;
  trig::sincos_cr(a, &sa, &ca);
  trig::sincos_cr(o, &so, &co);
  return v3(ca * so, sa, ca * co);
}

void query_lonlat_to_xyz(const double* lon, const double* lat, int64_t n, double* xyz) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    double fakela = lat[i] + 90;
    double realla = fakela > 180 ? 270 - fakela : fakela - 90;
    store3(xyz + 3 * i, ll_to_xyz(lon[i], realla));
  }
}

// true when the ring (unique vertices) is read as clockwise: at the first
// vertex of maximal latitude, the next vertex is not strictly left of the arc
// arriving there.
static bool ring_is_clockwise(const double* lo, const double* la, int n) {
  double best = -FLT_MAX;
  int m = 0;
  for (int i = 0; i < n; i++) {
    if (la[i] > best) {
      best = la[i];
      m = i;
    }
  }
  int prev = (m == 0) ? n - 1 : m - 1;
  int next = (m == n - 1) ? 0 : m + 1;
  V3 a0 = ll_to_xyz(lo[prev], la[prev]);
  V3 a1 = ll_to_xyz(lo[m], la[m]);
  V3 b = ll_to_xyz(lo[next], la[next]);
  V3 nrm = unit(cross(a0, sub(a1, a0)));
  return !(dot(nrm, b) > 0);
}

// xyz, when given, holds every vertex's point (parallel to lo/la), as the
// device loader computed it with the same conversion.
static int assemble(const double* lo, const double* la, const double* xyz, const std::vector<Ring>& rings,
                    Polygon* out, std::string* err) {
  int64_t total = 0;
  for (const Ring& r : rings) total += r.n;
  if (total <= 0 || total > (1LL << 30)) {
    *err = "polygon: vertex count out of range";
    return GAHHG_E_LIMIT;
  }
  out->N = (int32_t)total;
  out->flips = 0;
  out->pv.assign(3 * total, 0.0);
  out->pe.assign(2 * total, 0);
  int64_t k = 0;
  for (const Ring& r : rings) {
    double* pv = &out->pv[3 * k];
    if (xyz) {
      memcpy(pv, xyz + 3 * r.start, 24 * (size_t)r.n);
    } else {
#pragma omp parallel for schedule(static) if (r.n > 65536)
      for (int j = 0; j < r.n; j++) store3(pv + 3 * j, ll_to_xyz(lo[r.start + j], la[r.start + j]));
    }
    k += r.n;
  }
  // Edge list: forward rings link i -> i+1; reversed rings walk the indices
  // downwards; the last edge closes the ring.
  int32_t idx = 0;
  for (size_t ri = 0; ri < rings.size(); ri++) {
    const Ring& r = rings[ri];
    bool cw = ring_is_clockwise(lo + r.start, la + r.start, r.n);
    bool reverse = (ri == 0) ? cw : !cw;  // outer must be CCW, holes CW
    if (reverse && ri < 31) out->flips |= 1 << ri;
    int32_t s = idx;
    for (int j = 0; j < r.n; j++) {
      int32_t from = reverse ? s + (r.n - 1 - j) : s + j;
      int32_t to = reverse ? from - 1 : from + 1;
      if (j == r.n - 1) to = reverse ? s + r.n - 1 : s;
      out->pe[2 * idx] = from;
      out->pe[2 * idx + 1] = to;
      idx++;
    }
  }
  return GAHHG_OK;
}

void lonlat_to_xyz(const double* lon, const double* lat, int64_t n, double* xyz) {
  for (int64_t i = 0; i < n; i++) store3(xyz + 3 * i, ll_to_xyz(lon[i], lat[i]));
}

int polygon_from_lonlat(const double* lon, const double* lat, const int64_t* ring_off,
                        int32_t n_rings, Polygon* out, std::string* err) {
  return polygon_from_lonlat_xyz(lon, lat, nullptr, ring_off, n_rings, out, err);
}

int polygon_from_lonlat_xyz(const double* lon, const double* lat, const double* xyz, const int64_t* ring_off,
                            int32_t n_rings, Polygon* out, std::string* err) {
  if (n_rings < 1 || !lon || !lat || !ring_off) {
    *err = "polygon: need at least one ring";
    return GAHHG_E_ARG;
  }
  std::vector<Ring> rings;
  for (int r = 0; r < n_rings; r++) {
    int64_t cnt = ring_off[r + 1] - ring_off[r];
    if (cnt < 4) {
      *err = "polygon: ring " + std::to_string(r) + ": number of edges should be greater than 3";
      return GAHHG_E_FORMAT;
    }
    rings.push_back(Ring{ring_off[r], (int32_t)(cnt - 1)});
  }
  return assemble(lon, lat, xyz, rings, out, err);
}

namespace {
// Lines of a whole file read at once: a line ends at '\n' (not stored); a
// last line without one counts when it is not empty.
struct LineReader {
  std::string buf;
  size_t pos = 0;
  const char* line = nullptr;
  size_t len = 0;
  int64_t lineno = 0;
  bool open(const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    char tmp[1 << 16];
    size_t got;
    while ((got = fread(tmp, 1, sizeof tmp, f)) > 0) buf.append(tmp, got);
    fclose(f);
    return true;
  }
  bool next() {
    if (pos >= buf.size()) return false;
    const char* b = buf.data() + pos;
    const void* nl = memchr(b, '\n', buf.size() - pos);
    len = nl ? (size_t)((const char*)nl - b) : buf.size() - pos;
    line = b;
    pos += len + (nl ? 1 : 0);
    lineno++;
    return true;
  }
  std::string str() const { return std::string(line, len); }
};

// "%lf %lf" as sscanf reads it: the correctly rounded decimal fast path
// (gahhg_decimal.h, equal to strtod on its grammar), sscanf for the rest.
bool two_doubles(const LineReader& rd, double* x, double* y) {
  const uint8_t* b = (const uint8_t*)rd.line;
  const int64_t e = (int64_t)rd.len;
  int64_t i = dec::skip_space(b, 0, e);
  if (i < e && dec::parse_decimal(b, i, e, *x)) {
    i = dec::skip_space(b, i, e);
    if (i < e && dec::parse_decimal(b, i, e, *y)) return true;
  }
  return sscanf(rd.str().c_str(), "%lf %lf", x, y) == 2;
}
}  // namespace

int parse_polygon_txt(const char* path, PolygonText* out, std::string* err) {
  LineReader rd;
  if (!rd.open(path)) {
    *err = std::string("Can not open ") + path;
    return GAHHG_E_IO;
  }
  auto fail = [&](const char* what) {
    *err = std::string(path) + ": line " + std::to_string(rd.lineno) + ": " + what;
    return GAHHG_E_FORMAT;
  };
  int holes = 0;
  if (!rd.next() || sscanf(rd.str().c_str(), "POLYGON 1 %d", &holes) != 1 || holes < 0)
    return fail("wrong format");
  std::vector<double>& lo = out->lo;
  std::vector<double>& la = out->la;
  lo.clear();
  la.clear();
  out->rings.clear();
  for (int r = 0; r <= holes; r++) {
    int cnt = 0, tmp = 0;
    bool ok = rd.next();
    if (r == 0)
      ok = ok && sscanf(rd.str().c_str(), "0 %d", &cnt) == 1;
    else
      ok = ok && sscanf(rd.str().c_str(), "%d %d", &tmp, &cnt) == 2;
    if (!ok) return fail("wrong format");
    if (cnt < 4) return fail("number of edges should be greater than 3");
    out->rings.push_back(Ring{(int64_t)lo.size(), cnt - 1});
    for (int j = 0; j < cnt; j++) {
      double x, y;
      if (!rd.next() || !two_doubles(rd, &x, &y)) return fail("wrong format");
      if (j + 1 < cnt) {  // the closing duplicate is not stored
        lo.push_back(x);
        la.push_back(y);
      }
    }
  }
  return GAHHG_OK;
}

int polygon_from_text(const PolygonText& t, const double* xyz, Polygon* out, std::string* err) {
  return assemble(t.lo.data(), t.la.data(), xyz, t.rings, out, err);
}

int load_polygon_txt(const char* path, Polygon* out, std::string* err) {
  PolygonText t;
  const int rc = parse_polygon_txt(path, &t, err);
  return rc ? rc : polygon_from_text(t, nullptr, out, err);
}

int read_query_csv(const char* path, std::vector<double>* lon, std::vector<double>* lat,
                   std::string* err) {
  LineReader rd;
  if (!rd.open(path)) {
    *err = std::string("Can not open ") + path;
    return GAHHG_E_IO;
  }
  while (rd.next()) {
    double x, y;
    // "%lf,%lf": the decimal fast path, sscanf for what it leaves
    const uint8_t* b = (const uint8_t*)rd.line;
    const int64_t e = (int64_t)rd.len;
    int64_t i = dec::skip_space(b, 0, e);
    bool ok = i < e && dec::parse_decimal(b, i, e, x) && i < e && b[i] == ',';
    if (ok) {
      i = dec::skip_space(b, i + 1, e);
      ok = i < e && dec::parse_decimal(b, i, e, y);
    }
    if (!ok && sscanf(rd.str().c_str(), "%lf,%lf", &x, &y) != 2) {
      *err = std::string(path) + ": line " + std::to_string(rd.lineno) + ": wrong format";
      return GAHHG_E_FORMAT;
    }
    lon->push_back(x);
    lat->push_back(y);
  }
  return GAHHG_OK;
}

// Lune/Ketzner-format files.  The reference ships copies of its polygons and
// query points in the format of the comparators it was measured against
// (README.md:5, test/data_for_Lune): one "colatitude,longitude" pair per
// line, radians, the longitude counted from 180 deg West, so that
//   lat = 90 - deg(colatitude),  lon = (deg(longitude) mod 360) - 180
// with deg(x) = x * (180 / pi) and the mod of Python / numpy (the result
// takes the divisor's sign: fmod, plus 360 when fmod is negative).  Under
// this rule the format's Square_72, Radiometer and Tennessee are the TXT
// files' polygons.
namespace {
constexpr double kDegPerRad = 180.0 / 3.14159265358979323846264338327950288;

double lune_lat(double colat) { return 90.0 - colat * kDegPerRad; }

double lune_lon(double lon) {
  double m = fmod(lon * kDegPerRad, 360.0);
  if (m != 0.0) {
    if (m < 0.0) m += 360.0;
  } else {
    m = 0.0;  // +0 whatever the sign of the dividend
  }
  return m - 180.0;
}
}  // namespace

int read_lune_points(const char* path, std::vector<double>* lon, std::vector<double>* lat, std::string* err) {
  std::vector<double> colat, lo;
  const int rc = read_query_csv(path, &colat, &lo, err);  // the same "%lf,%lf" lines
  if (rc) return rc;
  lon->resize(lo.size());
  lat->resize(lo.size());
  for (size_t i = 0; i < lo.size(); i++) {
    (*lon)[i] = lune_lon(lo[i]);
    (*lat)[i] = lune_lat(colat[i]);
  }
  return GAHHG_OK;
}

// A polygon file is one ring, closed by repeating its first vertex (the
// format has no holes), then the point the comparators use as known inside.
int parse_polygon_lune(const char* path, PolygonText* out, double* ref_lonlat, bool* has_ref, std::string* err) {
  std::vector<double> colat, lo;
  const int rc = read_query_csv(path, &colat, &lo, err);
  if (rc) return rc;
  size_t close = 0;
  for (size_t i = 1; i < lo.size() && close == 0; i++)
    if (colat[i] == colat[0] && lo[i] == lo[0]) close = i;
  if (close == 0) {
    *err = std::string(path) + ": the ring is not closed";
    return GAHHG_E_FORMAT;
  }
  if (close < 3) {
    *err = std::string(path) + ": a ring needs at least 3 vertices";
    return GAHHG_E_FORMAT;
  }
  out->lo.resize(close);
  out->la.resize(close);
  for (size_t i = 0; i < close; i++) {
    out->lo[i] = lune_lon(lo[i]);
    out->la[i] = lune_lat(colat[i]);
  }
  out->rings.assign(1, Ring{0, (int32_t)close});
  *has_ref = close + 1 < lo.size();
  if (*has_ref) {
    ref_lonlat[0] = lune_lon(lo[close + 1]);
    ref_lonlat[1] = lune_lat(colat[close + 1]);
  }
  return GAHHG_OK;
}

// Level-0 grid.  The reference ships it only inside a compiled library
// (initgrid0, declared at source/common.h:162-201, called at
// source/GPISPAHHG.cu:1962); its shape is 42 cells (12 pentagons + 30
// hexagons), 80 vertices and 240 directed borders, each cell's borders CCW
// seen from outside so that n = normalize(v0 x (v1 - v0)) points inwards.
// This grid is the dual of the frequency-2 geodesic icosahedron: cell centres
// are the icosahedron's 12 vertices and 30 edge midpoints, cell corners the
// normalised centroids of the 80 sub-triangles.  Cell c's borders occupy
// [bstart, bstart + bnum) with id = position in the cell.
void build_grid0(double* v80, int32_t* bi, double* bn, int32_t* ci) {
  const double pi = 3.14159265358979323846;
  std::array<V3, 12> ico;
  const double lat = atan(0.5);
  ico[0] = v3(0.0, 1.0, 0.0);
  for (int k = 0; k < 5; k++) {
    double lon_up = (double)k * (2.0 * pi / 5.0);
    double lon_dn = ((double)k + 0.5) * (2.0 * pi / 5.0);
    ico[1 + k] = unit(v3(cos(lat) * sin(lon_up), sin(lat), cos(lat) * cos(lon_up)));
    ico[6 + k] = unit(v3(cos(-lat) * sin(lon_dn), sin(-lat), cos(-lat) * cos(lon_dn)));
  }
  ico[11] = v3(0.0, -1.0, 0.0);

  std::array<std::array<int, 3>, 20> face;
  for (int k = 0; k < 5; k++) {
    int k1 = (k + 1) % 5;
    face[k] = {0, 1 + k, 1 + k1};
    face[5 + k] = {1 + k, 6 + k, 1 + k1};
    face[10 + k] = {1 + k1, 6 + k, 6 + k1};
    face[15 + k] = {6 + k, 11, 6 + k1};
  }
  for (auto& f : face) {  // counter-clockwise seen from outside
    V3 a = ico[f[0]], b = ico[f[1]], c = ico[f[2]];
    if (dot(cross(sub(b, a), sub(c, a)), add(add(a, b), c)) < 0) std::swap(f[1], f[2]);
  }

  // centres: 12 vertices, then 30 edge midpoints in order of first use
  std::vector<V3> centre(ico.begin(), ico.end());
  std::vector<std::pair<int, int>> edges;
  std::array<std::array<int, 3>, 20> mid;  // centre id of edge (j, j+1) of face f
  for (int f = 0; f < 20; f++) {
    for (int j = 0; j < 3; j++) {
      int a = face[f][j], b = face[f][(j + 1) % 3];
      std::pair<int, int> key(std::min(a, b), std::max(a, b));
      auto it = std::find(edges.begin(), edges.end(), key);
      int e;
      if (it == edges.end()) {
        e = (int)edges.size();
        edges.push_back(key);
        centre.push_back(unit(add(ico[key.first], ico[key.second])));
      } else {
        e = (int)(it - edges.begin());
      }
      mid[f][j] = 12 + e;
    }
  }

  std::array<std::array<int, 3>, 80> tri;
  for (int f = 0; f < 20; f++) {
    int a = face[f][0], b = face[f][1], c = face[f][2];
    int ab = mid[f][0], bc = mid[f][1], ca = mid[f][2];
    tri[4 * f + 0] = {a, ab, ca};
    tri[4 * f + 1] = {ab, b, bc};
    tri[4 * f + 2] = {ca, bc, c};
    tri[4 * f + 3] = {ab, bc, ca};
  }
  for (int t = 0; t < 80; t++)
    store3(v80 + 3 * t, unit(add(add(centre[tri[t][0]], centre[tri[t][1]]), centre[tri[t][2]])));

  // Cell x: its sub-triangles in CCW order around x, starting from the
  // lowest-numbered one; the corner of triangle (x, y, z) is followed by the
  // triangle (x, z, w).
  int bpos = 0;
  for (int x = 0; x < 42; x++) {
    auto rot = [&](int t, int* y, int* z) {
      for (int j = 0; j < 3; j++)
        if (tri[t][j] == x) {
          *y = tri[t][(j + 1) % 3];
          *z = tri[t][(j + 2) % 3];
          return true;
        }
      return false;
    };
    std::vector<int> ring;
    int first = -1, y = -1, z = -1;
    for (int t = 0; t < 80 && first < 0; t++)
      if (rot(t, &y, &z)) first = t;
    ring.push_back(first);
    for (;;) {
      int nxt = -1, ny = -1, nz = -1;
      for (int t = 0; t < 80; t++)
        if (rot(t, &ny, &nz) && ny == z) {
          nxt = t;
          break;
        }
      if (nxt < 0 || nxt == first) break;
      ring.push_back(nxt);
      z = nz;
    }
    int n = (int)ring.size();
    int32_t* cc = ci + 7 * x;
    cc[0] = bpos;  // bstart
    cc[1] = n;     // bnum
    cc[2] = -1;    // rstart
    cc[3] = 0;     // rnum
    cc[4] = -1;    // upcell
    cc[5] = -1;    // subcell
    cc[6] = kUndefined;
    for (int i = 0; i < n; i++) {
      int32_t* b = bi + 6 * (bpos + i);
      b[0] = x;   // parent
      b[1] = -1;  // partner (below)
      b[2] = ring[i];
      b[3] = ring[(i + 1) % n];
      b[4] = -1;  // subcell
      b[5] = i;   // id
      V3 n3 = arc_normal(load3(v80 + 3 * b[2]), load3(v80 + 3 * b[3]));
      store3(bn + 3 * (bpos + i), n3);
    }
    bpos += n;
  }
  for (int i = 0; i < 240; i++)
    for (int j = 0; j < 240; j++)
      if (bi[6 * i + 2] == bi[6 * j + 3] && bi[6 * i + 3] == bi[6 * j + 2]) {
        bi[6 * i + 1] = j;
        break;
      }
}

}  // namespace gahhg
