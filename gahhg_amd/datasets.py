"""Seeded synthetic polygons and query sets.

Workloads (sizes in edges N and queries Q):
  star_1k      one star polygon, 1,000 + 200-vertex hole, 1M uniform queries
  fractal_100k 100k-vertex fractal ring + 5 holes of 2k, 10M queries
               (50% uniform, 50% within ~0.01 deg of an edge)
  fractal_1m   1M-vertex fractal ring (radius 40 deg, centre lat 60 deg, so it
               encloses the North Pole) + 3 holes of 50k, 100M queries
               (70% uniform, 30% near-edge)
  batch_256    256 polygons, log-uniform vertex counts in [16, 200k]
  fractal_512  512 polygons of the fractal_1m generator scaled to 4k-64k
               outer vertices, 32M queries (fractal_1m's mix per polygon)

Rings are produced as lon/lat degrees with the closing vertex repeated, in
the layout of the reference's polygon TXT files (README.md:41-69); the
library applies the reference's orientation rule on load.  Generation uses
numpy's PCG64 streams (seed per workload, queries seed + 1000).
"""
import math

import numpy as np

DEG = math.pi / 180.0


def destination(lon0, lat0, az, r):
    """Point at angular distance r (rad) and azimuth az (rad) from (lon0, lat0)
    in degrees; returns lon/lat degrees (lon wrapped to [-180, 180))."""
    p0 = lat0 * DEG
    l0 = lon0 * DEG
    lat = np.arcsin(np.sin(p0) * np.cos(r) + np.cos(p0) * np.sin(r) * np.cos(az))
    lon = l0 + np.arctan2(np.sin(az) * np.sin(r) * np.cos(p0), np.cos(r) - np.sin(p0) * np.sin(lat))
    lon = (np.degrees(lon) + 180.0) % 360.0 - 180.0
    return lon, np.degrees(lat)


def star_ring(rng, lon0, lat0, n, radius_deg, jitter=0.3):
    az = 2.0 * np.pi * np.arange(n) / n
    r = radius_deg * DEG * (1.0 + jitter * rng.uniform(-1.0, 1.0, n))
    return destination(lon0, lat0, az, r)


def _xyz(lon, lat):
    lo, la = np.radians(lon), np.radians(lat)
    return np.stack([np.cos(la) * np.sin(lo), np.sin(la), np.cos(la) * np.cos(lo)], axis=-1)


def _lonlat(p):
    p = p / np.linalg.norm(p, axis=-1, keepdims=True)
    lat = np.degrees(np.arcsin(np.clip(p[..., 1], -1.0, 1.0)))
    lon = np.degrees(np.arctan2(p[..., 0], p[..., 2]))
    return lon, lat


def fractal_ring(rng, lon0, lat0, n_target, radius_deg, base=64, amp=0.35):
    """Midpoint displacement of a `base`-gon: each level inserts the displaced
    midpoint of every segment (perpendicular offset N(0,1) * amp * length,
    amp halving per level); the last level displaces only the first
    segments, so the ring has exactly n_target vertices."""
    lon, lat = destination(lon0, lat0, 2.0 * np.pi * np.arange(base) / base, np.full(base, radius_deg * DEG))
    pts = _xyz(lon, lat)
    a = amp
    while len(pts) < n_target:
        n = len(pts)
        k = min(n, n_target - n)  # segments to split this level
        A = pts[:k]
        B = np.roll(pts, -1, axis=0)[:k]
        m = A + B
        m /= np.linalg.norm(m, axis=1, keepdims=True)
        L = np.arccos(np.clip(np.sum(A * B, axis=1), -1.0, 1.0))
        nrm = np.cross(A, B)
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        d = rng.standard_normal(k) * a * L
        mp = m * np.cos(d)[:, None] + nrm * np.sin(d)[:, None]
        out = np.empty((n + k, 3))
        out[0:2 * k:2] = A
        out[1:2 * k:2] = mp
        out[2 * k:] = pts[k:]
        pts = out
        a *= 0.5
    return _lonlat(pts)


def assemble(rings):
    """rings: list of (lon, lat) without closing vertex -> (lon, lat, ring_off)
    with every ring closed."""
    lons, lats, off = [], [], [0]
    for lo, la in rings:
        lons.append(np.concatenate([lo, lo[:1]]))
        lats.append(np.concatenate([la, la[:1]]))
        off.append(off[-1] + len(lo) + 1)
    return np.concatenate(lons), np.concatenate(lats), np.asarray(off, dtype=np.int64)


def write_polygon_txt(path, lon, lat, ring_off):
    """Reference polygon TXT layout with %.17g coordinates (exact round trip)."""
    with open(path, "w") as f:
        f.write(f"POLYGON 1 {len(ring_off) - 2}\n")
        for r in range(len(ring_off) - 1):
            s, e = ring_off[r], ring_off[r + 1]
            f.write(f"{r if r else 0} {e - s}\n")
            for k in range(s, e):
                f.write(f"{lon[k]:.17g} {lat[k]:.17g}\n")


def write_query_csv(path, lon, lat):
    with open(path, "w") as f:
        for a, b in zip(lon, lat):
            f.write(f"{a:.17g},{b:.17g}\n")


def _fixed_columns(v, decimals):
    """v as fixed-point text, one row of bytes per value, right-aligned with
    0 bytes as padding."""
    scale = 10 ** decimals
    a = np.round(np.abs(v) * scale).astype(np.int64)
    ip, fp = a // scale, a % scale
    ndig = np.ones(len(v), dtype=np.int64)
    for k in range(1, 19):
        ndig += ip >= 10 ** k
    maxd = int(ndig.max()) if len(v) else 1
    width = 1 + maxd + 1 + decimals
    m = np.zeros((len(v), width), dtype=np.uint8)
    for k in range(decimals):
        m[:, width - 1 - k] = ord("0") + (fp // 10 ** k) % 10
    m[:, width - 1 - decimals] = ord(".")
    for t in range(maxd):
        col = width - 2 - decimals - t
        m[:, col] = np.where(t < ndig, ord("0") + (ip // 10 ** t) % 10, 0)
    neg = np.signbit(v)
    rows = np.nonzero(neg)[0]
    m[rows, width - 2 - decimals - ndig[rows]] = ord("-")
    return m


def query_csv_bytes(lon, lat, decimals=6):
    """A query CSV ("lon,lat" per line, fixed-point with `decimals` digits,
    as the reference's query files are written), built with array
    operations for large point counts."""
    a = _fixed_columns(np.asarray(lon, dtype=np.float64), decimals)
    b = _fixed_columns(np.asarray(lat, dtype=np.float64), decimals)
    n = len(a)
    m = np.concatenate([a, np.full((n, 1), ord(","), np.uint8), b, np.full((n, 1), ord("\n"), np.uint8)], axis=1)
    flat = m.reshape(-1)
    return flat[flat != 0].tobytes()


def uniform_queries(rng, n):
    lon = rng.uniform(-180.0, 180.0, n)
    lat = np.degrees(np.arcsin(rng.uniform(-1.0, 1.0, n)))
    return lon, lat


def near_edge_queries(rng, lon, lat, ring_off, n, sigma_deg=0.005):
    """Points on random polygon edges (uniform t along the arc) pushed off the
    edge along its normal by N(0, sigma)."""
    closed = np.ones(len(lon), dtype=bool)
    closed[ring_off[1:] - 1] = False  # edge k -> k+1 exists unless k is a ring end
    starts = np.nonzero(closed)[0]
    e = starts[rng.integers(0, len(starts), n)]
    A = _xyz(lon[e], lat[e])
    B = _xyz(lon[e + 1], lat[e + 1])
    t = rng.uniform(0.0, 1.0, n)[:, None]
    P = (1.0 - t) * A + t * B
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    nrm = np.cross(A, B)
    nn = np.linalg.norm(nrm, axis=1, keepdims=True)
    nn[nn == 0] = 1.0
    nrm /= nn
    off = rng.standard_normal(n) * sigma_deg * DEG
    P = P * np.cos(off)[:, None] + nrm * np.sin(off)[:, None]
    return _lonlat(P)


def mixed_queries(rng, lon, lat, ring_off, n, frac_uniform):
    nu = int(round(n * frac_uniform))
    ul, ua = uniform_queries(rng, nu)
    el, ea = near_edge_queries(rng, lon, lat, ring_off, n - nu)
    ql = np.concatenate([ul, el])
    qa = np.concatenate([ua, ea])
    perm = rng.permutation(n)
    return ql[perm], qa[perm]


def star_1k(n_queries=1_000_000, seed=1):
    rng = np.random.default_rng(seed)
    outer = star_ring(rng, 20.0, 35.0, 1000, 15.0)
    hole = star_ring(rng, 20.0, 35.0, 200, 4.0)
    lon, lat, off = assemble([outer, hole])
    q = uniform_queries(np.random.default_rng(seed + 1000), n_queries)
    return dict(name="star_1k", lon=lon, lat=lat, ring_off=off, qlon=q[0], qlat=q[1])


def fractal_100k(n_queries=10_000_000, seed=2, n_outer=100_000):
    rng = np.random.default_rng(seed)
    lon0, lat0 = -40.0, 10.0
    outer = fractal_ring(rng, lon0, lat0, n_outer, 30.0)
    holes = []
    for j in range(5):
        hl, ha = destination(lon0, lat0, 2.0 * np.pi * j / 5, 12.0 * DEG)
        holes.append(star_ring(rng, float(hl), float(ha), 2000, 3.0))
    lon, lat, off = assemble([outer] + holes)
    q = mixed_queries(np.random.default_rng(seed + 1000), lon, lat, off, n_queries, 0.5)
    return dict(name="fractal_100k", lon=lon, lat=lat, ring_off=off, qlon=q[0], qlat=q[1])


def fractal_1m(n_queries=100_000_000, seed=3, n_outer=1_000_000, n_hole=50_000):
    rng = np.random.default_rng(seed)
    lon0, lat0 = 0.0, 60.0
    outer = fractal_ring(rng, lon0, lat0, n_outer, 40.0)
    holes = []
    for j in range(3):
        hl, ha = destination(lon0, lat0, 2.0 * np.pi * j / 3 + np.pi / 3, 15.0 * DEG)
        holes.append(star_ring(rng, float(hl), float(ha), n_hole, 4.0))
    lon, lat, off = assemble([outer] + holes)
    q = mixed_queries(np.random.default_rng(seed + 1000), lon, lat, off, n_queries, 0.7) if n_queries else (None, None)
    return dict(name="fractal_1m", lon=lon, lat=lat, ring_off=off, qlon=q[0], qlat=q[1])


def cap_queries(rng, lon0, lat0, radius_deg, n):
    """n points uniform in the spherical cap of the given radius."""
    ct = rng.uniform(np.cos(radius_deg * DEG), 1.0, n)
    az = rng.uniform(0.0, 2.0 * np.pi, n)
    return destination(lon0, lat0, az, np.arccos(ct))


def batch_polygons(n_poly=256, seed=4, n_min=16, n_max=200_000):
    """n_poly polygons: vertex counts log-uniform in [n_min, n_max], centres
    uniform on the sphere, radii U(1, 25) deg, 0-4 star holes of
    max(16, n/10) vertices at 0.4 R.  Rings of >= 1024 vertices are fractal
    (64-gon base), smaller ones star-shaped."""
    rng = np.random.default_rng(seed)
    polys = []
    for k in range(n_poly):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        lon0 = float(rng.uniform(-180.0, 180.0))
        lat0 = float(np.degrees(np.arcsin(rng.uniform(-1.0, 1.0))))
        R = float(rng.uniform(1.0, 25.0))
        if n >= 1024:
            outer = fractal_ring(rng, lon0, lat0, n, R)
        else:
            outer = star_ring(rng, lon0, lat0, n, R)
        h = int(rng.integers(0, 5))
        holes = []
        for j in range(h):
            hl, ha = destination(lon0, lat0, 2.0 * np.pi * j / max(h, 1), 0.4 * R * DEG)
            holes.append(star_ring(rng, float(hl), float(ha), max(16, n // 10), 0.12 * R))
        lon, lat, off = assemble([outer] + holes)
        polys.append(dict(lon=lon, lat=lat, ring_off=off, centre=(lon0, lat0), radius=R))
    return polys


def batch_256(n_queries=64_000_000, seed=4, n_poly=256):
    """The batch plus queries: Zipf(1.1) polygon choice over a random rank
    order, each point uniform in its polygon's cap of radius 1.3 R."""
    polys = batch_polygons(n_poly, seed)
    rng = np.random.default_rng(seed + 1000)
    rank = rng.permutation(n_poly)
    w = 1.0 / np.arange(1, n_poly + 1) ** 1.1
    w /= w.sum()
    cdf = np.cumsum(w)
    cdf[-1] = 1.0
    pid = rank[np.searchsorted(cdf, rng.random(n_queries), side="right")].astype(np.int32)
    qlon = np.empty(n_queries)
    qlat = np.empty(n_queries)
    order = np.argsort(pid, kind="stable")
    counts = np.bincount(pid, minlength=n_poly)
    start = 0
    for k in range(n_poly):
        cnt = counts[k]
        if cnt:
            c = polys[k]
            lo, la = cap_queries(rng, c["centre"][0], c["centre"][1], 1.3 * c["radius"], cnt)
            idx = order[start:start + cnt]
            qlon[idx] = lo
            qlat[idx] = la
        start += cnt
    return dict(name="batch_256", polys=polys, qlon=qlon, qlat=qlat, qpoly=pid)


def fractal_512(n_queries=32_000_000, seed=6, n_poly=512, n_min=4096, n_max=65536):
    """512 polygons of the fractal_1m generator (fractal outer ring on a
    64-gon, 3 star holes at 0.375 R of radius 0.1 R with n/20 vertices),
    scaled down to outer vertex counts log-uniform in [n_min, n_max] (the
    full 1M vertices would make 512M), centres uniform on the sphere, radii
    U(10, 40) deg.  Each polygon gets n_queries / n_poly points, fractal_1m's
    mix: 70% uniform in its cap of radius 1.3 R, 30% near its edges."""
    rng = np.random.default_rng(seed)
    polys = []
    for k in range(n_poly):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        lon0 = float(rng.uniform(-180.0, 180.0))
        lat0 = float(np.degrees(np.arcsin(rng.uniform(-1.0, 1.0))))
        R = float(rng.uniform(10.0, 40.0))
        outer = fractal_ring(rng, lon0, lat0, n, R)
        holes = []
        for j in range(3):
            hl, ha = destination(lon0, lat0, 2.0 * np.pi * j / 3 + np.pi / 3, 0.375 * R * DEG)
            holes.append(star_ring(rng, float(hl), float(ha), max(16, n // 20), 0.1 * R))
        lon, lat, off = assemble([outer] + holes)
        polys.append(dict(lon=lon, lat=lat, ring_off=off, centre=(lon0, lat0), radius=R))
    per = n_queries // n_poly
    qrng = np.random.default_rng(seed + 1000)
    qlon, qlat, qpoly = [], [], []
    for k, c in enumerate(polys):
        nu = int(round(0.7 * per))
        ul, ua = cap_queries(qrng, c["centre"][0], c["centre"][1], 1.3 * c["radius"], nu)
        el, ea = near_edge_queries(qrng, c["lon"], c["lat"], c["ring_off"], per - nu)
        qlon += [ul, el]
        qlat += [ua, ea]
        qpoly.append(np.full(per, k, dtype=np.int32))
    perm = qrng.permutation(per * n_poly)
    return dict(name="fractal_512", polys=polys, qlon=np.concatenate(qlon)[perm], qlat=np.concatenate(qlat)[perm],
                qpoly=np.concatenate(qpoly)[perm])


WORKLOADS = {"star_1k": star_1k, "fractal_100k": fractal_100k, "fractal_1m": fractal_1m, "batch_256": batch_256,
             "fractal_512": fractal_512}
