"""Polygons at the geometric edge cases of the path, as (name, lon, lat,
ring_off) with closing vertices included, plus query points that sit on
their features.  Shared by the oracle tests and the GPU parity tests."""
import numpy as np


def ring(lon, lat):
    lon = np.asarray(lon, dtype=np.float64)
    lat = np.asarray(lat, dtype=np.float64)
    return np.append(lon, lon[0]), np.append(lat, lat[0])


def assemble(rings):
    lon = np.concatenate([r[0] for r in rings])
    lat = np.concatenate([r[1] for r in rings])
    off = np.cumsum([0] + [len(r[0]) for r in rings]).astype(np.int64)
    return lon, lat, off


def circle(lon0, lat0, r_deg, n, ccw=True):
    """Small circle of radius r_deg around (lon0, lat0), n vertices."""
    t = np.linspace(0.0, 2 * np.pi, n, endpoint=False)
    if not ccw:
        t = -t
    p0, l0, r = np.radians(lat0), np.radians(lon0), np.radians(r_deg)
    lat = np.arcsin(np.sin(p0) * np.cos(r) + np.cos(p0) * np.sin(r) * np.cos(t))
    lon = l0 + np.arctan2(np.sin(t) * np.sin(r) * np.cos(p0), np.cos(r) - np.sin(p0) * np.sin(lat))
    lon = (np.degrees(lon) + 180.0) % 360.0 - 180.0
    return ring(lon, np.degrees(lat))


def cases():
    """(name, lon, lat, ring_off, simple): simple polygons (holes inside the
    outer ring, disjoint, no zero-length edge) have a geometric answer the
    hierarchy-free check gives; the others are inputs the reference accepts
    without checking, where only the algorithm itself (GPU against oracle)
    defines the labels."""
    out = []
    # cap around the north pole and around the south pole (ring of constant
    # latitude), each vertex's longitude on a 0.36 degree step
    lon = np.linspace(-180.0, 180.0, 1000, endpoint=False)
    out.append(("north_cap", *assemble([ring(lon, np.full(1000, 80.0))]), True))
    out.append(("south_cap", *assemble([ring(lon[::-1], np.full(1000, -75.0))]), True))
    # an annulus: a 40-degree cap with a 10-degree hole at its centre
    out.append(("annulus", *assemble([circle(30.0, 5.0, 40.0, 1500), circle(30.0, 5.0, 10.0, 400, ccw=False)]),
                True))
    # two latitude rings around the whole sphere (which side each ring
    # encloses is left to the orientation rule: the "hole" need not lie
    # inside the outer ring)
    out.append(("latitude_rings", *assemble([ring(lon, np.full(1000, 10.0)), ring(lon, np.full(1000, -10.0))]),
                False))
    # across the antimeridian
    out.append(("antimeridian", *assemble([circle(179.5, 20.0, 3.0, 400)]), True))
    # tiny: 1e-5 degree radius (deeper than the default depth cap)
    out.append(("tiny", *assemble([circle(12.3, 45.6, 1e-5, 64)]), True))
    # many holes: a 30-degree cap with 40 disjoint holes of 0.7 degrees
    rng = np.random.default_rng(3)
    rings = [circle(-60.0, -20.0, 30.0, 2000)]
    centres = []
    while len(centres) < 40:
        a, d = rng.uniform(0, 2 * np.pi), rng.uniform(0, 22.0)
        c = (-60.0 + d * np.cos(a), -20.0 + d * np.sin(a) * 0.8)
        if all(np.hypot((c[0] - e[0]) * np.cos(np.radians(c[1])), c[1] - e[1]) > 2.0 for e in centres):
            centres.append(c)
            rings.append(circle(c[0], c[1], 0.7, 50, ccw=False))
    out.append(("many_holes", *assemble(rings), True))
    # overlapping holes (parity and the nearest-edge classification disagree)
    rings = [circle(-60.0, -20.0, 30.0, 2000)] + [circle(-60.0 + 0.5 * k, -20.0, 1.0, 60, ccw=False)
                                                 for k in range(3)]
    out.append(("overlapping_holes", *assemble(rings), False))
    # clockwise ring (the orientation rule reverses it) with collinear runs
    # along the meridian lon = 7
    lo = np.array([7.0, 7.0, 7.0, 7.0, 37.0, 37.0])
    la = np.array([-3.0, 7.0, 17.0, 27.0, 27.0, -3.0])
    out.append(("collinear_cw", *assemble([ring(lo[::-1], la[::-1])]), True))
    # edges along the equator, which passes through level-0 cell centres: an
    # arc from such a centre starts on the edges' plane, where the
    # reference's epsilon rules (not the geometry) decide the crossings
    lo = np.array([0.0, 10.0, 20.0, 30.0, 40.0, 40.0, 20.0, 0.0])
    la = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 25.0, 30.0, 25.0])
    out.append(("equator_edges", *assemble([ring(lo[::-1], la[::-1])]), False))
    # a repeated vertex: a zero-length edge (its plane normal is 0/0)
    lo = np.array([0.0, 10.0, 10.0, 20.0, 30.0, 40.0, 40.0, 20.0, 0.0])
    la = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 25.0, 30.0, 25.0])
    out.append(("zero_length_edge", *assemble([ring(lo[::-1], la[::-1])]), False))
    # more than a hemisphere
    out.append(("large", *assemble([circle(100.0, -30.0, 140.0, 3000)]), True))
    return out


def feature_points(lon, lat, ring_off, rng, n_random=200_000):
    """Polygon vertices, edge midpoints, points 1e-12 rad off edges, the
    poles and the lon = +-180 meridian, plus uniform points (lon, lat)."""
    lo, la = [lon], [lat]
    xyz = np.stack([np.cos(np.radians(lat)) * np.sin(np.radians(lon)), np.sin(np.radians(lat)),
                    np.cos(np.radians(lat)) * np.cos(np.radians(lon))], axis=1)
    mids, offs = [], []
    for r in range(len(ring_off) - 1):
        a = xyz[ring_off[r]:ring_off[r + 1] - 1]
        b = xyz[ring_off[r] + 1:ring_off[r + 1]]
        m = a + b
        nm = np.linalg.norm(m, axis=1)
        ok = nm > 1e-12
        m = m[ok] / nm[ok, None]
        mids.append(m)
        nrm = np.cross(a[ok], b[ok])
        nn = np.linalg.norm(nrm, axis=1)
        good = nn > 1e-15
        nrm = nrm[good] / nn[good, None]
        for s in (1e-12, -1e-12):
            q = m[good] + s * nrm
            offs.append(q / np.linalg.norm(q, axis=1)[:, None])
    pts = np.concatenate(mids + offs) if mids else np.zeros((0, 3))
    plon = np.degrees(np.arctan2(pts[:, 0], pts[:, 2]))
    plat = np.degrees(np.arcsin(np.clip(pts[:, 1], -1, 1)))
    lo.append(plon)
    la.append(plat)
    lo.append(np.array([0.0, 0.0, 180.0, -180.0, 180.0, -180.0]))
    la.append(np.array([90.0, -90.0, 0.0, 0.0, 45.0, -45.0]))
    lo.append(rng.uniform(-180, 180, n_random))
    la.append(np.degrees(np.arcsin(rng.uniform(-1, 1, n_random))))
    return np.concatenate(lo), np.concatenate(la)
