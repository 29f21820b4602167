"""GPU parity: the gfx950 library against the CPU oracle, bit for bit.

P0  device FP64 primitives and predicates equal the oracle's bitwise;
P1  the hierarchy arrays built on the GPU equal the oracle's;
P2  query labels equal the oracle's (0 mismatches).
"""
import os
import sys

import numpy as np
import pytest

import gahhg_amd
from gahhg_amd import datasets
import oracle_lib as ol

pytestmark = pytest.mark.gpu
REF = os.path.join(os.path.dirname(__file__), "golden", "reference")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def refpath(n):
    return os.path.join(REF, n)


def _rand_unit(rng, n):
    x = rng.standard_normal((n, 3))
    return x / np.linalg.norm(x, axis=1)[:, None]


@pytest.fixture(scope="module")
def dev():
    g = gahhg_amd.GPISP(device=0)
    yield g
    g.close()


def test_kat_bitwise(dev):
    say("kat: 4 ops x 20000 cases")
    rng = np.random.default_rng(42)
    n = 20000
    a, b, c, d = (_rand_unit(rng, n) for _ in range(4))
    # near-degenerate cases: short arcs, points on the edge's great circle
    b[:2000] = a[:2000] + 1e-9 * _rand_unit(rng, 2000)
    m = a[2000:4000] + b[2000:4000]
    d[2000:4000] = m / np.linalg.norm(m, axis=1)[:, None]
    c[4000:6000] = a[4000:6000]
    d[6000:8000] = b[6000:8000] * (1 + 1e-12)
    x = np.concatenate([a, b, c, d], axis=1)
    L = ol.lib()
    out0 = dev.kat(0, x)
    out1 = dev.kat(1, x)
    out2 = dev.kat(2, x)
    out3 = dev.kat(3, x)
    bad = 0
    r3 = np.zeros(3)
    import ctypes
    res = ctypes.c_int32()
    for i in range(n):
        L.og_kat_normalize(np.ascontiguousarray(a[i]), r3)
        bad += not np.array_equal(r3, out0[i])
        L.og_kat_cross(np.ascontiguousarray(a[i]), np.ascontiguousarray(b[i]), r3)
        bad += not np.array_equal(r3, out1[i])
        e = L.og_kat_is_ee_inter(*(np.ascontiguousarray(v[i]) for v in (a, b, c, d)))
        bad += e != int(out2[i, 0])
        k = L.og_kat_edge_border_inter(*(np.ascontiguousarray(v[i]) for v in (a, b, c, d)), ctypes.byref(res))
        bad += (k != int(out3[i, 0])) or (res.value != int(out3[i, 1]))
    assert bad == 0


def _unit(x):
    return x / np.linalg.norm(x, axis=-1, keepdims=True)


def cp_prop_cases(pv, pe, rng, n_tie=5000, n_near=3000, n_vertex=500, n_arc=500):
    """getCPProp inputs that reach every branch of its loop: centres near a
    vertex with both edges at it among the refs (the shared-vertex tie rule),
    centres near an edge with up to 150 nearby refs (more than a 64-lane
    group), centres exactly at a vertex and exactly on an edge's great
    circle (zero distances and zero sides)."""
    N = len(pe)
    nxt = np.empty(N, dtype=np.int64)
    prv = np.empty(N, dtype=np.int64)
    nxt[pe[:, 0]] = np.arange(N)
    prv[pe[:, 1]] = np.arange(N)
    mid = _unit(pv[pe[:, 0]] + pv[pe[:, 1]])
    flat, rs_rn, centers, tie_vertex = [], [], [], []

    def add(refs, c, k=-1):
        rs_rn.append((len(flat), len(refs)))
        flat.extend(int(r) for r in refs)
        centers.append(c)
        tie_vertex.append(k)

    def nearest(c, m):
        return np.argsort(-(mid @ c))[:m]

    for _ in range(n_tie):
        k = int(rng.integers(N))
        c = _unit(pv[k] + 10 ** rng.uniform(-9, -2) * _unit(rng.normal(size=3)))
        others = [e for e in nearest(c, 8) if e not in (prv[k], nxt[k])][:int(rng.integers(0, 7))]
        refs = [prv[k], nxt[k]] + others
        rng.shuffle(refs)
        add(refs, c, k)
    for _ in range(n_near):
        e = int(rng.integers(N))
        c = _unit(mid[e] + 10 ** rng.uniform(-8, -1) * _unit(rng.normal(size=3)))
        refs = list(nearest(c, int(rng.integers(1, 151))))
        rng.shuffle(refs)
        add(refs, c)
    for _ in range(n_vertex):
        k = int(rng.integers(N))
        refs = [prv[k], nxt[k]] + list(nearest(pv[k], int(rng.integers(0, 5))))
        rng.shuffle(refs)
        add(refs, pv[k].copy(), k)
    for _ in range(n_arc):
        e = int(rng.integers(N))
        refs = list(nearest(mid[e], int(rng.integers(1, 9))))
        if e not in refs:
            refs.append(e)
        rng.shuffle(refs)
        add(refs, mid[e].copy())
    return (np.array(flat, dtype=np.int32), np.array(rs_rn, dtype=np.int32), np.array(centers),
            np.array(tie_vertex))


def test_kat_cp_prop_bitwise(dev):
    """getCPProp (device_getCPProp, source/GPISPAHHG.cu:61-171) as
    setBottomNECell runs it on the GPU -- a lane group per cell replaying
    the selection -- against the oracle, case by case: type, chosen edge,
    nearest vertex and squared distance bit for bit, for every group size."""
    say("kat: getCPProp, NA0, 9000 cases x 7 group sizes")
    dev.load_polygon_txt(refpath("NA0.txt")).build()
    o = ol.Oracle().load_txt(refpath("NA0.txt"))
    pv, pe = o.polygon()
    refs, rs_rn, centers, tie_vertex = cp_prop_cases(pv, pe, np.random.default_rng(11))
    want = np.zeros((len(rs_rn), 3), dtype=np.int32)
    want_md = np.zeros(len(rs_rn))
    for k, (rs, rn) in enumerate(rs_rn):
        t, sl, nv, md = ol.cp_prop(pv, pe, refs[rs:rs + rn], centers[k])
        want[k] = (t, sl, nv)
        want_md[k] = md
    # the cases reach the branches: both types, vertex and interior
    # features, the shared vertex of a tie case chosen
    assert set(np.unique(want[:, 0])) == {1, 2}
    assert np.mean(want[:, 2] == -1) > 0.2 and np.mean(want[:, 2] >= 0) > 0.2
    tie = tie_vertex >= 0
    assert np.mean(want[tie, 2] == tie_vertex[tie]) > 0.25
    assert np.all(want[:, 1] >= 0)
    for G in (1, 2, 4, 8, 16, 32, 64):
        got, md = dev.kat_cp_prop(refs, rs_rn, centers, G)
        bad = np.nonzero((got != want).any(axis=1) | (md.view(np.uint64) != want_md.view(np.uint64)))[0]
        assert len(bad) == 0, (G, bad[:5], got[bad[:3]], want[bad[:3]])


def compare_grids(gg, og, levels_o):
    assert gg["depth"] == len(levels_o)
    lv = [[L[k] for k in ("vstart", "vnum", "bstart", "bnum", "cstart", "cnum", "rnum")] for L in levels_o]
    assert gg["levels"].tolist() == lv
    for key in ("v", "cells", "borders", "normals", "refs"):
        x, y = gg[key], og[key]
        assert x.shape == y.shape, key
        if key in ("v", "normals"):
            diff = np.nonzero((x.view(np.int64) != y.view(np.int64)).any(axis=1))[0]
        else:
            diff = np.nonzero((x != y).reshape(len(x), -1).any(axis=1))[0]
        assert len(diff) == 0, f"{key}: {len(diff)} rows differ, first {diff[:5]}: gpu {x[diff[:2]]} oracle {y[diff[:2]]}"


POLYS = ["Square_72.txt", "Radiometer.txt", "NA0.txt", "Mars_001.txt", "Mars_002.txt", "Mars_100.txt",
         "Mars_300.txt", "Mars_624.txt", "Tennessee_Tennessee_00005.txt", "Tennessee_Tennessee_02999.txt"]


@pytest.mark.parametrize("name", POLYS)
def test_grid_and_labels_reference_polygons(dev, name):
    say("grid+labels", name)
    dev.load_polygon_txt(refpath(name)).build()
    o = ol.Oracle().load_txt(refpath(name)).build()
    compare_grids(dev.grid(), o.grid(), o.levels())
    st = dev.stats()
    assert st["depth"] == o.counts()["depth"]
    assert st["gpu_space_bytes"] == o.space_bytes()
    assert abs(st["gamma"] - o.gamma()) == 0
    qfile = {"Square_72.txt": "Square_72_query.csv", "Radiometer.txt": "Radiometer_query.csv"}.get(name, "Mars_query.csv")
    lon, lat = ol.load_query_csv(refpath(qfile))
    rlon, rlat = datasets.uniform_queries(np.random.default_rng(7), 200000)
    xyz = ol.query_lonlat_to_xyz(np.concatenate([lon, rlon]), np.concatenate([lat, rlat]))
    g_lab = dev.query_xyz(xyz)
    o_lab = o.query_xyz(xyz)
    assert np.count_nonzero(g_lab != o_lab) == 0
    assert dev.stats()["inside"] == int((o_lab == 2).sum())


def test_star_1k_million_queries(dev):
    say("star_1k 1M")
    w = datasets.star_1k()
    dev.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    o = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    compare_grids(dev.grid(), o.grid(), o.levels())
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    g_lab = dev.query_xyz(xyz)
    o_lab = o.query_xyz(xyz)
    assert len(g_lab) == 1_000_000
    assert np.count_nonzero(g_lab != o_lab) == 0
    assert (g_lab == 0).sum() == 0


def test_fractal_grid_parity(dev):
    say("fractal 20k")
    w = datasets.fractal_100k(n_queries=300_000, n_outer=20_000)
    dev.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    o = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    compare_grids(dev.grid(), o.grid(), o.levels())
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    assert np.count_nonzero(dev.query_xyz(xyz) != o.query_xyz(xyz)) == 0


@pytest.mark.parametrize("depth", [1, 3, 8])
def test_fixed_depth_parity(dev, depth):
    say("fixed depth", depth)
    g = gahhg_amd.GPISP(device=0, max_depth=depth, auto_depth=False)
    g.load_polygon_txt(refpath("Radiometer.txt")).build()
    o = ol.Oracle().load_txt(refpath("Radiometer.txt")).build(max_depth=depth, auto_depth=False)
    compare_grids(g.grid(), o.grid(), o.levels())
    lon, lat = ol.load_query_csv(refpath("Radiometer_query.csv"))
    xyz = ol.query_lonlat_to_xyz(lon, lat)
    assert np.count_nonzero(g.query_xyz(xyz) != o.query_xyz(xyz)) == 0
    g.close()


def test_device_resident_query_and_edge_cases(dev):
    say("device query")
    dev.load_polygon_txt(refpath("NA0.txt")).build()
    lon, lat = datasets.uniform_queries(np.random.default_rng(1), 100000)
    xyz = gahhg_amd.query_lonlat_to_xyz(lon, lat)
    host = dev.query_xyz(xyz)
    dx = dev.alloc(xyz.nbytes)
    dl = dev.alloc(4 * len(host))
    dx.upload(xyz)
    dev.query_device(dx, len(host), dl)
    out = dl.download(np.zeros(len(host), dtype=np.int32))
    assert (out == host).all()
    # empty query
    assert dev.query_xyz(np.zeros((0, 3))).size == 0
    dx.free()
    dl.free()


def test_error_paths(tmp_path):
    say("errors")
    g = gahhg_amd.GPISP(device=0)
    with pytest.raises(gahhg_amd.GahhgError) as e:
        g.query_xyz(np.zeros((3, 3)))
    assert e.value.code == -4
    with pytest.raises(gahhg_amd.GahhgError) as e:
        g.build()
    assert e.value.code == -4
    with pytest.raises(gahhg_amd.GahhgError) as e:
        g.load_polygon_txt(tmp_path / "missing.txt")
    assert e.value.code == -2
    bad = tmp_path / "bad.txt"
    bad.write_text("POLYGON 1 0\n0 3\n0 0\n1 1\n0 0\n")
    with pytest.raises(gahhg_amd.GahhgError) as e:
        g.load_polygon_txt(bad)
    assert e.value.code == -3
    g.close()


def test_report_format(dev, tmp_path):
    say("report")
    dev.load_polygon_txt(refpath("Square_72.txt")).build()
    lon, lat = ol.load_query_csv(refpath("Square_72_query.csv"))
    dev.query_lonlat(lon, lat)
    p = tmp_path / "report.txt"
    dev.write_report(p, "Square_72.txt", "Square_72_query.csv")
    txt = p.read_text()
    assert "Polygon file: Square_72.txt" in txt
    assert "Number of polygon vertice: 4" in txt
    assert "Depth of GAHHG: 1" in txt
    assert "Number of query points: 10000" in txt
    st = dev.stats()
    assert f"Test result: inside {st['inside']}, outside {10000 - st['inside']}" in txt


def bin_lattice(B, S, face_axis=2, sign=1.0, centre=(0.0, 0.0), span=40):
    """Points on the bin and sub-bin boundary lines (u, v = -1 + 2 k / (B S))
    of one cube face around a given (u, v) centre."""
    k0 = np.floor((np.asarray(centre) + 1.0) * 0.5 * B * S).astype(int)
    ks = np.arange(-span, span + 1)
    uu = -1.0 + 2.0 * (k0[0] + ks) / (B * S)
    vv = -1.0 + 2.0 * (k0[1] + ks) / (B * S)
    U, V = np.meshgrid(uu, vv)
    uv = np.stack([U.ravel(), V.ravel()], axis=1)
    uv = uv[(np.abs(uv) <= 1.0).all(axis=1)]
    p = np.zeros((len(uv), 3))
    other = [a for a in range(3) if a != face_axis]
    p[:, face_axis] = sign
    p[:, other[0]] = uv[:, 0]
    p[:, other[1]] = uv[:, 1]
    return p / np.linalg.norm(p, axis=1)[:, None]


@pytest.mark.parametrize("layout,bins,group,sub,sort", [
    (1, 0, 0, 64, 0), (0, 0, 0, 64, 0), (0, 16, 0, 64, 0), (0, 1024, 0, 64, 0), (0, 256, 0, 1, 0),
    (0, 256, 0, 256, 0), (0, 256, 1, 64, 0), (0, 256, 2, 64, 0), (0, 256, 4, 64, 0), (0, 256, 8, 64, 0),
    (0, 256, 16, 64, 0), (0, 256, 32, 64, 0), (0, 256, 64, 64, 0), (0, 256, 0, 64, 2), (0, 0, 8, 1, 2),
    (0, 1024, 64, 8, 2)])
def test_query_paths_agree(layout, bins, group, sub, sort):
    say("query path", layout, bins, group, sub, sort)
    w = datasets.fractal_100k(n_queries=400_000, n_outer=30_000)
    g = gahhg_amd.GPISP(device=0, query_layout=layout, query_bins=bins, query_group=group, query_sub=sub,
                        query_sub_budget=64, query_sort=sort)
    g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    o = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    # exact cube-face, bin and sub-bin boundaries too, on a polygon vertex's
    # face (x axis dominant for lon -40, lat 10)
    extra = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1],
                      [1, 1, 0], [1, -1, 1], [-1, 1, 1], [0.5, 1, 0]], dtype=np.float64)
    extra /= np.linalg.norm(extra, axis=1)[:, None]
    v = ol.query_lonlat_to_xyz(w["lon"][:1], w["lat"][:1])[0]
    ax = int(np.argmax(np.abs(v)))
    uvc = [v[a] / abs(v[ax]) for a in range(3) if a != ax]
    lat = [bin_lattice(max(bins, 1), s, ax, np.sign(v[ax]), uvc) for s in (1, 2, 8, 64)]
    xyz = np.concatenate([xyz, extra] + lat)
    assert np.count_nonzero(g.query_xyz(xyz) != o.query_xyz(xyz)) == 0
    g.close()


@pytest.mark.parametrize("graph", [0, 1])
def test_persistent_launch_agrees(graph):
    """The persistent work-stealing walk (query_persistent=1: a resident
    grid whose waves take 64-point tiles from per-XCD queues and steal from
    the others) labels every point as the oracle does: a single polygon with
    grouped points (more tiles than the grid has waves, a count known only
    on the device) and a batch, recorded as a graph or launched directly."""
    say("persistent launch, graph", graph)
    w = datasets.fractal_100k(n_queries=1_200_000, n_outer=30_000)
    o = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    want = o.query_xyz(xyz)
    with gahhg_amd.GPISP(device=0, query_persistent=1, query_graph=graph, query_sort=2) as g:
        g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
        for _ in range(2):
            assert np.count_nonzero(g.query_xyz(xyz) != want) == 0
        # fewer points than one tile, and an empty query
        assert np.array_equal(g.query_xyz(xyz[:37]), want[:37])
        assert g.query_xyz(np.zeros((0, 3))).size == 0
    ps = datasets.batch_polygons(n_poly=6, seed=33, n_min=50, n_max=20_000)
    items = [(p["lon"], p["lat"], p["ring_off"]) for p in ps]
    rng = np.random.default_rng(8)
    lon, lat = datasets.uniform_queries(rng, 400_000)
    bx = gahhg_amd.query_lonlat_to_xyz(lon, lat)
    poly = rng.integers(0, len(items), len(lon)).astype(np.int32)
    with gahhg_amd.GPISP(device=0, query_persistent=1, query_graph=graph, query_sort=2) as g:
        g.set_polygons_lonlat(items).build()
        got = g.query_batch_xyz(bx, poly)
    for k, it in enumerate(items):
        sel = poly == k
        ok = ol.Oracle().set_lonlat(*it).build()
        assert np.count_nonzero(got[sel] != ok.query_xyz(np.ascontiguousarray(bx[sel]))) == 0, k


def test_cli_matches_oracle(tmp_path):
    say("cli")
    import subprocess
    exe = os.path.join(os.path.dirname(gahhg_amd.LIB_PATH), "gahhg")
    rep, res = tmp_path / "report.txt", tmp_path / "result.csv"
    r = subprocess.run([exe, "/polyfile", refpath("Radiometer.txt"), "/tpfile", refpath("Radiometer_query.csv"),
                        "/reportfile", str(rep), "/resultfile", str(res), "/unknownflag", "x"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.array([int(x) for x in res.read_text().split()])
    o = ol.Oracle().load_txt(refpath("Radiometer.txt")).build()
    lon, lat = ol.load_query_csv(refpath("Radiometer_query.csv"))
    exp = (o.query_xyz(ol.query_lonlat_to_xyz(lon, lat)) == 2).astype(int)
    assert (got == exp).all()
    assert "Depth of GAHHG: 4" in rep.read_text()
    # AHHG-style fixed depth
    r = subprocess.run([exe, "/polyfile", refpath("Radiometer.txt"), "/tpfile", refpath("Radiometer_query.csv"),
                        "/reportfile", str(rep), "/resultfile", str(res), "/maxdepth", "6"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    assert "Depth of GAHHG: 6" in rep.read_text().split("-----\n")[-2] or "Depth of GAHHG: 6" in rep.read_text()


def test_batch_matches_single_polygon_builds():
    say("batch of reference + synthetic polygons")
    from gahhg_amd import datasets as ds
    items = []
    # reference polygons re-read as lon/lat rings
    for name in ["Square_72.txt", "Radiometer.txt", "NA0.txt", "Mars_300.txt", "Mars_001.txt", "Tennessee_Tennessee_00500.txt"]:
        lon, lat, roff = [], [], [0]
        with open(refpath(name)) as f:
            holes = int(f.readline().split()[2])
            for r in range(holes + 1):
                cnt = int(f.readline().split()[1])
                for _ in range(cnt):
                    a, b = f.readline().split()[:2]
                    lon.append(float(a))
                    lat.append(float(b))
                roff.append(len(lon))
        items.append((np.array(lon), np.array(lat), np.array(roff)))
    w = ds.star_1k(n_queries=10)
    items.append((w["lon"], w["lat"], w["ring_off"]))
    w2 = ds.fractal_100k(n_queries=10, n_outer=8000)
    items.append((w2["lon"], w2["lat"], w2["ring_off"]))
    g = gahhg_amd.GPISP(device=0)
    g.set_polygons_lonlat(items).build()
    assert g.n_polygons == len(items)
    rng = np.random.default_rng(3)
    nq = 400_000
    qlon, qlat = ds.uniform_queries(rng, nq)
    xyz = ol.query_lonlat_to_xyz(qlon, qlat)
    poly = rng.integers(0, len(items), nq).astype(np.int32)
    got = g.query_batch_xyz(xyz, poly)
    for k, (lon, lat, roff) in enumerate(items):
        o = ol.Oracle().set_lonlat(lon, lat, roff).build()
        compare_grids(g.polygon_grid(k), o.grid(), o.levels())
        ps = g.polygon_stats(k)
        assert ps["depth"] == o.counts()["depth"] and ps["gpu_space_bytes"] == o.space_bytes()
        sel = poly == k
        assert np.count_nonzero(got[sel] != o.query_xyz(xyz[sel])) == 0, k
    # the plain query path and the sorted accelerated path on the same batch
    gp = gahhg_amd.GPISP(device=0, query_layout=1, query_bins=0)
    gp.set_polygons_lonlat(items).build()
    assert (gp.query_batch_xyz(xyz, poly) == got).all()
    gs = gahhg_amd.GPISP(device=0, query_sort=2)
    gs.set_polygons_lonlat(items).build()
    assert (gs.query_batch_xyz(xyz, poly) == got).all()
    gs.close()
    # start tables of one size for all polygons (budget 0), of sizes
    # between the default's and the full one (budget 40), and of sizes that
    # are not powers of two (no shared level-0 parts)
    for budget, bins in ((0, 256), (40, 256), (8, 100)):
        gt = gahhg_amd.GPISP(device=0, query_table_budget=budget, query_bins=bins)
        gt.set_polygons_lonlat(items).build()
        assert (gt.query_batch_xyz(xyz, poly) == got).all(), (budget, bins)
        gt.close()
    g.close()
    gp.close()


def test_refs_beyond_flag_scratch():
    """Regression: a polygon whose ref count R exceeds the per-cell/per-edge
    scratch (dense star hole, R ~ 4.9 N) must build and match the oracle."""
    say("polygon 68 of the 256-polygon batch, alone (R > 4 N)")
    from gahhg_amd import datasets as ds
    p = ds.batch_polygons(n_poly=69)[68]
    o = ol.Oracle().set_lonlat(p["lon"], p["lat"], p["ring_off"]).build()
    assert o.counts()["R"] > 4 * o.n_edges
    g = gahhg_amd.GPISP(device=0)
    g.set_polygon_lonlat(p["lon"], p["lat"], p["ring_off"]).build()
    compare_grids(g.grid(), o.grid(), o.levels())
    rng = np.random.default_rng(2)
    lon, lat = ds.cap_queries(rng, p["centre"][0], p["centre"][1], 1.3 * p["radius"], 300_000)
    xyz = ol.query_lonlat_to_xyz(lon, lat)
    assert np.count_nonzero(g.query_xyz(xyz) != o.query_xyz(xyz)) == 0
    g.close()


def test_leaf_index_agrees():
    """The angular leaf index (query_index=1, default) skips only records
    that cannot count: labels equal the full-stream scan (query_index=0) and
    the oracle on points 1e-6 .. 1e-12 rad either side of the edges, at and
    around the bottom-cell centres (the arcs' origin), and uniform points."""
    say("leaf index vs full scan")
    w = datasets.fractal_100k(n_queries=300_000, n_outer=30_000)
    g1 = gahhg_amd.GPISP(device=0, query_index=1)
    g0 = gahhg_amd.GPISP(device=0, query_index=0)
    for g in (g0, g1):
        g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    o = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    pv, pe = o.polygon()
    rng = np.random.default_rng(17)
    pts = [gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])]
    e = rng.choice(len(pe), 40_000)
    a, b = pv[pe[e, 0]], pv[pe[e, 1]]
    t = rng.uniform(0, 1, (len(e), 1))
    m = a * (1 - t) + b * t
    n = np.cross(a, b)
    n /= np.linalg.norm(n, axis=1)[:, None]
    for off in (1e-6, 1e-8, 1e-9, 1e-10, 1e-12):
        for sgn in (1, -1):
            q = m + sgn * off * n
            pts.append(q / np.linalg.norm(q, axis=1)[:, None])
    gg = g1.grid()
    cells, lv = gg["cells"], gg["levels"]
    bot = np.arange(lv[-1][4], lv[-1][4] + lv[-1][5])
    bot = bot[cells[bot, 3] >= 16][:20_000]
    cen = np.zeros((len(bot), 3))
    for k in range(6):
        has = cells[bot, 1] > k
        cen[has] += gg["v"][gg["borders"][cells[bot[has], 0] + k, 2]]
    cen /= np.linalg.norm(cen, axis=1)[:, None]
    pts.append(cen)
    for s in (1e-14, 1e-11, 1e-8, 1e-6):
        q = cen + s * rng.standard_normal(cen.shape)
        pts.append(q / np.linalg.norm(q, axis=1)[:, None])
    xyz = np.concatenate(pts)
    l1, l0 = g1.query_xyz(xyz), g0.query_xyz(xyz)
    assert np.count_nonzero(l1 != l0) == 0
    assert np.count_nonzero(l1 != o.query_xyz(xyz)) == 0
    g0.close()
    g1.close()


def test_pairs_query_and_device_id_check():
    """gahhg_query_pairs_xyz (every polygon x the same points) equals the
    batch query with the points repeated per polygon; device polygon ids
    outside the batch are rejected before any query kernel runs."""
    say("pairs query and device polygon-id check")
    items = ["Square_72.txt", "Radiometer.txt", "Mars_300.txt"]
    g = gahhg_amd.GPISP(device=0)
    g.load_polygons_txt([refpath(n) for n in items]).build()
    P = g.n_polygons
    rng = np.random.default_rng(9)
    lon, lat = datasets.uniform_queries(rng, 7001)
    xyz = gahhg_amd.query_lonlat_to_xyz(lon, lat)
    pairs = g.query_pairs_xyz(xyz)
    assert pairs.shape == (P, len(xyz))
    rep = g.query_batch_xyz(np.tile(xyz, (P, 1)), np.repeat(np.arange(P, dtype=np.int32), len(xyz)))
    assert (pairs.reshape(-1) == rep).all()
    for k, name in enumerate(items):
        o = ol.Oracle().load_txt(refpath(name)).build()
        assert np.count_nonzero(pairs[k] != o.query_xyz(xyz)) == 0, name
    n = len(xyz)
    dx, dp, dl = g.alloc(xyz.nbytes), g.alloc(4 * n), g.alloc(4 * n)
    dx.upload(xyz)
    for bad in (P, -1):
        ids = np.zeros(n, np.int32)
        ids[n // 2] = bad
        dp.upload(ids)
        with pytest.raises(gahhg_amd.GahhgError):
            g.query_batch_device(dx, dp, n, dl)
    ids = np.full(n, P - 1, np.int32)
    dp.upload(ids)
    g.query_batch_device(dx, dp, n, dl)
    assert (dl.download(np.empty(n, np.int32)) == pairs[P - 1]).all()
    for b in (dx, dp, dl):
        b.free()
    g.close()
