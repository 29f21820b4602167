"""Trace recorder and per-polygon fixed depths on the GPU.

* The traced query returns the same labels as the untraced one, and its leaf
  cell for every point is the leaf the oracle's first-match descent reaches.
* Its top-level cell and level fields agree with the exported hierarchy.
* The build trace's counts agree with the exported hierarchy.
* Fixed depths chosen per polygon of a batch give, for each polygon, the
  hierarchy and labels of the oracle built at that fixed depth.
"""
import sys

import numpy as np
import pytest

import gahhg_amd
from gahhg_amd import datasets
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def roots_and_levels(cells, levels):
    """Top-level ancestor and level of every cell of an exported grid."""
    up = cells[:, 4].copy()
    root = np.arange(len(cells))
    for _ in range(40):
        nxt = np.where(up[root] >= 0, up[root], root)
        if np.array_equal(nxt, root):
            break
        root = nxt
    lev = np.zeros(len(cells), dtype=np.int64)
    for l, L in enumerate(levels):
        lev[L[4]:L[4] + L[5]] = l
    return root, lev


def test_trace_leaf_matches_oracle():
    say("trace: fractal 20k ring, 400k mixed queries")
    w = datasets.fractal_100k(n_queries=400_000, n_outer=20_000)
    g = gahhg_amd.GPISP(device=0, query_bins=256)
    g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    o = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    lab, tr = g.trace_xyz(xyz)
    assert (lab == g.query_xyz(xyz)).all()
    assert np.count_nonzero(lab != o.query_xyz(xyz)) == 0
    leaf = o.query_leaf(xyz)
    assert np.count_nonzero(tr[:, 0] != leaf) == 0
    gg = g.grid()
    root, lev = roots_and_levels(gg["cells"], gg["levels"])
    ok = leaf >= 0
    assert (tr[ok, 1] == root[leaf[ok]]).all()
    assert (tr[ok, 2] == lev[leaf[ok]]).all()
    # table-resolved points end in leaves without refs; points with leaf
    # records end in leaves with refs
    rnum = gg["cells"][:, 3]
    assert (tr[tr[:, 3] == -2, 5] == 0).all()
    assert (rnum[leaf[tr[:, 5] > 0]] > 0).all()
    assert (tr[:, 4] >= 0).all()
    # records = refs + one closing record per chain >= refs + 1
    b = tr[:, 5] > 0
    assert (tr[b, 5] >= rnum[leaf[b]] + 1).all()
    g.close()


def test_build_trace_counts():
    say("build trace: 3-polygon batch")
    ps = datasets.batch_polygons(n_poly=3, seed=9)
    items = [(p["lon"], p["lat"], p["ring_off"]) for p in ps]
    g = gahhg_amd.GPISP(device=0)
    g.set_polygons_lonlat(items).build()
    cells, edges = g.build_trace()
    for k in range(len(items)):
        pg = g.polygon_grid(k)
        c = pg["cells"]
        root, lev = roots_and_levels(c, pg["levels"])
        for l in range(pg["depth"]):
            sel = lev == l
            assert cells[k, l, :, 0].tolist() == np.bincount(root[sel], minlength=42).tolist()
            ref = sel & (c[:, 5] != -1)
            assert cells[k, l, :, 1].tolist() == np.bincount(root[ref], minlength=42).tolist()
        assert cells[k, pg["depth"]:].sum() == 0
        bot = lev == pg["depth"] - 1
        assert cells[k, :, :, 3].sum() == c[bot, 3].sum() == len(pg["refs"])
        n_edges = len(items[k][0]) - (len(items[k][2]) - 1)
        assert edges[k, :, 0].sum() == n_edges
        assert (edges[k, :, 1] >= 0).all() and (edges[k, :, 2] >= 0).all()
        check_edge_aggregates(edges[k], pg, items[k])
    g.close()


def check_edge_aggregates(got, pg, item):
    """Edge aggregates of one polygon against a host restatement: each edge
    goes to the first level-0 cell whose border planes all pass the inside
    test at its first endpoint; per cell the edge count (exact), the sum of
    arc lengths and the sum of turning angles to the ring's next edge, in
    units of 2^-40 rad, each edge truncated (at most one unit per edge
    apart, from the last bit of acos)."""
    pv, pe = ol.Oracle().set_lonlat(*item).polygon()
    a, b = pv[pe[:, 0]], pv[pe[:, 1]]
    c, nrm = pg["cells"], pg["normals"]
    root = np.full(len(pe), -1)
    for i in range(41, -1, -1):  # backwards: the first matching cell wins
        inside = np.ones(len(pe), dtype=bool)
        for k in range(c[i, 1]):
            n = nrm[c[i, 0] + k]
            inside &= ~((n[0] * a[:, 0] + n[1] * a[:, 1] + n[2] * a[:, 2]) < -1e-8)
        root[inside] = i
    ok = root >= 0
    cnt = np.bincount(root[ok], minlength=42)
    assert got[:, 0].astype(np.int64).tolist() == cnt.tolist()
    fix = float(2 ** 40)
    ln = np.arccos(np.clip(a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1] + a[:, 2] * b[:, 2], -1.0, 1.0))
    d = b - a
    ne = np.cross(a, d)
    ne /= np.sqrt((ne * ne).sum(axis=1))[:, None]
    turn = np.zeros(len(pe))
    nxt = np.zeros(len(pe), dtype=bool)
    nxt[:-1] = pe[:-1, 1] == pe[1:, 0]
    dn = (ne[:-1] * ne[1:]).sum(axis=1)
    turn[:-1] = np.where(nxt[:-1], np.arccos(np.clip(dn, -1.0, 1.0)), 0.0)
    for col, val, tol in ((1, ln, 0.0), (2, turn, 1e-9)):
        exp = np.zeros(42, dtype=np.int64)
        np.add.at(exp, root[ok], np.floor(val[ok] * fix).astype(np.int64))
        # the binding returns the sums in rad: the fixed-point integers
        # (below 2^53) times 2^-40, exactly
        dif = np.abs(np.rint(got[:, col] * fix).astype(np.int64) - exp)
        assert (dif <= cnt + 1 + tol * exp).all(), (col, dif.max())


def test_per_polygon_fixed_depths():
    say("fixed depths per polygon: batch of 4")
    ps = datasets.batch_polygons(n_poly=4, seed=12)
    items = [(p["lon"], p["lat"], p["ring_off"]) for p in ps]
    depths = [2, 0, 6, 9]
    g = gahhg_amd.GPISP(device=0)
    g.set_polygons_lonlat(items).set_polygon_depths(depths).build()
    rng = np.random.default_rng(5)
    for k, (lon, lat, roff) in enumerate(items):
        o = ol.Oracle().set_lonlat(lon, lat, roff)
        o.build(max_depth=depths[k], auto_depth=False) if depths[k] else o.build()
        pg = g.polygon_grid(k)
        assert pg["depth"] == o.counts()["depth"]
        for key in ("cells", "borders", "refs"):
            assert np.array_equal(pg[key], o.grid()[key]), (k, key)
        assert np.array_equal(pg["v"].view(np.int64), o.grid()["v"].view(np.int64))
        qlon, qlat = datasets.cap_queries(rng, ps[k]["centre"][0], ps[k]["centre"][1], 1.3 * ps[k]["radius"], 50_000)
        xyz = ol.query_lonlat_to_xyz(qlon, qlat)
        got = g.query_batch_xyz(xyz, np.full(len(qlon), k, dtype=np.int32))
        assert np.count_nonzero(got != o.query_xyz(xyz)) == 0, k
    # setting the polygons again clears the depths: the rule applies
    g.set_polygons_lonlat(items).build()
    o = ol.Oracle().set_lonlat(*items[0]).build()
    assert g.polygon_stats(0)["depth"] == o.counts()["depth"]
    g.close()


def test_depth_argument_errors():
    g = gahhg_amd.GPISP(device=0)
    w = datasets.star_1k(n_queries=10)
    g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"])
    with pytest.raises(gahhg_amd.GahhgError):
        g.set_polygon_depths([40])
    with pytest.raises(gahhg_amd.GahhgError):
        g.set_polygon_depths([-1])
    g.close()


def test_autotune_keeps_labels():
    say("autotune: 5-polygon batch")
    ps = datasets.batch_polygons(n_poly=5, seed=31)
    items = [(p["lon"], p["lat"], p["ring_off"]) for p in ps]
    g = gahhg_amd.GPISP(device=0)
    g.set_polygons_lonlat(items).build()
    rng = np.random.default_rng(4)
    qx, qp = [], []
    for k, p in enumerate(ps):
        lon, lat = datasets.cap_queries(rng, p["centre"][0], p["centre"][1], 1.3 * p["radius"], 60_000)
        qx.append(gahhg_amd.query_lonlat_to_xyz(lon, lat))
        qp.append(np.full(len(lon), k, dtype=np.int32))
    xyz = np.concatenate(qx)
    poly = np.concatenate(qp)
    ref = g.query_batch_xyz(xyz, poly)
    dx, dp, dl = g.alloc(xyz.nbytes), g.alloc(4 * len(poly)), g.alloc(4 * len(poly))
    dx.upload(xyz)
    dp.upload(poly)
    G, B, us = g.autotune_device(dx, dp, len(poly), dl)
    assert G in (0, 4, 8, 16, 32, 64) and B in (64, 128, 256) and us > 0
    assert (dl.download(np.empty(len(poly), dtype=np.int32)) == ref).all()
    assert (g.query_batch_xyz(xyz, poly) == ref).all()  # with the tuned settings
    for b in (dx, dp, dl):
        b.free()
    g.close()


def test_autotune_layout_keeps_labels():
    """The layout autotuner re-packs the query layout per candidate; labels
    stay the oracle's for every candidate and after the choice."""
    say("layout autotune: fractal polygon and a 5-polygon batch")
    w = datasets.fractal_100k(n_queries=400_000, n_outer=20_000)
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    n = len(xyz)
    g = gahhg_amd.GPISP(device=0)
    g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    ref = g.query_xyz(xyz)
    o = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    assert np.count_nonzero(ref != o.query_xyz(xyz)) == 0
    dx, dl = g.alloc(xyz.nbytes), g.alloc(4 * n)
    dx.upload(xyz)
    for passes in (1.0, 1000.0):
        r = g.autotune_layout_device(dx, None, n, dl, passes=passes)
        assert r["query_bins"] in (128, 256) and r["query_sub"] in (1, 8, 16) and r["query_ms"] > 0
        assert (dl.download(np.empty(n, dtype=np.int32)) == ref).all()
        assert (g.query_xyz(xyz) == ref).all()
    for b in (dx, dl):
        b.free()
    g.close()
    ps = datasets.batch_polygons(n_poly=5, seed=31)
    items = [(p["lon"], p["lat"], p["ring_off"]) for p in ps]
    g = gahhg_amd.GPISP(device=0)
    g.set_polygons_lonlat(items).build()
    rng = np.random.default_rng(8)
    lon, lat = datasets.uniform_queries(rng, 300_000)
    xyz = gahhg_amd.query_lonlat_to_xyz(lon, lat)
    poly = rng.integers(0, 5, len(lon)).astype(np.int32)
    ref = g.query_batch_xyz(xyz, poly)
    dx, dp, dl = g.alloc(xyz.nbytes), g.alloc(4 * len(poly)), g.alloc(4 * len(poly))
    dx.upload(xyz)
    dp.upload(poly)
    r = g.autotune_layout_device(dx, dp, len(poly), dl, passes=10.0)
    assert r["query_table_budget"] in (0, 8, 32)
    assert (dl.download(np.empty(len(poly), dtype=np.int32)) == ref).all()
    assert (g.query_batch_xyz(xyz, poly) == ref).all()
    for b in (dx, dp, dl):
        b.free()
    g.close()


def test_grid_save_load(tmp_path):
    """A saved hierarchy loads into a fresh context with identical arrays,
    statistics and labels (single polygon and batch)."""
    import os
    say("grid save/load: NA0 and a 4-polygon batch")
    na0 = os.path.join(os.path.dirname(__file__), "golden", "reference", "NA0.txt")
    ps = datasets.batch_polygons(n_poly=4, seed=41)
    items = [(p["lon"], p["lat"], p["ring_off"]) for p in ps]
    rng = np.random.default_rng(6)
    lon, lat = datasets.uniform_queries(rng, 200_000)
    xyz = gahhg_amd.query_lonlat_to_xyz(lon, lat)
    poly = rng.integers(0, 4, len(lon)).astype(np.int32)
    for case in ("single", "batch"):
        a = gahhg_amd.GPISP(device=0)
        if case == "single":
            a.load_polygon_txt(na0).build()
        else:
            a.set_polygons_lonlat(items).build()
        f = tmp_path / f"{case}.grid"
        a.save_grid(f)
        b = gahhg_amd.GPISP(device=0).load_grid(f)
        assert b.n_polygons == a.n_polygons
        for k in range(a.n_polygons):
            ga, gb = a.polygon_grid(k), b.polygon_grid(k)
            for key in ("cells", "borders", "refs", "levels"):
                assert np.array_equal(ga[key], gb[key]), (case, k, key)
            assert np.array_equal(ga["v"].view(np.int64), gb["v"].view(np.int64))
            sa, sb = a.polygon_stats(k), b.polygon_stats(k)
            for key in ("depth", "gpu_space_bytes", "vnum", "bnum", "cnum", "rnum"):
                assert sa[key] == sb[key], (case, k, key)
        if case == "single":
            assert (a.query_xyz(xyz) == b.query_xyz(xyz)).all()
        else:
            assert (a.query_batch_xyz(xyz, poly) == b.query_batch_xyz(xyz, poly)).all()
        a.close()
        b.close()
    bad = tmp_path / "bad.grid"
    bad.write_bytes(b"not a grid")
    with pytest.raises(gahhg_amd.GahhgError):
        gahhg_amd.GPISP(device=0).load_grid(bad)
