"""The device-driven build (build_device=1: the density rule, the next
level's layout, the dispatch sizes and the classification fixpoint all
decided on the GPU, one stream of launches with no host wait between levels)
and its recorded HIP graph (build_graph=1), against the oracle:

D1  the host-driven loop, the device-driven loop run eagerly and the graph
    replays all give the oracle's grid (every array bitwise) and labels, for
    a reference polygon, a 1200-edge star and a batch of eight polygons;
    the second build of a polygon set records the graph (a one-off build
    is not recorded) and the third and later builds are replays;
D2  an arena too small for the hierarchy -- a 4-edge polygon at fixed depth
    9 needs far more cells, borders and refs than the reference's initial
    capacities (createAHHGInit, source/GPISPAHHG.cu:1949-1953) -- is found
    on the device; the build grows it and runs again, and equals the oracle;
D3  a replay whose polygons need more levels than it recorded (depth 3
    recorded, depth 7 asked) stops short on the device, is redone eagerly
    and equals the oracle at the new depth;
D4  a polygon set replaced by another of the same shape (same edge count,
    new vertices) builds the new polygon's grid, replayed or not.
"""
import os
import sys

import numpy as np
import pytest

import gahhg_amd
from gahhg_amd import datasets
import oracle_lib as ol
from test_gpu_parity import compare_grids

pytestmark = pytest.mark.gpu
REF = os.path.join(os.path.dirname(__file__), "golden", "reference")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_polygon(g, k, o, xyz):
    pg = g.polygon_grid(k)
    og = o.grid()
    assert pg["depth"] == o.counts()["depth"]
    for key in ("cells", "borders", "refs"):
        assert np.array_equal(pg[key], og[key]), (k, key)
    assert np.array_equal(pg["v"].view(np.int64), og["v"].view(np.int64)), k
    assert np.array_equal(pg["normals"].view(np.int64), og["normals"].view(np.int64)), k
    got = g.query_batch_xyz(xyz, np.full(len(xyz), k, dtype=np.int32))
    assert np.count_nonzero(got != o.query_xyz(xyz)) == 0, k


@pytest.mark.parametrize("mode", ["host", "device", "graph"])
def test_build_modes_equal_oracle(mode):
    kw = {"host": dict(build_device=0), "device": dict(build_graph=0), "graph": {}}[mode]
    rng = np.random.default_rng(3)
    lon, lat = datasets.uniform_queries(rng, 100_000)
    xyz = gahhg_amd.query_lonlat_to_xyz(lon, lat)
    # a reference polygon and the star, one at a time
    star = datasets.star_1k(n_queries=10)
    singles = [("NA0", ol.Oracle().load_txt(os.path.join(REF, "NA0.txt")).build(),
                lambda g: g.load_polygon_txt(os.path.join(REF, "NA0.txt"))),
               ("star_1k", ol.Oracle().set_lonlat(star["lon"], star["lat"], star["ring_off"]).build(),
                lambda g: g.set_polygon_lonlat(star["lon"], star["lat"], star["ring_off"]))]
    for name, o, load in singles:
        with gahhg_amd.GPISP(device=0, **kw) as g:
            load(g)
            for rep in range(3):
                say(f"D1 {mode} {name} build {rep}")
                g.build()
                st = g.stats()
                assert st["build_replayed"] == (1 if mode == "graph" and rep > 1 else 0)
                assert st["build_attempts"] == (0 if mode == "host" else 1)
                compare_grids(g.grid(), o.grid(), o.levels())
                assert np.count_nonzero(g.query_xyz(xyz) != o.query_xyz(xyz)) == 0
    # a batch
    ps = datasets.batch_polygons(n_poly=8, seed=21, n_min=50, n_max=20_000)
    items = [(p["lon"], p["lat"], p["ring_off"]) for p in ps]
    oracles = [ol.Oracle().set_lonlat(*it).build() for it in items]
    with gahhg_amd.GPISP(device=0, **kw) as g:
        g.set_polygons_lonlat(items)
        for rep in range(3):
            say(f"D1 {mode} batch of 8 build {rep}")
            g.build()
            assert g.stats()["build_replayed"] == (1 if mode == "graph" and rep > 1 else 0)
            for k, o in enumerate(oracles):
                check_polygon(g, k, o, xyz[:20_000])


def test_arena_grows_on_the_device():
    path = os.path.join(REF, "Square_72.txt")
    o = ol.Oracle().load_txt(path).build(max_depth=9, auto_depth=False)
    cnt = o.counts()
    assert cnt["C"] > 4 * (4 + 84 + 42) and cnt["R"] > 1024  # beyond the initial capacities
    lon, lat = ol.load_query_csv(os.path.join(REF, "Square_72_query.csv"))
    xyz = ol.query_lonlat_to_xyz(lon, lat)
    with gahhg_amd.GPISP(device=0, max_depth=9, auto_depth=False) as g:
        g.load_polygon_txt(path)
        say("D2 first build: the arena grows")
        g.build()
        assert g.stats()["build_attempts"] >= 2
        compare_grids(g.grid(), o.grid(), o.levels())
        assert np.count_nonzero(g.query_xyz(xyz) != o.query_xyz(xyz)) == 0
        say("D2 rebuild: sized from the last build, recorded")
        g.build()
        st = g.stats()
        assert st["build_attempts"] == 1 and st["build_replayed"] == 0
        compare_grids(g.grid(), o.grid(), o.levels())
        say("D2 rebuild: replayed")
        g.build()
        st = g.stats()
        assert st["build_attempts"] == 1 and st["build_replayed"] == 1
        compare_grids(g.grid(), o.grid(), o.levels())


def test_replay_needing_more_levels_is_redone():
    star = datasets.star_1k(n_queries=10)
    it = (star["lon"], star["lat"], star["ring_off"])
    o7 = ol.Oracle().set_lonlat(*it).build(max_depth=7, auto_depth=False)
    with gahhg_amd.GPISP(device=0) as g:
        g.set_polygons_lonlat([it]).set_polygon_depths([3])
        g.build()
        g.build()
        g.build()
        assert g.stats()["build_replayed"] == 1  # a graph of 3 levels
        say("D3 depth 3 recorded, depth 7 asked")
        g.set_polygon_depths([7]).build()
        st = g.stats()
        # the replay stops short; the eager run after it may also have to
        # grow the arena sized from the depth-3 build
        assert st["build_replayed"] == 0 and st["build_attempts"] >= 2
        pg = g.polygon_grid(0)
        assert pg["depth"] == 7
        for key in ("cells", "borders", "refs"):
            assert np.array_equal(pg[key], o7.grid()[key]), key
        g.build()
        assert g.stats()["build_replayed"] == 1


def test_same_shape_new_vertices():
    rng = np.random.default_rng(9)
    a = datasets.star_1k(n_queries=10)
    lon_b = np.array(a["lon"], dtype=np.float64)
    lat_b = np.array(a["lat"], dtype=np.float64)
    # the same ring moved and rotated: same edge count, new vertices
    lon_b = (lon_b + 47.0 + 180.0) % 360.0 - 180.0
    lat_b = np.clip(-lat_b * 0.8 + 5.0, -89.0, 89.0)
    lon_b[-1], lat_b[-1] = lon_b[0], lat_b[0]
    lon, lat = datasets.uniform_queries(rng, 200_000)
    xyz = gahhg_amd.query_lonlat_to_xyz(lon, lat)
    oa = ol.Oracle().set_lonlat(a["lon"], a["lat"], a["ring_off"]).build()
    ob = ol.Oracle().set_lonlat(lon_b, lat_b, a["ring_off"]).build()
    with gahhg_amd.GPISP(device=0) as g:
        for rep, (lo, la, o) in enumerate([(a["lon"], a["lat"], oa), (lon_b, lat_b, ob), (a["lon"], a["lat"], oa)]):
            say(f"D4 polygon {rep}")
            g.set_polygon_lonlat(lo, la, a["ring_off"])
            for _ in range(3):
                g.build()
                compare_grids(g.grid(), o.grid(), o.levels())
                assert np.count_nonzero(g.query_xyz(xyz) != o.query_xyz(xyz)) == 0
