"""Learned start-table depth: every start-table candidate gives the labels of
the density-rule hierarchy (the oracle's), alone and in batches, with and
without the model; the table the build reports is the one asked for.

L1  one polygon, each of the 8 candidates fixed: labels equal the oracle's on
    uniform and near-edge points, table layout as requested;
L2  a batch with per-polygon candidates 0..4: labels equal the per-polygon
    oracle, per-polygon bins as requested; the polygons without a table
    (candidate 0) start most points from the shared level-0 lookup (a
    level-0 cell) instead of the level-0 scan, with the same labels;
L3  the table model (trained one when present, else random weights) with a
    query hint: labels equal the oracle's; the features are finite.
"""
import os
import sys

import numpy as np
import pytest

import gahhg_amd
from gahhg_amd import datasets, learned_tables
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mixed_points(w, n, seed):
    rng = np.random.default_rng(seed)
    lon, lat = datasets.mixed_queries(rng, w["lon"], w["lat"], w["ring_off"], n, 0.5)
    return gahhg_amd.query_lonlat_to_xyz(lon, lat)


def test_every_candidate_same_labels_single():
    w = datasets.fractal_100k(n_queries=10, n_outer=20_000)
    xyz = mixed_points(w, 1_200_000, 3)
    o = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    ref = o.query_xyz(xyz)
    with gahhg_amd.GPISP(device=0, query_sub=16) as g:
        g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"])
        for c in range(len(gahhg_amd.TABLE_BINS)):
            say(f"L1 candidate {c}: bins {gahhg_amd.TABLE_BINS[c]} sub {gahhg_amd.TABLE_SUB[c]}")
            g.set_table_choice([c]).build()
            b, s = g.table_layout()[0]
            assert b == gahhg_amd.TABLE_BINS[c] and 1 <= s <= gahhg_amd.TABLE_SUB[c]
            assert g.stats()["layout_ms"] > 0
            assert np.count_nonzero(g.query_xyz(xyz) != ref) == 0, c


def test_batch_candidates_same_labels():
    polys = datasets.batch_polygons(n_poly=10, seed=33, n_min=100, n_max=30_000)
    items = [(p["lon"], p["lat"], p["ring_off"]) for p in polys]
    rng = np.random.default_rng(8)
    n = 1_500_000
    pid = rng.integers(0, len(polys), n).astype(np.int32)
    lon, lat = np.empty(n), np.empty(n)
    for k, p in enumerate(polys):
        m = pid == k
        lon[m], lat[m] = datasets.cap_queries(rng, p["centre"][0], p["centre"][1], 1.3 * p["radius"], int(m.sum()))
    xyz = gahhg_amd.query_lonlat_to_xyz(lon, lat)
    cand = np.arange(len(polys), dtype=np.int32) % gahhg_amd.TABLE_CAND_BATCH
    with gahhg_amd.GPISP(device=0) as g:
        g.set_polygons_lonlat(items).set_table_choice(cand).build()
        lay = g.table_layout()
        assert [int(b) for b in lay[:, 0]] == [gahhg_amd.TABLE_BINS[c] for c in cand]
        say("L2 batch of 10 with candidates 0..4")
        got = g.query_batch_xyz(xyz, pid)
        say("L2 traced: start levels of the table-less polygons")
        tlab, tr = g.trace_xyz(xyz, pid)
        assert (tlab == got).all()
        none = cand[pid] == 0
        assert none.any()
        start = tr[none, 3]
        # level 0 = the shared lookup's cell, -1 = the level-0 scan; a bin
        # straddling level-0 cells falls back to the scan
        assert np.count_nonzero(start == 0) >= 0.8 * np.count_nonzero(start >= -1), np.unique(start)
    for k, it in enumerate(items):
        sel = pid == k
        ref = ol.Oracle().set_lonlat(*it).build().query_xyz(np.ascontiguousarray(xyz[sel]))
        assert np.count_nonzero(got[sel] != ref) == 0, k
    with gahhg_amd.GPISP(device=0) as g:  # no polygon gets a table
        g.set_polygons_lonlat(items).set_table_choice(np.zeros(len(items), np.int32)).build()
        assert (g.table_layout()[:, 0] == 0).all()
        say("L2 batch of 10 without tables")
        assert (g.query_batch_xyz(xyz, pid) == got).all()


def test_model_choice_keeps_labels():
    if os.path.exists(learned_tables.MODEL_PATH):
        params, cost, _ = learned_tables.load_model()
    else:
        params = gahhg_amd.learned.init_params(3)
        cost = {"a_ns": [1e3 * (c + 1) for c in range(8)], "b_ns_per_edge": [0.1 * c for c in range(8)]}
    w = datasets.fractal_100k(n_queries=10, n_outer=20_000)
    xyz = mixed_points(w, 1_100_000, 4)
    ref = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build().query_xyz(xyz)
    for npts in (1e4, 1e8):
        with gahhg_amd.GPISP(device=0, query_sub=16) as g:
            g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"])
            learned_tables.apply(g, (params, cost), xyz, None, npts)
            g.build()
            say(f"L3 model at {npts:.0e} expected points: table {g.table_layout()[0].tolist()}")
            assert np.isfinite(g.table_features()).all()
            assert np.count_nonzero(g.query_xyz(xyz) != ref) == 0
