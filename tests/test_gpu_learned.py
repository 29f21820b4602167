"""Learned refinement depth in the build.

* Decision rule: a model whose prediction does not depend on stop/refine
  stops as soon as its window opens (never deeper than the density rule);
  one that predicts a large saving refines to the top of the window (2
  levels past the density rule's stop).  The density rule is untouched when
  no model is set.
* Whatever depth the model picks, the hierarchy and the labels equal the
  oracle's built at that fixed depth (0 mismatches, same hierarchy).
* The features a build uses are the ones the post-build entry point
  returns for training.
"""
import sys

import numpy as np
import pytest

import gahhg_amd
from gahhg_amd import datasets, learned
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def const_model(b3):
    p = np.zeros(learned.N_PARAMS, dtype=np.float32)
    p[5312] = b3
    return p


def refine_model(gain):
    """y = b3 - gain * k: refining always predicts 2^gain times faster points."""
    p = np.zeros(learned.N_PARAMS, dtype=np.float32)
    p[0 * 16 + 1] = 1.0          # hidden-1 unit 0 = relu(k)
    p[1088 + 0 * 64 + 0] = 1.0   # hidden-2 unit 0 = relu(hidden-1 unit 0)
    p[5248 + 0] = -gain          # w3
    p[5312] = 10.0
    return p


def items(n, seed):
    ps = datasets.batch_polygons(n_poly=n, seed=seed)
    return ps, [(p["lon"], p["lat"], p["ring_off"]) for p in ps]


def test_decision_extremes():
    ps, it = items(3, 21)
    cost = [1e6, 1.0, 1.0, 1.0, 1000.0]
    rule = [ol.Oracle().set_lonlat(*x).build(max_depth=9).counts()["depth"] for x in it]
    g = gahhg_amd.GPISP(device=0, max_depth=9)
    # no predicted saving: stops as soon as the window opens, never deeper
    # than the density rule
    g.set_polygons_lonlat(it).set_depth_model(const_model(3.0), cost).build()
    d0 = [g.polygon_stats(k)["depth"] for k in range(3)]
    assert all(1 <= a <= b for a, b in zip(d0, rule)), (d0, rule)
    # a large predicted saving: refines to the top of the window, 2 levels
    # past the density rule's stop (max_depth caps it)
    g.set_depth_model(refine_model(8.0), cost).build()
    d1 = [g.polygon_stats(k)["depth"] for k in range(3)]
    assert d1 == [min(r + 2, 9) for r in rule], (d1, rule)
    # model off: the density rule
    g.set_depth_model(None, None).build()
    assert [g.polygon_stats(k)["depth"] for k in range(3)] == rule
    # fixed depths still win over the model
    g.set_polygons_lonlat(it).set_polygon_depths([2, 0, 4]).set_depth_model(refine_model(8.0), cost).build()
    assert [g.polygon_stats(k)["depth"] for k in range(3)] == [2, min(rule[1] + 2, 9), 4]
    g.close()


def test_learned_hierarchy_matches_oracle_at_chosen_depth():
    say("learned: 6-polygon batch, random model")
    ps, it = items(6, 22)
    params = learned.init_params(5) * 0.5
    params[5312] = 4.0
    g = gahhg_amd.GPISP(device=0)
    rng = np.random.default_rng(9)
    qx, qp = [], []
    for k, p in enumerate(ps):
        lon, lat = datasets.cap_queries(rng, p["centre"][0], p["centre"][1], 1.3 * p["radius"], 40_000)
        qx.append(gahhg_amd.query_lonlat_to_xyz(lon, lat))
        qp.append(np.full(len(lon), k, dtype=np.int32))
    xyz = np.concatenate(qx)
    poly = np.concatenate(qp)
    g.set_polygons_lonlat(it).set_query_hint(xyz, poly, 100.0)
    g.set_depth_model(params, [0.0, 1.0, 5e-3, 2e-2, 5e3]).build()
    got = g.query_batch_xyz(xyz, poly)
    depths = [g.polygon_stats(k)["depth"] for k in range(len(it))]
    say("depths", depths)
    for k in range(len(it)):
        o = ol.Oracle().set_lonlat(*it[k]).build(max_depth=depths[k], auto_depth=False)
        pg = g.polygon_grid(k)
        for key in ("cells", "borders", "refs"):
            assert np.array_equal(pg[key], o.grid()[key]), (k, key)
        sel = poly == k
        assert np.count_nonzero(got[sel] != o.query_xyz(xyz[sel])) == 0, k
    g.close()


def test_features_after_build():
    ps, it = items(2, 23)
    g = gahhg_amd.GPISP(device=0)
    g.set_polygons_lonlat(it).set_polygon_depths([6, 6]).build()
    cells, edges = g.build_trace()
    for level in (0, 2, 4):
        x = g.depth_features(level)
        assert x.shape == (2, 42, 2, 12)
        assert np.all(x[..., 0] == np.float32(level / 16.0))
        assert np.all(x[:, :, 0, 1] == 0) and np.all(x[:, :, 1, 1] == 1)
        for k in range(2):
            n_edges = len(it[k][0]) - (len(it[k][2]) - 1)
            assert np.allclose(x[k, :, :, 10], np.log2(1 + n_edges) / 20)
            assert np.allclose(x[k, :, 0, 2], np.log2(1 + edges[k, :, 0]) / 20, rtol=1e-6)
            assert np.allclose(x[k, :, 0, 3], np.log2(1 + cells[k, level, :, 1]) / 20, rtol=1e-6)
            assert np.allclose(x[k, :, 0, 4], np.log2(1 + cells[k, level, :, 0]) / 20, rtol=1e-6)
            assert abs(x[k, :, 0, 11].sum() - 1.0) < 1e-5  # default point shares
    g.close()
