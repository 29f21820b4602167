"""Parity at the benchmark configurations' full sizes (the GPU library
against the CPU oracle, bit for bit, density-rule hierarchy).

C2  fractal_100k: 100k-vertex fractal ring + 5 holes of 2k (110,000 edges);
    every grid array and all 10M labels;
C3  fractal_1m: 1M-vertex fractal ring around the pole + 3 holes of 50k
    (1,150,000 edges); every grid array and all 100M labels;
C4  batch_256: the 256-polygon batch (16..200k vertices) in one build and
    one query pass over its 64M Zipf-skewed points; grids of the largest
    and of random polygons, labels of every point of 40 polygons;
C5  fractal_512: 512 fractal polygons (4k..64k outer vertices) in one batch
    with 32M points; grids and labels of a sample of polygons.
"""
import sys

import numpy as np
import pytest

import gahhg_amd
from gahhg_amd import datasets
import oracle_lib as ol
from test_gpu_parity import compare_grids

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_labels_chunked(g_lab, o, xyz, chunk=20_000_000):
    bad = 0
    for s in range(0, len(xyz), chunk):
        bad += int(np.count_nonzero(g_lab[s:s + chunk] != o.query_xyz(np.ascontiguousarray(xyz[s:s + chunk]))))
    return bad


@pytest.mark.timeout(300)
def test_c2_fractal_100k_full():
    say("C2 fractal_100k: generating 10M points")
    w = datasets.fractal_100k()
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    with gahhg_amd.GPISP(device=0) as g:
        g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
        say(f"C2 built: {g.stats()['depth']} levels; querying 10M")
        lab = g.query_xyz(xyz)
        o = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
        assert o.n_edges == 110_000
        compare_grids(g.grid(), o.grid(), o.levels())
        assert g.stats()["gpu_space_bytes"] == o.space_bytes()
    say("C2 oracle labels")
    assert check_labels_chunked(lab, o, xyz) == 0
    assert (lab == 0).sum() == 0


@pytest.mark.timeout(400)
def test_c3_fractal_1m_full():
    say("C3 fractal_1m: generating 100M points")
    w = datasets.fractal_1m()
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    del w["qlon"], w["qlat"]
    say("C3 generated; building")
    with gahhg_amd.GPISP(device=0) as g:
        g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
        st = g.stats()
        say(f"C3 built: {st['depth']} levels in {st['pre_ms']:.1f} ms; querying 100M")
        lab = g.query_xyz(xyz)
        say("C3 oracle build")
        o = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
        assert o.n_edges == 1_150_000
        assert st["depth"] == o.counts()["depth"]
        compare_grids(g.grid(), o.grid(), o.levels())
        assert st["gpu_space_bytes"] == o.space_bytes()
    say("C3 oracle labels (100M)")
    assert len(lab) == 100_000_000
    assert check_labels_chunked(lab, o, xyz) == 0


def _batch_check(name, polys, xyz, qp, sample_grids, sample_labels, **kw):
    items = [(p["lon"], p["lat"], p["ring_off"]) for p in polys]
    with gahhg_amd.GPISP(device=0, **kw) as g:
        g.set_polygons_lonlat(items).build()
        say(f"{name} built: {g.stats()['pre_ms']:.1f} ms; querying {len(qp)} points")
        n = len(qp)
        dx, dp, dl = g.alloc(xyz.nbytes), g.alloc(4 * n), g.alloc(4 * n)
        dx.upload(xyz)
        dp.upload(qp)
        g.query_batch_device(dx, dp, n, dl)
        lab = dl.download(np.empty(n, np.int32))
        for b in (dx, dp, dl):
            b.free()
        for k in sample_grids:
            say(f"{name} grid of polygon {k}")
            o = ol.Oracle().set_lonlat(*items[k]).build()
            compare_grids(g.polygon_grid(k), o.grid(), o.levels())
            assert g.polygon_stats(k)["gpu_space_bytes"] == o.space_bytes()
    order = np.argsort(qp, kind="stable")
    cnt = np.bincount(qp, minlength=len(items))
    start = np.concatenate([[0], np.cumsum(cnt)])
    checked = bad = 0
    for k in sample_labels:
        idx = order[start[k]:start[k + 1]]
        if not len(idx):
            continue
        o = ol.Oracle().set_lonlat(*items[k]).build()
        bad += int(np.count_nonzero(o.query_xyz(np.ascontiguousarray(xyz[idx])) != lab[idx]))
        checked += len(idx)
    say(f"{name}: {checked} labels checked, {bad} mismatches")
    return checked, bad, lab


@pytest.mark.timeout(400)
def test_c4_batch_256_full():
    say("C4 batch_256: generating 64M points")
    w = datasets.batch_256()
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    del w["qlon"], w["qlat"]
    polys, qp = w["polys"], w["qpoly"]
    sizes = np.array([len(p["lon"]) for p in polys])
    cnt = np.bincount(qp, minlength=len(polys))
    rng = np.random.default_rng(44)
    largest = list(np.argsort(-sizes)[:4])
    grids = largest + list(rng.choice(len(polys), 8, replace=False))
    labels = sorted(set(largest + list(np.argsort(-cnt)[:6]) + list(rng.choice(len(polys), 30, replace=False))))
    checked, bad, lab = _batch_check("C4", polys, xyz, qp, grids, labels)
    assert checked >= 1_000_000 and bad == 0
    assert (lab == 0).sum() == 0


@pytest.mark.timeout(400)
def test_c5_fractal_512_full():
    say("C5 fractal_512: generating 32M points")
    w = datasets.fractal_512()
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    del w["qlon"], w["qlat"]
    polys, qp = w["polys"], w["qpoly"]
    sizes = np.array([len(p["lon"]) for p in polys])
    rng = np.random.default_rng(55)
    grids = list(np.argsort(-sizes)[:2]) + list(rng.choice(len(polys), 6, replace=False))
    labels = list(rng.choice(len(polys), 40, replace=False))
    checked, bad, _ = _batch_check("C5", polys, xyz, qp, grids, labels)
    assert checked >= 1_000_000 and bad == 0
