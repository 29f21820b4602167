"""The allocation-free query entry point (gahhg_query_batch_xyz_device_noalloc).

NA1  single polygon, grouped pass (start table, compaction, radix sort):
     labels equal gahhg_query_batch_xyz_device's bit for bit on two calls in
     a row, the second with fewer points; the context's device memory does
     not grow across them (the allocating form's does);
NA2  a batch with polygon ids, the same comparison, plus the oracle;
NA3  more than GAHHG_NOALLOC_MAX_POINTS points is refused (GAHHG_E_LIMIT)
     before anything runs.
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


def run_both(g, xyz, poly, n):
    """Labels of the first n points from the allocating and the
    allocation-free entry points (same device buffers of points)."""
    d_xyz = g.alloc(xyz.nbytes)
    d_xyz.upload(xyz)
    d_poly = None
    if poly is not None:
        d_poly = g.alloc(poly.nbytes)
        d_poly.upload(poly)
    d_a = g.alloc(4 * len(xyz))
    d_b = g.alloc(4 * len(xyz))
    g.query_device_noalloc(d_xyz, d_poly, n, d_b)
    st_b = g.stats()
    if poly is None:
        g.query_device(d_xyz, n, d_a)
    else:
        g.query_batch_device(d_xyz, d_poly, n, d_a)
    st_a = g.stats()
    a = d_a.download(np.empty(len(xyz), np.int32))[:n]
    b = d_b.download(np.empty(len(xyz), np.int32))[:n]
    for d in (d_xyz, d_a, d_b) + ((d_poly,) if d_poly is not None else ()):
        d.free()
    return a, b, st_a, st_b


def test_noalloc_single_polygon_two_calls():
    w = datasets.star_1k(n_queries=2_000_000, seed=1)
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    with gahhg_amd.GPISP(device=0) as g:
        g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
        before = g.stats()["device_bytes"]
        d_xyz = g.alloc(xyz.nbytes)
        d_xyz.upload(xyz)
        d_lab = g.alloc(4 * len(xyz))
        outs = []
        for n in (2_000_000, 1_300_000):  # both above the grouping threshold, the second smaller
            say(f"NA1 noalloc star_1k n={n}")
            g.query_device_noalloc(d_xyz, None, n, d_lab)
            outs.append((n, d_lab.download(np.empty(len(xyz), np.int32))[:n], g.stats()))
        assert g.stats()["device_bytes"] == before, "the allocation-free pass grew the context's buffers"
        for n, lab, st in outs:
            say(f"NA1 allocating star_1k n={n}")
            g.query_device(d_xyz, n, d_lab)
            ref = d_lab.download(np.empty(len(xyz), np.int32))[:n]
            assert np.array_equal(lab, ref), n
            assert st["inside"] == int((ref == 2).sum()) and st["n_queries"] == n
        assert g.stats()["device_bytes"] > before  # the allocating form grows its buffers
        d_xyz.free()
        d_lab.free()
    o = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    assert np.array_equal(outs[1][1], o.query_xyz(np.ascontiguousarray(xyz[:outs[1][0]])))


def test_noalloc_batch_two_calls():
    polys = datasets.batch_polygons(n_poly=8, seed=21, n_min=200, n_max=20_000)
    rng = np.random.default_rng(5)
    n = 1_500_000
    pid = rng.integers(0, len(polys), n).astype(np.int32)
    lon = np.empty(n)
    lat = np.empty(n)
    for k, c in enumerate(polys):
        m = pid == k
        lon[m], lat[m] = datasets.cap_queries(rng, c["centre"][0], c["centre"][1], 1.3 * c["radius"], int(m.sum()))
    xyz = gahhg_amd.query_lonlat_to_xyz(lon, lat)
    with gahhg_amd.GPISP(device=0, query_sort=2) as g:
        g.set_polygons_lonlat([(c["lon"], c["lat"], c["ring_off"]) for c in polys]).build()
        for m in (n, 400_000):
            say(f"NA2 batch of 8 n={m}")
            a, b, st_a, st_b = run_both(g, xyz, pid, m)
            assert np.array_equal(a, b), m
            assert st_a["inside"] == st_b["inside"] and st_a["unlocated"] == st_b["unlocated"]
        sub = np.arange(0, 400_000, 7)
        ref = np.empty(len(sub), np.int32)
        for k, c in enumerate(polys):
            m = pid[sub] == k
            o = ol.Oracle().set_lonlat(c["lon"], c["lat"], c["ring_off"]).build()
            ref[m] = o.query_xyz(np.ascontiguousarray(xyz[sub][m]))
        assert np.array_equal(b[sub], ref)


def test_noalloc_refuses_too_many_points():
    w = datasets.star_1k(n_queries=1000, seed=1)
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    with gahhg_amd.GPISP(device=0) as g:
        g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
        d_xyz = g.alloc(xyz.nbytes)
        d_xyz.upload(xyz)
        d_lab = g.alloc(4 * len(xyz))
        say("NA3 refuse n = max + 1 (checked on the host before any launch)")
        with pytest.raises(gahhg_amd.GahhgError) as e:
            g.query_device_noalloc(d_xyz, None, gahhg_amd.NOALLOC_MAX_POINTS + 1, d_lab)
        assert e.value.code == -7
        g.query_device_noalloc(d_xyz, None, len(xyz), d_lab)  # the context still works
        lab = d_lab.download(np.empty(len(xyz), np.int32))
        assert np.array_equal(lab, g.query_xyz(xyz))
