"""The grouping sort and the graph-recorded query pass.

S1  the device radix sort (count read from device memory, launches sized
    for a larger bound) equals numpy's stable argsort on the key bits in
    use, for sizes around the 64-pair wave and 4096-pair tile boundaries,
    few and many distinct keys, and keys with bits above end_bit set;
G1  query_graph = 1: every replay of the recorded pass equals the
    kernel-by-kernel query (and the oracle), with other GPU work and host
    synchronisation between replays; a new count, new buffers or a rebuild
    record the pass again.
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


def stable_ref(keys, vals, end_bit):
    masked = keys.astype(np.uint64) & np.uint64((1 << end_bit) - 1)
    o = np.argsort(masked, kind="stable")
    return keys[o], vals[o]


def test_sort_pairs_matches_stable_argsort():
    g = gahhg_amd.GPISP(device=0)
    rng = np.random.default_rng(11)
    sizes = [0, 1, 2, 63, 64, 65, 4095, 4096, 4097, 3 * 4096 + 17, 100_003, 1_500_000]
    for n in sizes:
        for end_bit in (8, 13, 24, 31, 32):
            for kind in ("random", "few", "sorted", "reversed"):
                if kind == "random":
                    keys = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
                elif kind == "few":
                    keys = (rng.integers(0, 3, n) * 0x01010101).astype(np.uint32)
                elif kind == "sorted":
                    keys = np.sort(rng.integers(0, 2 ** 32, n, dtype=np.uint64)).astype(np.uint32)
                else:
                    keys = np.sort(rng.integers(0, 2 ** 32, n, dtype=np.uint64))[::-1].astype(np.uint32)
                vals = np.arange(n, dtype=np.uint32)
                slack = 0 if (n + end_bit) % 2 else 5000
                say(f"sort n={n} end_bit={end_bit} {kind} bound=n+{slack}")
                k, v = g.sort_pairs(keys, vals, end_bit=end_bit, n_bound=n + slack)
                rk, rv = stable_ref(keys, vals, end_bit)
                assert (k == rk).all() and (v == rv).all(), (n, end_bit, kind)
    with pytest.raises(gahhg_amd.GahhgError):
        g.sort_pairs(np.zeros(4, np.uint32), np.zeros(4, np.uint32), end_bit=0)
    with pytest.raises(gahhg_amd.GahhgError):
        g.sort_pairs(np.zeros(4, np.uint32), np.zeros(4, np.uint32), end_bit=8, n_bound=3)
    g.close()


def _batch_items():
    items = []
    for name in ["Square_72.txt", "Radiometer.txt", "NA0.txt", "Mars_300.txt"]:
        lon, lat, roff = [], [], [0]
        with open(os.path.join(REF, name)) as f:
            holes = int(f.readline().split()[2])
            for _ in range(holes + 1):
                cnt = int(f.readline().split()[1])
                for _ in range(cnt):
                    a, b = f.readline().split()[:2]
                    lon.append(float(a))
                    lat.append(float(b))
                roff.append(len(lon))
        items.append((np.array(lon), np.array(lat), np.array(roff)))
    w = datasets.star_1k(n_queries=10)
    items.append((w["lon"], w["lat"], w["ring_off"]))
    w2 = datasets.fractal_100k(n_queries=10, n_outer=8000)
    items.append((w2["lon"], w2["lat"], w2["ring_off"]))
    return items


def test_graph_replays_match_eager_batch():
    items = _batch_items()
    eager = gahhg_amd.GPISP(device=0, query_sort=2)
    graph = gahhg_amd.GPISP(device=0, query_sort=2, query_graph=1)
    for g in (eager, graph):
        g.set_polygons_lonlat(items).build()
    rng = np.random.default_rng(5)
    n = 300_000
    qlon, qlat = datasets.uniform_queries(rng, n)
    xyz = ol.query_lonlat_to_xyz(qlon, qlat)
    poly = rng.integers(0, len(items), n).astype(np.int32)
    ref = eager.query_batch_xyz(xyz, poly)
    d_xyz, d_poly, d_lab = graph.alloc(xyz.nbytes), graph.alloc(poly.nbytes), graph.alloc(4 * n)
    d_xyz.upload(xyz)
    d_poly.upload(poly)
    sentinel = np.full(n, -7, np.int32)
    other_lon, other_lat = datasets.uniform_queries(rng, 200_000)
    other = ol.query_lonlat_to_xyz(other_lon, other_lat)
    other_poly = rng.integers(0, len(items), other.shape[0]).astype(np.int32)
    for rep in range(4):
        say(f"graph replay {rep}")
        d_lab.upload(sentinel)
        graph.query_batch_device(d_xyz, d_poly, n, d_lab)
        got = d_lab.download(np.empty(n, np.int32))
        assert (got == ref).all(), rep
        # other GPU work on both contexts between replays
        eager.query_batch_xyz(other, other_poly)
        graph_other = graph.query_batch_xyz(other, other_poly)
        assert (graph_other == eager.query_batch_xyz(other, other_poly)).all()
    # a smaller count records the pass again
    m = n // 3 + 1
    d_lab.upload(sentinel)
    graph.query_batch_device(d_xyz, d_poly, m, d_lab)
    got = d_lab.download(np.empty(n, np.int32))
    assert (got[:m] == ref[:m]).all() and (got[m:] == -7).all()
    # a rebuild moves the layout: recorded again
    graph.build()
    for rep in range(2):
        d_lab.upload(sentinel)
        graph.query_batch_device(d_xyz, d_poly, n, d_lab)
        assert (d_lab.download(np.empty(n, np.int32)) == ref).all(), rep
    # and the labels are the oracle's
    sel = rng.choice(n, 60_000, replace=False)
    for k, (lon, lat, roff) in enumerate(items):
        o = ol.Oracle().set_lonlat(lon, lat, roff).build()
        s = sel[poly[sel] == k]
        assert np.count_nonzero(ref[s] != o.query_xyz(xyz[s])) == 0, k
    for b in (d_xyz, d_poly, d_lab):
        b.free()
    eager.close()
    graph.close()


def test_graph_single_polygon_with_sub_tables():
    w = datasets.fractal_100k(n_queries=1_200_000, n_outer=20_000)
    eager = gahhg_amd.GPISP(device=0)
    graph = gahhg_amd.GPISP(device=0, query_graph=1)
    for g in (eager, graph):
        g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    n = xyz.shape[0]
    ref = eager.query_xyz(xyz)
    d_xyz, d_lab = graph.alloc(xyz.nbytes), graph.alloc(4 * n)
    d_xyz.upload(xyz)
    for rep in range(3):
        say(f"single-polygon graph replay {rep}")
        d_lab.upload(np.full(n, -7, np.int32))
        graph.query_device(d_xyz, n, d_lab)
        assert (d_lab.download(np.empty(n, np.int32)) == ref).all(), rep
        assert (graph.query_xyz(xyz[: n // 2]) == ref[: n // 2]).all()
    o = ol.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"]).build()
    sel = np.random.default_rng(1).choice(n, 100_000, replace=False)
    assert np.count_nonzero(ref[sel] != o.query_xyz(xyz[sel])) == 0
    d_xyz.free()
    d_lab.free()
    eager.close()
    graph.close()
