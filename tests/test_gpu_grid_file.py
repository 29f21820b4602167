"""Grid files (gahhg_save_grid / gahhg_load_grid) against the oracle.

F1  a saved hierarchy, loaded into a fresh context, equals the oracle's
    hierarchy array for array (single polygon and batch), and the loaded
    context's labels equal the oracle's;
F2  a file whose arrays would send a kernel out of range (a cell with 9
    borders, a centre child outside the cells, a centre child on its own
    level, a ref past the last edge, a border vertex past the last vertex)
    or that is cut short is refused with GAHHG_E_FORMAT before anything is
    uploaded, and the context keeps working.
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


def test_loaded_grid_equals_oracle(tmp_path):
    ps = datasets.batch_polygons(n_poly=5, seed=41, n_min=50, n_max=40_000)
    items = [(p["lon"], p["lat"], p["ring_off"]) for p in ps]
    rng = np.random.default_rng(6)
    lon, lat = datasets.uniform_queries(rng, 300_000)
    xyz = gahhg_amd.query_lonlat_to_xyz(lon, lat)
    poly = rng.integers(0, len(items), len(lon)).astype(np.int32)
    # single polygon: the reference's NA0
    say("F1 NA0 saved, loaded, against the oracle")
    with gahhg_amd.GPISP(device=0) as a:
        a.load_polygon_txt(os.path.join(REF, "NA0.txt")).build()
        a.save_grid(tmp_path / "na0.grid")
    o = ol.Oracle().load_txt(os.path.join(REF, "NA0.txt")).build()
    with gahhg_amd.GPISP(device=0) as b:
        b.load_grid(tmp_path / "na0.grid")
        compare_grids(b.grid(), o.grid(), o.levels())
        assert b.stats()["gpu_space_bytes"] == o.space_bytes()
        assert np.count_nonzero(b.query_xyz(xyz) != o.query_xyz(xyz)) == 0
    # batch
    say("F1 batch of 5 saved, loaded, against the oracle")
    with gahhg_amd.GPISP(device=0) as a:
        a.set_polygons_lonlat(items).build()
        a.save_grid(tmp_path / "batch.grid")
    with gahhg_amd.GPISP(device=0) as b:
        b.load_grid(tmp_path / "batch.grid")
        got = b.query_batch_xyz(xyz, poly)
        for k, it in enumerate(items):
            ok = ol.Oracle().set_lonlat(*it).build()
            compare_grids(b.polygon_grid(k), ok.grid(), ok.levels())
            sel = poly == k
            assert np.count_nonzero(got[sel] != ok.query_xyz(np.ascontiguousarray(xyz[sel]))) == 0, k


def _counts(g):
    cnt = np.zeros(5, dtype=np.int64)
    g._check(g.L.gahhg_grid_counts(g.h, cnt))
    return [int(x) for x in cnt]  # depth, V, B, C, R


def test_tampered_grid_files_are_refused(tmp_path):
    say("F2 tampered grid files")
    path = tmp_path / "good.grid"
    with gahhg_amd.GPISP(device=0) as a:
        a.load_polygon_txt(os.path.join(REF, "Radiometer.txt")).build()
        a.save_grid(path)
        _, V, B, C, R = _counts(a)
        N = a.stats()["n_edges"]
    raw = bytearray(path.read_bytes())
    size = len(raw)
    # the arrays end the file (grid_arrays order), followed by two int64
    end = size - 16
    refs = end - 4 * R
    c_subcell = refs - 4 * 4 * C  # c_subcell, c_type, c_poly, c_root
    c_bnum = refs - 8 * 4 * C
    b_arrays = c_bnum - 4 * C - 24 * B - 6 * 4 * B  # b_parent .. b_id, then b_n
    b_v0 = b_arrays + 2 * 4 * B

    def put(buf, off, k, val):
        buf[off + 4 * k: off + 4 * k + 4] = np.int32(val).tobytes()

    with gahhg_amd.GPISP(device=0) as a:
        a.load_grid(path)
        sub = a.grid()["cells"][:, 5]
    parent = int(np.nonzero(sub >= 0)[0][0])  # a refined cell
    cases = {
        "nine borders": (c_bnum, 0, 9),
        "centre child outside": (c_subcell, parent, C + 5),
        "centre child on its own level": (c_subcell, parent, parent),
        "ref past the last edge": (refs, 0, N + 100),
        "border vertex past the last vertex": (b_v0, 3, V),
    }
    with gahhg_amd.GPISP(device=0) as g:
        for name, (off, k, val) in cases.items():
            bad = bytearray(raw)
            put(bad, off, k, val)
            p = tmp_path / "bad.grid"
            p.write_bytes(bytes(bad))
            say(f"F2 {name}")
            with pytest.raises(gahhg_amd.GahhgError) as e:
                g.load_grid(p)
            assert e.value.code == -3, name
        p = tmp_path / "short.grid"
        p.write_bytes(bytes(raw[:size - 100]))
        with pytest.raises(gahhg_amd.GahhgError) as e:
            g.load_grid(p)
        assert e.value.code == -3
        # the untouched file still loads and answers like the oracle
        g.load_grid(path)
        o = ol.Oracle().load_txt(os.path.join(REF, "Radiometer.txt")).build()
        lon, lat = ol.load_query_csv(os.path.join(REF, "Radiometer_query.csv"))
        xyz = ol.query_lonlat_to_xyz(lon, lat)
        assert np.count_nonzero(g.query_xyz(xyz) != o.query_xyz(xyz)) == 0
