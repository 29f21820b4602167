"""The six Lune/Ketzner-format polygons the reference ships
(test/data_for_Lune/testpolygons: Antarctic, NA7_nohole, Radiometer,
Square_72, Tennessee, Zigzag; 4 to 102,401 edges), loaded through the
library's reader (gahhg_load_polygon_lune) and built on the GPU, against the
oracle fed by the numpy reader of the format:

L1  the hierarchy equals the oracle's array for array (vertices, cells,
    borders, normals, refs bitwise) at the density-rule depth;
L2  the labels of all 30,000 points of the three Lune query files
    (testqueries: Radiometer, Square_72, Tennessee, read by
    gahhg_read_lune_points) and of the file's own reference point equal the
    oracle's, on the start-table path and on the plain descent.

For Antarctic, NA7_nohole and Zigzag the file's reference point (the North
Pole) is outside under the reference's ring-orientation rule: the reference
keeps the side its orientation test normalises to (isPolyClockwise,
source/GPISPPolygon.cu:280-313, applied in loadPolyFromTXT), not the
comparators' convention, so only parity is asserted for it.
"""
import os
import sys

import numpy as np
import pytest

import gahhg_amd
from gahhg_amd import formats
import lune_numpy as ln
import oracle_lib as ol
from test_gpu_parity import compare_grids

pytestmark = pytest.mark.gpu
LUNE = os.path.join(os.path.dirname(__file__), "golden", "lune")
POLYGONS = ["Antarctic.csv.gz", "NA7_nohole.csv.gz", "Radiometer.csv", "Square_72.csv", "Tennessee.csv.gz",
            "Zigzag.csv.gz"]
QUERIES = ["Radiometer_Queries.csv", "Square_72_Queries.csv", "Tennessee_Queries.csv.gz"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def queries(tmp_path_factory):
    d = tmp_path_factory.mktemp("lune")
    parts = [formats.read_lune_points(ln.unpacked(os.path.join(LUNE, q), d)) for q in QUERIES]
    lon = np.concatenate([p[0] for p in parts])
    lat = np.concatenate([p[1] for p in parts])
    assert len(lon) == 30_000
    return d, lon, lat


@pytest.mark.parametrize("name", POLYGONS)
def test_lune_polygon_equals_oracle(name, queries):
    d, qlon, qlat = queries
    path = ln.unpacked(os.path.join(LUNE, name), d)
    lon, lat, off, ref = ln.read_polygon(path)
    o = ol.Oracle().set_lonlat(lon, lat, off).build()
    lon_all = np.append(qlon, ref[0])
    lat_all = np.append(qlat, ref[1])
    xyz = gahhg_amd.query_lonlat_to_xyz(lon_all, lat_all)
    want = o.query_xyz(xyz)
    for layout in (1, 0):
        say(f"L1/L2 {name}: {len(lon) - 1} edges, query_layout={layout}")
        with gahhg_amd.GPISP(device=0, query_layout=layout) as g:
            got_ref = g.load_polygon_lune(path)
            assert got_ref == ref
            g.build()
            compare_grids(g.grid(), o.grid(), o.levels())
            assert g.stats()["gpu_space_bytes"] == o.space_bytes()
            got = g.query_xyz(xyz)
            bad = np.nonzero(got != want)[0]
            assert len(bad) == 0, (name, layout, bad[:10])
            # the host lon/lat entry point converts the same way
            assert np.array_equal(g.query_lonlat(lon_all, lat_all), want)
    say(f"{name}: depth {len(o.levels())}, inside {int(np.sum(want[:-1] == 2))} of 30000, "
        f"reference point label {int(want[-1])}")
