"""GPU parity at the geometric edge cases (tests/edge_polygons.py): caps
around either pole, an annulus, a ring across the antimeridian, a polygon
1e-5 degrees wide (depth 15), 40 holes, a clockwise ring with collinear
vertices, more than a hemisphere, and inputs where the algorithm alone
defines the labels (latitude rings around the whole sphere, overlapping
holes, a zero-length edge, edges through level-0 cell centres).  Queries
include every vertex, every edge midpoint and points 1e-12 rad either side
of each edge, the poles and the lon = +-180 meridian.

For each: the hierarchy arrays equal the oracle's bitwise, the walk and
classification counters agree, and the labels equal the oracle's (0
mismatches) whether the points are converted on the host or loaded as text
on the device.
"""
import sys

import numpy as np
import pytest

import edge_polygons as ep
import gahhg_amd
from gahhg_amd import datasets
import oracle_lib as ol
from test_gpu_parity import compare_grids

pytestmark = pytest.mark.gpu

CASES = ep.cases()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_edge_case_polygon(case):
    name, lon, lat, off, _ = case
    say("edge case", name)
    g = gahhg_amd.GPISP(device=0)
    g.set_polygon_lonlat(lon, lat, off).build()
    o = ol.Oracle().set_lonlat(lon, lat, off).build()
    compare_grids(g.grid(), o.grid(), o.levels())
    st, d = g.stats(), o.diag()
    for key in ("walk_steps_mark", "walk_steps_dispatch", "walk_partner_missing", "walk_unlocated",
                "undefined_leaves_snapshot", "undefined_leaves_final", "mark_passes"):
        assert st[key] == d[key], key
    assert st["orientation_flipped"] == o.flips
    qlo, qla = ep.feature_points(lon, lat, off, np.random.default_rng(7))
    xyz = ol.query_lonlat_to_xyz(qlo, qla)
    exp = o.query_xyz(xyz)
    assert np.count_nonzero(g.query_lonlat(qlo, qla) != exp) == 0
    v, n = g.parse_query_csv_device(datasets.query_csv_bytes(qlo, qla, decimals=9))
    dl = g.alloc(4 * n)
    g.query_device(v, n, dl)
    lab = dl.download(np.empty(n, dtype=np.int32))
    dxyz = v.download(np.empty((n, 3)))
    assert np.count_nonzero(lab != o.query_xyz(dxyz)) == 0
    dl.free()
    g.close()
