"""Series sweep: the reference's full Mars series (624 polygon files, each
tested against Mars_query.csv's 10000 points, test/explist.bat:20-643) as one
batch on the GPU.

* Every (polygon, point) label equals the oracle's for that polygon built on
  its own (0 mismatches over 6.24M pairs).
* Each polygon's hierarchy equals the oracle's (every 25th polygon checked
  array by array; all depths checked).
* The report holds one block per polygon in file order with the polygon's
  file name, edge count, depth and inside/outside counts; each result file
  holds that polygon's labels as "1"/"0" lines.
"""
import os
import re
import sys

import numpy as np
import pytest

import gahhg_amd
from gahhg_amd import sweep
import oracle_lib as ol

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
MARS = os.path.join(HERE, "golden", "mars")
QCSV = os.path.join(HERE, "golden", "reference", "Mars_query.csv")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_mars_series_sweep(tmp_path):
    files = sweep.series_files(MARS)
    assert len(files) == 624
    say("sweep: 624 Mars polygons x 10000 points")
    rep = tmp_path / "report.txt"
    r = sweep.run(files, QCSV, rep, tmp_path / "res")
    lab = r["labels"]
    assert lab.shape == (624, 10000)
    lon, lat = ol.load_query_csv(QCSV)
    xyz = ol.query_lonlat_to_xyz(lon, lat)
    g = gahhg_amd.GPISP(device=0).load_polygons_txt(files[::25]).build()
    bad = []
    for k, f in enumerate(files):
        o = ol.Oracle().load_txt(f).build()
        if o.counts()["depth"] != r["depth"][k] or o.n_edges != r["n_edges"][k]:
            bad.append((k, "depth/edges"))
        n = np.count_nonzero(lab[k] != o.query_xyz(xyz))
        if n:
            bad.append((k, n))
        if k % 25 == 0:
            pg = g.polygon_grid(k // 25)
            for key in ("cells", "borders", "refs"):
                if not np.array_equal(pg[key], o.grid()[key]):
                    bad.append((k, key))
    g.close()
    assert not bad, bad[:10]
    assert int(r["inside"].sum()) == int((lab == gahhg_amd.LABEL_IN).sum())

    text = rep.read_text()
    names = re.findall(r"Polygon file: (\S+)", text)
    assert names == [os.path.basename(f) for f in files]
    assert [int(x) for x in re.findall(r"Number of polygon vertice: (\d+)", text)] == r["n_edges"].tolist()
    assert [int(x) for x in re.findall(r"Depth of GAHHG: (\d+)", text)] == r["depth"].tolist()
    io = [(int(a), int(b)) for a, b in re.findall(r"Test result: inside (\d+), outside (\d+)", text)]
    assert io == [(int(i), 10000 - int(i)) for i in r["inside"]]
    assert set(re.findall(r"Query point file: (\S+)", text)) == {"Mars_query.csv"}
    for k in (0, 99, 623):
        stem = os.path.splitext(os.path.basename(files[k]))[0]
        got = (tmp_path / "res" / f"{stem}_result.csv").read_text().split()
        assert got == ["1" if v == gahhg_amd.LABEL_IN else "0" for v in lab[k]]
