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


def test_command_list_runs(tmp_path):
    """A command list in the reference's syntax (test/explist.bat's `various`,
    Mars and Tennessee lines) laid out as the reference's test folder: every
    run that loads gives the oracle's labels for its own polygon and query
    file; the NA0 run, whose query file NA_query.csv the reference does not
    ship, is skipped with nothing written, as the reference's process exits."""
    import shutil
    gold = os.path.join(HERE, "golden")
    ref = os.path.join(gold, "reference")
    for d in ("polygon/various", "polygon/Mars", "polygon/Tennessee", "query"):
        (tmp_path / d).mkdir(parents=True)
    for n in ("Square_72", "Radiometer", "Tennessee", "NA0"):
        shutil.copy(os.path.join(ref, n + ".txt"), tmp_path / "polygon" / "various")
    for n in ("Square_72", "Radiometer", "Tennessee", "Mars"):
        shutil.copy(os.path.join(ref, n + "_query.csv"), tmp_path / "query")
    mars = ["001", "417", "624"]
    for n in mars:
        shutil.copy(os.path.join(gold, "mars", n + ".txt"), tmp_path / "polygon" / "Mars")
    tn = ["00005", "00500", "01000", "01500", "02000", "02500", "02999"]
    for n in tn:
        shutil.copy(os.path.join(ref, f"Tennessee_Tennessee_{n}.txt"),
                    tmp_path / "polygon" / "Tennessee" / f"Tennessee_{n}.txt")
    lines = ["md report", "md report\\GAHHG\\various", "md report\\GAHHG\\Mars", "md report\\GAHHG\\Tennessee"]
    various = [("Square_72", "Square_72"), ("Radiometer", "Radiometer"), ("Tennessee", "Tennessee"), ("NA0", "NA")]
    for p, q in various:
        lines.append(f".\\GPISP.exe /polyfile .\\polygon\\various\\{p}.txt /tpfile .\\query\\{q}_query.csv  "
                     f"/reportfile .\\report\\GAHHG\\various\\report.txt   /resultfile .\\report\\GAHHG\\various\\{p}_result.csv ")
    for n in mars:
        lines.append(f".\\GPISP.exe /polyfile .\\polygon\\Mars\\{n}.txt /tpfile .\\query\\Mars_query.csv "
                     f"/reportfile .\\report\\GAHHG\\Mars\\report.txt /resultfile .\\report\\GAHHG\\Mars\\{n}_result.csv")
    for n in tn:
        lines.append(f".\\GPISP.exe /polyfile .\\polygon\\Tennessee\\Tennessee_{n}.txt /tpfile .\\query\\Tennessee_query.csv "
                     f"/reportfile .\\report\\GAHHG\\Tennessee\\report.txt "
                     f"/resultfile .\\report\\GAHHG\\Tennessee\\Tennessee_{n}_result.csv")
    lst = tmp_path / "explist.bat"
    lst.write_text("\r\n".join(lines) + "\r\n")
    say("command list: 4 various + 3 Mars + 7 Tennessee runs")
    jobs, dirs = sweep.parse_command_list(lst)
    assert len(jobs) == 14 and len(dirs) == 4
    for d in dirs:
        os.makedirs(d, exist_ok=True)
    r = sweep.run_jobs(jobs)
    assert [i for i, _ in r["skipped"]] == [3] and "NA_query.csv" in r["skipped"][0][1]
    assert r["runs"][3] is None
    for i, j in enumerate(jobs):
        if i == 3:
            assert not os.path.exists(j.resultfile)
            continue
        lon, lat = ol.load_query_csv(j.tpfile)
        o = ol.Oracle().load_txt(j.polyfile).build()
        run = r["runs"][i]
        assert np.count_nonzero(run["labels"] != o.query_xyz(ol.query_lonlat_to_xyz(lon, lat))) == 0, j.polyfile
        assert run["depth"] == o.counts()["depth"]
        got = open(j.resultfile).read().split()
        assert got == ["1" if v == gahhg_amd.LABEL_IN else "0" for v in run["labels"]]
    rep = tmp_path / "report" / "GAHHG"
    assert re.findall(r"Polygon file: (\S+)", (rep / "various" / "report.txt").read_text()) == \
        ["Square_72.txt", "Radiometer.txt", "Tennessee.txt"]
    assert re.findall(r"Polygon file: (\S+)", (rep / "Mars" / "report.txt").read_text()) == [n + ".txt" for n in mars]
    text = (rep / "Tennessee" / "report.txt").read_text()
    assert re.findall(r"Polygon file: (\S+)", text) == [f"Tennessee_{n}.txt" for n in tn]
    io = [(int(a), int(b)) for a, b in re.findall(r"Test result: inside (\d+), outside (\d+)", text)]
    assert io == [(r["runs"][7 + k]["inside"], len(r["runs"][7 + k]["labels"]) - r["runs"][7 + k]["inside"])
                  for k in range(7)]
