"""Series sweep: every polygon of a series against one query file, as a batch.

The reference runs one GPISP process per polygon file (test/explist.bat:20-643
for the Mars series: ``GPISP.exe /polyfile polygon/Mars/NNN.txt /tpfile
query/Mars_query.csv /reportfile report/GAHHG/Mars/report.txt /resultfile
report/GAHHG/Mars/NNN_result.csv``, one line per file).  Each process loads
its polygon (GPISPPolygon::loadPolyFromTXT, source/GPISPPolygon.cu:32), builds
and classifies the hierarchy, tests every point (GPISP::test), appends one
report block (GPISP::exportStat2TXT, source/GPISPAHHG.cu:2115-2161) and
writes one result file (GPISPTP::exportResult2CSV, source/GPISPTP.cu:81-99).

Here the series is one batch: one build with shared launches, one batched
query over every (polygon, point) pair, then the same report blocks (in file
order) and result files.  A batch has no per-polygon build or query time;
each block reports the polygon's share: build time split by edge count, query
time split evenly (every polygon gets the same points).

    python -m gahhg_amd.sweep --polydir DIR --tpfile Q.csv \\
        [--reportfile R.txt] [--resultdir OUT]
"""
import argparse
import glob
import os
import time

import numpy as np

import gahhg_amd


def series_files(polydir):
    """The polygon files of a series directory, in name order (the order of
    the reference's command list)."""
    files = sorted(glob.glob(os.path.join(polydir, "*.txt")))
    if not files:
        raise FileNotFoundError(f"no polygon files in {polydir}")
    return files


def run(poly_paths, tpfile, reportfile=None, resultdir=None, device=0, **gpisp_kw):
    """Builds every polygon of poly_paths as one batch and tests every point
    of tpfile against each.  Returns a dict with labels (P x Q, int32, the
    library's 0/1/2 = undefined/out/in), inside (P,), depth (P,), n_edges (P,),
    build_ms, query_ms (the batched query, as GPISP::test times it) and
    wall_ms (load + build + query + download, host clock)."""
    poly_paths = [str(x) for x in poly_paths]
    t0 = time.perf_counter()
    lon, lat = gahhg_amd.read_query_csv(tpfile)
    xyz = gahhg_amd.query_lonlat_to_xyz(lon, lat)
    P, Q = len(poly_paths), len(lon)
    g = gahhg_amd.GPISP(device=device, **gpisp_kw)
    try:
        g.load_polygons_txt(poly_paths).build()
        build_ms = g.stats()["pre_ms"]
        poly = np.repeat(np.arange(P, dtype=np.int32), Q)
        labels = g.query_batch_xyz(np.tile(xyz, (P, 1)), poly).reshape(P, Q)
        query_ms = g.stats()["query_ms"]
        wall_ms = (time.perf_counter() - t0) * 1e3
        ps = [g.polygon_stats(k) for k in range(P)]
        inside = np.count_nonzero(labels == gahhg_amd.LABEL_IN, axis=1)
        n_edges = np.array([s["n_edges"] for s in ps], dtype=np.int64)
        if reportfile is not None:
            share = n_edges / max(int(n_edges.sum()), 1)
            qname = os.path.basename(str(tpfile))
            for k, path in enumerate(poly_paths):
                g.write_polygon_report(k, reportfile, os.path.basename(path), qname, Q, int(inside[k]),
                                       query_ms / P, build_ms * float(share[k]))
    finally:
        g.close()
    if resultdir is not None:
        os.makedirs(resultdir, exist_ok=True)
        for k, path in enumerate(poly_paths):
            stem = os.path.splitext(os.path.basename(path))[0]
            gahhg_amd.write_result_csv(os.path.join(resultdir, f"{stem}_result.csv"), labels[k])
    return {"labels": labels, "inside": inside, "depth": np.array([s["depth"] for s in ps]),
            "n_edges": n_edges, "build_ms": build_ms, "query_ms": query_ms, "wall_ms": wall_ms}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--polydir", required=True)
    ap.add_argument("--tpfile", required=True)
    ap.add_argument("--reportfile")
    ap.add_argument("--resultdir")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    files = series_files(a.polydir)
    r = run(files, a.tpfile, a.reportfile, a.resultdir, a.device)
    P, Q = r["labels"].shape
    print(f"{P} polygons x {Q} points: build {r['build_ms']:.2f} ms, query {r['query_ms']:.3f} ms, "
          f"total {r['wall_ms']:.1f} ms; {int(r['inside'].sum())} (polygon, point) pairs inside")


if __name__ == "__main__":
    main()
