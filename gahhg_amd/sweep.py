"""Series sweep: many (polygon file, query file) runs as one batch.

The reference runs one GPISP process per polygon file from a command list
(test/explist.bat: lines 6-18 the `various` set, 20-643 the Mars series, one
``GPISP.exe /polyfile polygon/Mars/NNN.txt /tpfile query/Mars_query.csv
/reportfile report/GAHHG/Mars/report.txt /resultfile
report/GAHHG/Mars/NNN_result.csv`` per file, 645-3142 the Tennessee series).
Each process loads its polygon (GPISPPolygon::loadPolyFromTXT,
source/GPISPPolygon.cu:32) and points (GPISPTP::loadTPFromCSV), returns -1
without output when either fails (source/GPISP.cu:47-55), builds and
classifies the hierarchy, tests every point (GPISP::test), appends one report
block (GPISP::exportStat2TXT, source/GPISPAHHG.cu:2115-2161) and writes one
result file (GPISPTP::exportResult2CSV, source/GPISPTP.cu:81-99).

Here the runs of a list become one batch: every polygon built with shared
launches, every run's points tested in one batched query, then the same
report blocks (in list order, each appended to its run's report file) and
result files.  A run whose polygon or query file does not load is skipped
with its reason, as the reference's process exits.  A batch has no per-run
build or query time: each block reports the run's share, build time split by
edge count and query time by point count.

    python -m gahhg_amd.sweep --explist test/explist.bat [--base test]
    python -m gahhg_amd.sweep --polydir DIR --tpfile Q.csv \\
        [--reportfile R.txt] [--resultdir OUT]
"""
import argparse
import glob
import os
import shlex
import time
from dataclasses import dataclass

import numpy as np

import gahhg_amd

# GPISPParam defaults (source/GPISPParam.h:66-77), used for a flag a line omits
DEFAULTS = {"polyfile": "NA0.txt", "tpfile": "NA_query.csv", "reportfile": "report.txt",
            "resultfile": "result.csv"}


@dataclass
class Job:
    polyfile: str
    tpfile: str
    reportfile: str = None
    resultfile: str = None


def _local(path, base):
    path = path.replace("\\", "/")
    if path.startswith("./"):
        path = path[2:]
    return path if os.path.isabs(path) else os.path.join(base, path)


def parse_command_list(path, base=None):
    """The GPISP runs of a command list in the reference's syntax (one
    ``... /polyfile P /tpfile Q /reportfile R /resultfile S`` per line, paths
    relative to base, default the list's directory; Windows separators
    accepted).  ``md DIR`` lines are returned as directories to create;
    other lines are ignored.  Returns (jobs, dirs)."""
    base = os.path.dirname(os.path.abspath(path)) if base is None else base
    jobs, dirs = [], []
    with open(path, encoding="utf-8", errors="replace") as f:
        for line in f:
            tok = shlex.split(line.replace("\\", "/"), posix=True)
            if not tok:
                continue
            if tok[0].lower() == "md" and len(tok) > 1:
                dirs.append(_local(tok[1], base))
                continue
            if "/polyfile" not in tok:
                continue
            val = dict(DEFAULTS)
            i = 1
            while i < len(tok):
                key = tok[i][1:] if tok[i].startswith("/") else None
                if key in val and i + 1 < len(tok):
                    val[key] = tok[i + 1]
                    i += 1
                i += 1
            jobs.append(Job(*(_local(val[k], base) for k in ("polyfile", "tpfile", "reportfile", "resultfile"))))
    return jobs, dirs


def series_files(polydir):
    """The polygon files of a series directory, in name order (the order of
    the reference's command list)."""
    files = sorted(glob.glob(os.path.join(polydir, "*.txt")))
    if not files:
        raise FileNotFoundError(f"no polygon files in {polydir}")
    return files


def series_jobs(poly_paths, tpfile, reportfile=None, resultdir=None):
    """One run per polygon file, all against tpfile; result files named
    <polygon stem>_result.csv in resultdir."""
    out = []
    for p in poly_paths:
        res = None
        if resultdir is not None:
            res = os.path.join(resultdir, os.path.splitext(os.path.basename(p))[0] + "_result.csv")
        out.append(Job(str(p), str(tpfile), None if reportfile is None else str(reportfile), res))
    return out


def _load_polygons(g, paths):
    """Loads the batch; on a failure finds the files that do not load (each
    on its own) and returns their reasons so that the rest can run."""
    try:
        g.load_polygons_txt(paths)
        return {}
    except gahhg_amd.GahhgError:
        pass
    bad = {}
    for p in paths:
        try:
            g.load_polygon_txt(p)
        except gahhg_amd.GahhgError as e:
            bad[p] = str(e)
    return bad


def run_jobs(jobs, device=0, **gpisp_kw):
    """Runs a list of jobs as one batch.  Returns a dict: runs (per job in
    order: None if skipped, else labels (int32, the library's 0/1/2 =
    undefined/out/in), inside, depth, n_edges), skipped [(index, reason)],
    build_ms, query_ms (the batched query, as GPISP::test times it) and
    wall_ms (file reads + build + query + label download, host clock; the
    report and result files are written afterwards)."""
    t0 = time.perf_counter()
    skipped = []
    points = {}
    for q in sorted({j.tpfile for j in jobs}):
        try:
            lon, lat = gahhg_amd.read_query_csv(q)
            points[q] = gahhg_amd.query_lonlat_to_xyz(lon, lat)
        except gahhg_amd.GahhgError as e:
            points[q] = str(e)
    live = []
    for i, j in enumerate(jobs):
        if isinstance(points[j.tpfile], str):
            skipped.append((i, points[j.tpfile]))
        elif not os.path.isfile(j.polyfile):
            skipped.append((i, f"cannot open {j.polyfile}"))
        else:
            live.append(i)
    runs = [None] * len(jobs)
    out = {"runs": runs, "skipped": skipped, "build_ms": 0.0, "query_ms": 0.0}
    g = gahhg_amd.GPISP(device=device, **gpisp_kw)
    try:
        bad = _load_polygons(g, [jobs[i].polyfile for i in live]) if live else {}
        if bad:
            skipped += [(i, bad[jobs[i].polyfile]) for i in live if jobs[i].polyfile in bad]
            live = [i for i in live if jobs[i].polyfile not in bad]
            if live:
                g.load_polygons_txt([jobs[i].polyfile for i in live])
        skipped.sort()
        if not live:
            out["wall_ms"] = (time.perf_counter() - t0) * 1e3
            return out
        g.build()
        out["build_ms"] = build_ms = g.stats()["pre_ms"]
        cnt = np.array([len(points[jobs[i].tpfile]) for i in live], dtype=np.int64)
        tps = {jobs[i].tpfile for i in live}
        if len(tps) == 1:
            # a series (one query file for every polygon): the points go to
            # the device once and every polygon answers all of them
            labels = g.query_pairs_xyz(points[tps.pop()]).reshape(-1)
        else:
            xyz = np.concatenate([points[jobs[i].tpfile] for i in live])
            poly = np.repeat(np.arange(len(live), dtype=np.int32), cnt)
            labels = g.query_batch_xyz(xyz, poly)
        out["query_ms"] = query_ms = g.stats()["query_ms"]
        out["wall_ms"] = (time.perf_counter() - t0) * 1e3
        off = np.concatenate([[0], np.cumsum(cnt)])
        ps = [g.polygon_stats(k) for k in range(len(live))]
        n_edges = np.array([s["n_edges"] for s in ps], dtype=np.int64)
        e_share = n_edges / max(int(n_edges.sum()), 1)
        q_share = cnt / max(int(cnt.sum()), 1)
        for k, i in enumerate(live):
            lab = labels[off[k]:off[k + 1]]
            inside = int(np.count_nonzero(lab == gahhg_amd.LABEL_IN))
            runs[i] = {"labels": lab, "inside": inside, "depth": ps[k]["depth"], "n_edges": int(n_edges[k])}
            j = jobs[i]
            if j.reportfile is not None:
                g.write_polygon_report(k, j.reportfile, os.path.basename(j.polyfile), os.path.basename(j.tpfile),
                                       int(cnt[k]), inside, query_ms * float(q_share[k]),
                                       build_ms * float(e_share[k]))
            if j.resultfile is not None:
                gahhg_amd.write_result_csv(j.resultfile, lab)
    finally:
        g.close()
    return out


def run(poly_paths, tpfile, reportfile=None, resultdir=None, device=0, **gpisp_kw):
    """Every polygon of poly_paths against every point of tpfile, as one
    batch.  Returns run_jobs' timings plus labels (P x Q), inside (P,),
    depth (P,) and n_edges (P,); raises if a file does not load."""
    if resultdir is not None:
        os.makedirs(resultdir, exist_ok=True)
    r = run_jobs(series_jobs(poly_paths, tpfile, reportfile, resultdir), device, **gpisp_kw)
    if r["skipped"]:
        i, why = r["skipped"][0]
        raise gahhg_amd.GahhgError(-2, f"{len(r['skipped'])} runs did not load; first: {why}")
    runs = r["runs"]
    r["labels"] = np.stack([x["labels"] for x in runs])
    for key in ("inside", "depth", "n_edges"):
        r[key] = np.array([x[key] for x in runs], dtype=np.int64)
    return r


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--explist", help="command list in the reference's syntax")
    ap.add_argument("--base", help="directory the list's paths are relative to (default: the list's)")
    ap.add_argument("--polydir")
    ap.add_argument("--tpfile")
    ap.add_argument("--reportfile")
    ap.add_argument("--resultdir")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    if a.explist:
        jobs, dirs = parse_command_list(a.explist, a.base)
        for d in dirs:
            os.makedirs(d, exist_ok=True)
    elif a.polydir and a.tpfile:
        if a.resultdir:
            os.makedirs(a.resultdir, exist_ok=True)
        jobs = series_jobs(series_files(a.polydir), a.tpfile, a.reportfile, a.resultdir)
    else:
        ap.error("give --explist, or --polydir and --tpfile")
    r = run_jobs(jobs, a.device)
    for i, why in r["skipped"]:
        print(f"skipped {jobs[i].polyfile}: {why}")
    done = [x for x in r["runs"] if x is not None]
    pairs = sum(len(x["labels"]) for x in done)
    print(f"{len(done)} runs, {pairs} (polygon, point) pairs: build {r['build_ms']:.2f} ms, "
          f"query {r['query_ms']:.3f} ms, total {r.get('wall_ms', 0.0):.1f} ms; "
          f"{sum(x['inside'] for x in done)} pairs inside")
    return 1 if r["skipped"] else 0


if __name__ == "__main__":
    raise SystemExit(main())
