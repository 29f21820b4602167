"""Benchmark: point-in-spherical-polygon queries on one MI355X.

Workload (default): fractal_1m -- one polygon with a 1M-vertex fractal outer
ring (radius 40 deg, centred at lat 60 deg, enclosing the North Pole) and 3
star holes of 50k vertices (N = 1,150,000 edges), queried with 100M FP64
points (70% uniform on the sphere, 30% within ~0.01 deg of an edge).

A step is one query pass over all points, inputs resident in HBM.  The
hierarchy build (create + classify, the reference's "pre-processing") is
timed separately (median of several builds) and reported as build_ms.
Labels of a sample are checked against the CPU oracle before timing.

Prints one JSON line.  --gpus N > 1 is not supported (one GPU per box).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "queries/s & build ms (1M-vtx polygon), batched+learned vs plain, bit-exact"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def query_kw(args):
    kw = {} if args.query_xcd is None else {"query_xcd": args.query_xcd}
    if args.table_budget is not None:
        kw["query_table_budget"] = args.table_budget
    if args.query_graph:
        kw["query_graph"] = 1
    return kw


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def per_query_flops(qc, n):
    """Algorithmic FP64 operations per query of the reference query kernel,
    from exact oracle counters: 5 per plane test, (3 bnum + 14) for the
    centre of a boundary leaf, 64 per edge-crossing test."""
    return (5 * qc["plane_tests"] + 3 * qc["bnum_sum"] + 14 * qc["boundary_hits"] + 64 * qc["ref_tests"]) / n


def measured_traffic(depth, Q, prm):
    """HBM bytes per k_query launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (calibrated against a copy of known
    size), when the hierarchy, the point count and the query path (prm: the
    context's parameters) match; else None."""
    import re
    path = os.path.join(ROOT, "profiles", "query_fractal_1m_traffic.md")
    if not os.path.exists(path) or Q != 100_000_000:
        return None
    m = re.search(r"```json\n(.*?)\n```", open(path).read(), re.S)
    if not m:
        return None
    d = json.loads(m.group(1))
    same = (d.get("depth") == depth and d.get("query_sort_bits") == prm.query_sort_bits and d.get("table_pass")
            and d.get("leaf_index") == bool(prm.query_index) and d.get("query_sub") == prm.query_sub
            and d.get("query_bins") == prm.query_bins and d.get("query_xcd", 0) == prm.query_xcd
            and d.get("descent") == "node records" and d.get("leaf_stage") == "per-lane prefetch")
    return round(float(d["k_query_traffic_bytes"]), 0) if same else None


def load_depth_model(args):
    """The trained refinement-depth model (None when --rule or absent)."""
    if args.rule:
        return None
    from gahhg_amd import learned
    if not os.path.exists(learned.MODEL_PATH):
        return None
    params, cost, meta = learned.load_model()
    return params, cost


def apply_model(g, model, xyz, poly, Q, sample=262_144):
    """Query hint (every k-th point standing for k points) and the model."""
    params, cost = model
    k = max(1, len(xyz) // sample)
    g.set_query_hint(xyz[::k], None if poly is None else poly[::k], float(k))
    g.set_depth_model(params, [float(Q), 1.0, cost["ns_edge"], cost["ns_cell"], cost["ns_level"]])


def copy_256mib(g):
    """One device-to-device copy of 256 MiB (the runtime's copy kernel reads
    and writes exactly 256 MiB): the counter reference for traffic."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = g.alloc(256 << 20), g.alloc(256 << 20)
    rc = hip.hipMemcpyDtoD(b.ptr, a.ptr, ctypes.c_size_t(256 << 20))
    hip.hipDeviceSynchronize()
    log(f"calibration copy: 256 MiB device-to-device, rc {rc}")
    a.free()
    b.free()


def median_build(g, n):
    b = []
    for i in range(n + 1):
        g.build()
        s = g.stats()
        if i:
            b.append((s["pre_ms"], s["create_ms"], s["classify_ms"]))
    b.sort()
    return b[len(b) // 2]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--queries", type=int, default=100_000_000)
    ap.add_argument("--builds", type=int, default=5)
    ap.add_argument("--cpu-sample", type=int, default=4_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--bins", type=int, default=256)
    ap.add_argument("--query-xcd", type=int, default=None,
                    help="blocks per XCD run in the query kernel (0 = plain order; default: the library's)")
    ap.add_argument("--table-budget", type=int, default=None,
                    help="batches: start-table entries per polygon edge (0 = full tables)")
    ap.add_argument("--query-graph", action="store_true",
                    help="replay each query pass as a recorded HIP graph")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-queries", type=int, default=64_000_000)
    ap.add_argument("--no-series", action="store_true")
    ap.add_argument("--no-f512", action="store_true", help="skip the 512-polygon learned-vs-rule A/B")
    ap.add_argument("--f512-queries", type=int, default=32_000_000)
    ap.add_argument("--no-load", action="store_true")
    ap.add_argument("--load-lines", type=int, default=20_000_000)
    ap.add_argument("--rule", action="store_true", help="density-rule depth only (no learned model)")
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="also run one 256 MiB device-to-device copy (a known-size reference for "
                         "FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import gahhg_amd
    from gahhg_amd import datasets

    t0 = time.time()
    w = datasets.fractal_1m(n_queries=args.queries)
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    del w["qlon"], w["qlat"]
    Q = xyz.shape[0]
    log(f"generated N={len(w['lon']) - (len(w['ring_off']) - 1)} edges, Q={Q} in {time.time() - t0:.1f}s")

    model = load_depth_model(args)
    # the density-rule build of the accelerated path (the learned build is
    # compared against it)
    g = gahhg_amd.GPISP(device=0, query_bins=args.bins, **query_kw(args))
    g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"])
    rule_pre, rule_create, rule_classify = median_build(g, args.builds)
    rule_depth = g.stats()["depth"]
    log(f"rule build: depth {rule_depth} pre {rule_pre:.2f} ms (create {rule_create:.2f}, classify {rule_classify:.2f})")
    if model is not None:
        apply_model(g, model, xyz, None, Q)
    pre_ms, create_ms, classify_ms = median_build(g, args.builds)
    st = g.stats()
    log(f"build ({'learned' if model else 'rule'} depth): depth {st['depth']} pre {pre_ms:.2f} ms "
        f"(create {create_ms:.2f}, classify {classify_ms:.2f})")

    d_xyz = g.alloc(xyz.nbytes)
    d_lab = g.alloc(4 * Q)
    d_xyz.upload(xyz)
    if args.pmc_calibrate:
        copy_256mib(g)

    # parity spot check against the oracle on a sample (checker only)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    o = oracle_lib.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"])
    tb = time.time()
    if model is not None and st["depth"] != rule_depth:
        o.build(max_depth=int(st["depth"]), auto_depth=False)  # the same hierarchy
    else:
        o.build()
    cpu_build_s = time.time() - tb
    g.query_device(d_xyz, Q, d_lab)
    labels = d_lab.download(np.empty(Q, dtype=np.int32))
    ns = min(Q, args.cpu_sample)
    idx = np.random.default_rng(0).choice(Q, ns, replace=False) if ns < Q else np.arange(Q)
    samp = np.ascontiguousarray(xyz[idx])
    nthreads = os.cpu_count() or 1
    tq = time.time()
    ref, qc = o.query_xyz(samp, nthreads=0, with_counts=True)
    cpu_q_s = time.time() - tq
    mism = int(np.count_nonzero(labels[idx] != ref))
    log(f"parity sample {ns}: mismatches {mism}; cpu {ns / cpu_q_s / 1e6:.2f} Mq/s on {nthreads} threads")
    if mism:
        log("LABEL MISMATCH against the oracle")

    for _ in range(args.warmup):
        g.query_device(d_xyz, Q, d_lab)
    kernel_ms, sort_ms, gpu_ms = [], [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        g.query_device(d_xyz, Q, d_lab)  # synchronises on return
        s = g.stats()
        kernel_ms.append(s["query_kernel_ms"])
        sort_ms.append(s["query_sort_ms"])
        gpu_ms.append(s["query_ms"])
    t_end = time.perf_counter()
    ms_per_step = (t_end - t_start) * 1e3 / args.steps
    value = Q / (ms_per_step * 1e-3)
    kms = float(np.mean(kernel_ms))
    log(f"query: {ms_per_step:.3f} ms/step wall, gpu {np.mean(gpu_ms):.3f} ms = grouping {np.mean(sort_ms):.3f} "
        f"+ k_query {kms:.3f}")

    # the plain path: the reference's kernel structure (linear level-0 scan,
    # descent, per-query leaf centre and per-ref edge planes) on the same grid
    plain = None
    if not args.no_plain:
        gp = gahhg_amd.GPISP(device=0, query_layout=1, query_bins=0)
        gp.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"])
        pb = []
        for i in range(args.builds + 1):
            gp.build()
            if i:
                pb.append(gp.stats()["pre_ms"])
        dl2 = gp.alloc(4 * Q)
        dx2 = gp.alloc(xyz.nbytes)
        dx2.upload(xyz)
        gp.query_device(dx2, Q, dl2)
        pl = []
        for _ in range(3):
            gp.query_device(dx2, Q, dl2)
            pl.append(gp.stats()["query_ms"])
        lab2 = dl2.download(np.empty(Q, dtype=np.int32))
        plain = {"queries_per_s": round(Q / (float(np.median(pl)) * 1e-3), 1), "kernel_ms": round(float(np.median(pl)), 3),
                 "build_ms": round(float(np.median(pb)), 3),
                 # differences can only come from a different depth (points
                 # within 1e-8 of an edge, whose label depends on the hierarchy)
                 "label_differences": int(np.count_nonzero(lab2 != labels)),
                 "same_depth": int(gp.stats()["depth"]) == int(st["depth"])}
        log(f"plain path: {plain}")
        dx2.free()
        dl2.free()
        gp.close()

    fq = per_query_flops(qc, ns)
    achieved_tf = fq * Q / (kms * 1e-3) / 1e12
    achieved_gbs = 28.0 * Q / (kms * 1e-3) / 1e9
    frac_fp = achieved_tf / FP64_PEAK_TFLOPS
    frac_hbm = achieved_gbs / HBM_PEAK_GBS
    if frac_fp >= frac_hbm:
        roof = {"bound": "mfma", "achieved": round(achieved_tf, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(frac_fp, 4), "traffic": None,
                "note": "FP64 vector ALU (no MFMA is used); algorithmic FP64 ops/query "
                        f"{fq:.1f} from exact oracle counters; compulsory HBM {achieved_gbs:.1f} GB/s at 28 B/query"}
    else:
        roof = {"bound": "hbm", "achieved": round(achieved_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(frac_hbm, 4), "traffic": None}
    roof["kernel"] = "k_query"
    roof["kernel_ms"] = round(kms, 4)
    roof["traffic"] = measured_traffic(int(st["depth"]), Q,
                                       gahhg_amd.default_params(query_bins=args.bins, **query_kw(args)))
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "fractal_1m: one polygon, 1M-vertex fractal ring (r=40 deg, centre lat 60 deg) "
                               "+ 3 holes x 50k vertices; 100M FP64 queries (70% uniform, 30% near-edge)",
                   "n_edges": int(st["n_edges"]), "n_queries": int(Q),
                   "path": "single-polygon GPU: learned depth; start tables answer what they resolve, the rest grouped in Z-order (device radix sort) and walked on node records with angular-indexed leaf tests",
                   "parallelism": "1 GPU"},
        "build_ms": round(pre_ms, 3), "create_ms": round(create_ms, 3), "classify_ms": round(classify_ms, 3),
        "depth": int(st["depth"]), "depth_rule": "learned" if model is not None else "density",
        "density_rule_build": {"depth": int(rule_depth), "build_ms": round(rule_pre, 3)},
        "grid_bytes": int(st["gpu_space_bytes"]),
        "parity": {"sample": int(ns), "mismatches": mism},
        "query_bins": args.bins,
        "query_gpu_ms": round(float(np.mean(gpu_ms)), 4), "query_grouping_ms": round(float(np.mean(sort_ms)), 4),
        "plain_path": plain,
        "speedup_vs_plain": round(value / plain["queries_per_s"], 3) if plain else None,
        "build_ratio_vs_plain": round(pre_ms / plain["build_ms"], 3) if plain else None,
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = {"value": round(ns / cpu_q_s, 1), "unit": "queries/s", "cores": nthreads,
                               "kind": "port",
                               "sample": f"CPU oracle, same polygon, {ns} of the {Q} queries on {nthreads} threads; "
                                         f"oracle build {cpu_build_s * 1e3:.0f} ms"}
    d_xyz.free()
    d_lab.free()
    g.close()
    del xyz, labels
    if not args.no_batch:
        out["batch_256"] = bench_batch(args, oracle_lib)
    if not args.no_f512:
        out["fractal_512"] = bench_fractal512(args, oracle_lib)
    if not args.no_series:
        out["series_mars"] = bench_series(args, oracle_lib)
    if not args.no_load:
        out["load_path"] = bench_load(args)
    print(json.dumps(out), flush=True)


def bench_load(args):
    """A query CSV of --load-lines points ("lon,lat" with 6 decimals, as the
    reference's query files) from the file to device-resident xyz: the host
    loader (sscanf per line, conversion on the host threads, copy to the
    device) against the device loader (file bytes copied once, split, parse
    and conversion on the GPU).  Both give the same bits."""
    import shutil
    import tempfile

    import gahhg_amd
    from gahhg_amd import datasets

    rng = np.random.default_rng(17)
    n = args.load_lines
    t0 = time.time()
    lon = rng.uniform(-180.0, 180.0, n)
    lat = np.degrees(np.arcsin(rng.uniform(-1.0, 1.0, n)))
    text = datasets.query_csv_bytes(lon, lat, decimals=6)
    del lon, lat
    tmp = tempfile.mkdtemp(prefix="gahhg_load_")
    path = os.path.join(tmp, "queries.csv")
    with open(path, "wb") as f:
        f.write(text)
    nbytes = len(text)
    del text
    log(f"load_path: {n} lines, {nbytes / 1e6:.1f} MB of text written in {time.time() - t0:.1f}s")
    g = gahhg_amd.GPISP(device=0)
    try:
        # host loader
        t = time.perf_counter()
        hlon, hlat = gahhg_amd.read_query_csv(path)
        t_read = time.perf_counter()
        hxyz = gahhg_amd.query_lonlat_to_xyz(hlon, hlat)
        t_conv = time.perf_counter()
        d = g.alloc(hxyz.nbytes)
        d.upload(hxyz)
        t_up = time.perf_counter()
        host = {"read_parse_ms": round((t_read - t) * 1e3, 1), "convert_ms": round((t_conv - t_read) * 1e3, 1),
                "upload_ms": round((t_up - t_conv) * 1e3, 1), "total_ms": round((t_up - t) * 1e3, 1),
                "convert_threads": os.cpu_count()}
        d.free()
        del hlon, hlat
        # device loader
        infos = []
        for _ in range(3):
            v, m = g.load_query_csv_device(path)
            infos.append(g.load_info)
        dxyz = v.download(np.empty((m, 3)))
        same = bool(m == len(hxyz) and np.array_equal(dxyz.view(np.int64), hxyz.view(np.int64)))
        del dxyz, hxyz
        tot = float(np.median([i["total_ms"] for i in infos]))
        par = float(np.median([i["parse_ms"] for i in infos]))
        rd = float(np.median([i["read_ms"] for i in infos]))
        res = {"lines": n, "text_bytes": nbytes, "host_loader": host,
               "device_loader": {"total_ms": round(tot, 2), "read_copy_ms": round(rd, 2), "parse_ms": round(par, 3),
                                 "host_lines": int(infos[-1]["host_lines"]),
                                 "parse_lines_per_s": round(n / (par * 1e-3), 1),
                                 "parse_text_GB_per_s": round(nbytes / (par * 1e-3) / 1e9, 2)},
               "speedup_total": round(host["total_ms"] / tot, 2), "bit_identical": same}
    finally:
        g.close()
        shutil.rmtree(tmp, ignore_errors=True)
    log(f"load_path: {res}")
    return res


def bench_series(args, oracle_lib):
    """The reference's Mars series (624 polygon files x the 10000 points of
    Mars_query.csv): one batch (file reads, build, query, label download) against
    one context per polygon in turn (the reference starts one process per file;
    process and device start-up are not counted here)."""
    import gahhg_amd
    from gahhg_amd import sweep

    gold = os.path.join(ROOT, "tests", "golden")
    files = sweep.series_files(os.path.join(gold, "mars"))
    qcsv = os.path.join(gold, "reference", "Mars_query.csv")
    sweep.run(files[:8], qcsv, **query_kw(args))  # warm the library and device
    walls = []
    for _ in range(3):
        r = sweep.run(files, qcsv, **query_kw(args))
        walls.append(r["wall_ms"])
    lab = r["labels"]
    lon, lat = gahhg_amd.read_query_csv(qcsv)
    xyz = gahhg_amd.query_lonlat_to_xyz(lon, lat)
    t = time.perf_counter()
    g = gahhg_amd.GPISP(device=0)
    pb = pq = 0.0
    mism = 0
    for k, f in enumerate(files):
        g.load_polygon_txt(f).build()
        pb += g.stats()["pre_ms"]
        mism += int(np.count_nonzero(g.query_xyz(xyz) != lab[k]))
        pq += g.stats()["query_ms"]
    g.close()
    per_wall = (time.perf_counter() - t) * 1e3
    omism = 0
    for k, f in enumerate(files):
        omism += int(np.count_nonzero(oracle_lib.Oracle().load_txt(f).build().query_xyz(xyz) != lab[k]))
    P, Q = lab.shape
    res = {"n_polygons": P, "n_points": Q, "pairs": P * Q,
           "batched": {"build_ms": round(r["build_ms"], 3), "query_ms": round(r["query_ms"], 3),
                       "wall_ms": round(float(np.median(walls)), 1),
                       "pairs_per_s": round(P * Q / (float(np.median(walls)) * 1e-3), 1)},
           "per_polygon": {"build_ms_sum": round(pb, 3), "query_ms_sum": round(pq, 3), "wall_ms": round(per_wall, 1)},
           "wall_speedup": round(per_wall / float(np.median(walls)), 2),
           "label_differences_vs_per_polygon": mism, "oracle_mismatches": omism}
    log(f"series_mars: {res}")
    return res


def bench_fractal512(args, oracle_lib):
    """fractal_512 (512 polygons of the fractal_1m generator scaled to 4k-64k
    outer vertices, 32M points): the learned depth against the density rule
    as one batch each, on build ms, queries/s and grid bytes; labels of the
    two must be equal and match the oracle on a sample of polygons."""
    import gahhg_amd
    from gahhg_amd import datasets

    t0 = time.time()
    w = datasets.fractal_512(n_queries=args.f512_queries)
    items = [(p["lon"], p["lat"], p["ring_off"]) for p in w["polys"]]
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    qp = w["qpoly"]
    Q = len(qp)
    P = len(items)
    n_edges = int(sum(len(p[0]) - (len(p[2]) - 1) for p in items))
    log(f"fractal_512: {P} polygons, {n_edges} edges, {Q} queries generated in {time.time() - t0:.1f}s")
    model = load_depth_model(args)
    res = {"n_polygons": P, "n_edges": n_edges, "n_queries": Q}
    labs, depths = {}, {}
    for mode in ("rule", "learned"):
        if mode == "learned" and model is None:
            continue
        g = gahhg_amd.GPISP(device=0, query_bins=args.bins, **query_kw(args))
        g.set_polygons_lonlat(items)
        if mode == "learned":
            apply_model(g, model, xyz, qp, Q)
        build_ms = median_build(g, 3)[0]
        st = [g.polygon_stats(k) for k in range(P)]
        depths[mode] = [s["depth"] for s in st]
        grid = int(sum(s["gpu_space_bytes"] for s in st))
        dx, dp, dl = g.alloc(xyz.nbytes), g.alloc(4 * Q), g.alloc(4 * Q)
        dx.upload(xyz)
        dp.upload(qp)
        g.query_batch_device(dx, dp, Q, dl)
        qms = []
        for _ in range(5):
            g.query_batch_device(dx, dp, Q, dl)
            qms.append(g.stats()["query_ms"])
        labs[mode] = dl.download(np.empty(Q, dtype=np.int32))
        for b in (dx, dp, dl):
            b.free()
        g.close()
        q = float(np.median(qms))
        res[mode] = {"build_ms": round(build_ms, 3), "query_gpu_ms": round(q, 3),
                     "queries_per_s": round(Q / (q * 1e-3), 1), "grid_bytes": grid,
                     "mean_depth": round(float(np.mean(depths[mode])), 3)}
        log(f"fractal_512 {mode}: {res[mode]}")
    # oracle on the points of 24 polygons (checker only)
    rng = np.random.default_rng(7)
    check = rng.choice(P, min(P, 24), replace=False)
    final = labs.get("learned", labs["rule"])
    fd = depths.get("learned", depths["rule"])
    om = 0
    on = 0
    for k in check:
        sel = np.nonzero(qp == k)[0][:20000]
        o = oracle_lib.Oracle().set_lonlat(*items[k]).build(max_depth=int(fd[k]), auto_depth=False)
        om += int(np.count_nonzero(final[sel] != o.query_xyz(np.ascontiguousarray(xyz[sel]))))
        on += len(sel)
    res["oracle_sample"] = on
    res["oracle_mismatches"] = om
    if "learned" in res:
        r, l = res["rule"], res["learned"]
        res["label_differences_learned_vs_rule"] = int(np.count_nonzero(labs["rule"] != labs["learned"]))
        res["depth_changes"] = int(sum(a != b for a, b in zip(depths["rule"], depths["learned"])))
        res["learned_vs_rule"] = {"build": round(l["build_ms"] / r["build_ms"], 4),
                                  "queries_per_s": round(l["queries_per_s"] / r["queries_per_s"], 4),
                                  "grid_bytes": round(l["grid_bytes"] / r["grid_bytes"], 4)}
    log(f"fractal_512: {res}")
    return res


def bench_batch(args, oracle_lib):
    """256 polygons (log-uniform 16..200k vertices), 64M Zipf-skewed queries:
    one batched build + one query launch, against building and querying each
    polygon on its own (the reference runs one polygon per process)."""
    import gahhg_amd
    from gahhg_amd import datasets

    t0 = time.time()
    w = datasets.batch_256(n_queries=args.batch_queries)
    items = [(p["lon"], p["lat"], p["ring_off"]) for p in w["polys"]]
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    qp = w["qpoly"]
    Q = len(qp)
    n_edges = int(sum(len(p[0]) - (len(p[2]) - 1) for p in items))
    log(f"batch_256: {len(items)} polygons, {n_edges} edges, {Q} queries generated in {time.time() - t0:.1f}s")
    model = load_depth_model(args)
    g = gahhg_amd.GPISP(device=0, query_bins=args.bins, **query_kw(args))
    g.set_polygons_lonlat(items)
    rule_build = median_build(g, 3)[0]
    rule_depths = [g.polygon_stats(k)["depth"] for k in range(len(items))]
    if model is not None:
        apply_model(g, model, xyz, qp, Q)
    build_ms = median_build(g, 3)[0]
    depths = [g.polygon_stats(k)["depth"] for k in range(len(items))]
    log(f"batch_256: build {build_ms:.2f} ms ({'learned' if model else 'rule'} depths), rule build {rule_build:.2f} ms; "
        f"depth changes vs rule: {sum(a != b for a, b in zip(depths, rule_depths))}")
    dx = g.alloc(xyz.nbytes)
    dp = g.alloc(4 * Q)
    dl = g.alloc(4 * Q)
    dx.upload(xyz)
    dp.upload(qp)
    g.query_batch_device(dx, dp, Q, dl)
    ks = []
    tq = time.perf_counter()
    kk = []
    for _ in range(max(3, args.steps // 2)):
        g.query_batch_device(dx, dp, Q, dl)
        ks.append(g.stats()["query_ms"])
        kk.append(g.stats()["query_kernel_ms"])
    wall_ms = (time.perf_counter() - tq) * 1e3 / len(ks)
    lab = dl.download(np.empty(Q, dtype=np.int32))
    dx.free()
    dp.free()
    dl.free()
    g.close()
    # per-polygon runs: build each polygon alone, query its points (grouped)
    order = np.argsort(qp, kind="stable")
    cnt = np.bincount(qp, minlength=len(items))
    gp = gahhg_amd.GPISP(device=0, query_layout=1, query_bins=0)
    pb = pq = 0.0
    start = 0
    mism = 0
    for k, it in enumerate(items):
        gp.set_polygon_lonlat(*it).build()
        pb += gp.stats()["pre_ms"]
        idx = order[start:start + cnt[k]]
        start += cnt[k]
        if len(idx):
            l2 = gp.query_xyz(xyz[idx])
            pq += gp.stats()["query_ms"]
            mism += int(np.count_nonzero(l2 != lab[idx]))
    gp.close()
    # oracle check on a sample of polygons' points
    rng = np.random.default_rng(1)
    omism = 0
    checked = 0
    for k in rng.choice(len(items), 24, replace=False):
        idx = order[int(cnt[:k].sum()):int(cnt[:k + 1].sum())][:20000]
        if not len(idx):
            continue
        o = oracle_lib.Oracle().set_lonlat(*items[k]).build(max_depth=depths[k], auto_depth=False)
        omism += int(np.count_nonzero(o.query_xyz(np.ascontiguousarray(xyz[idx])) != lab[idx]))
        checked += len(idx)
    res = {"n_polygons": len(items), "n_edges": n_edges, "n_queries": Q,
           "depth_rule": "learned" if model is not None else "density",
           "density_rule_build_ms": round(rule_build, 3),
           "depths_changed_vs_density_rule": int(sum(a != b for a, b in zip(depths, rule_depths))),
           "batched": {"build_ms": round(build_ms, 3), "query_gpu_ms": round(float(np.median(ks)), 3),
                       "query_kernel_ms": round(float(np.median(kk)), 3),
                       "query_wall_ms": round(wall_ms, 3), "queries_per_s": round(Q / (float(np.median(ks)) * 1e-3), 1)},
           "per_polygon_plain": {"build_ms_sum": round(pb, 3), "query_kernel_ms_sum": round(pq, 3),
                                 "queries_per_s": round(Q / (pq * 1e-3), 1) if pq else None},
           "build_ratio": round(build_ms / pb, 4) if pb else None,
           "query_speedup": round(pq / float(np.median(ks)), 3) if pq else None,
           "label_differences_vs_per_polygon": mism,
           "per_polygon_same_depths": depths == rule_depths,
           "oracle_sample": checked, "oracle_mismatches": omism}
    log(f"batch_256: {res}")
    return res


if __name__ == "__main__":
    main()
