"""Benchmark: point-in-spherical-polygon queries on one MI355X.

Workload (default): fractal_1m -- one polygon with a 1M-vertex fractal outer
ring (radius 40 deg, centred at lat 60 deg, enclosing the North Pole) and 3
star holes of 50k vertices (N = 1,150,000 edges), queried with 100M FP64
points (70% uniform on the sphere, 30% within ~0.01 deg of an edge).

The hierarchy is the density rule's (the reference's), so every label is the
reference algorithm's; all 100M labels are checked against the CPU oracle.
The learned start-table model picks the depth of the query's start table
(label-neutral).  A step is one query pass over all points (recorded as a HIP
graph and replayed), inputs resident in HBM.  The build (create + classify,
the reference's "pre-processing") is reported cold (first build of a
context) and warm (median of several).

Also reported: the small and medium single polygons (star_1k, 1M points;
fractal_100k, 10M points) with the CPU oracle's build and query beside the
GPU's; the 256-polygon batch and the 512-polygon fractal batch (learned
tables vs the size rule, batched vs one polygon at a time); the Mars series;
the device text loader.

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


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def threads():
    """Host threads the CPU oracle uses (OpenMP's default on this host)."""
    return int(os.environ.get("OMP_NUM_THREADS", 0)) or (os.cpu_count() or 1)


def omp_set_threads(n):
    """Sets the OpenMP thread count of the calling thread (the oracle's
    parallel build and query regions)."""
    import ctypes
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
        return True
    except OSError:
        return False


def ref_flops(qc):
    """Reference-equivalent FP64 operations of the reference query kernel
    from exact oracle counters: 5 per plane test, (3 bnum + 14) for the
    centre of a boundary leaf, 64 per edge-crossing test."""
    return 5 * qc["plane_tests"] + 3 * qc["bnum_sum"] + 14 * qc["boundary_hits"] + 64 * qc["ref_tests"]


def oracle_full(o, xyz, chunk=10_000_000):
    """Oracle labels of every point and the summed FP64 counters."""
    labels = np.empty(len(xyz), np.int32)
    tot = None
    for s in range(0, len(xyz), chunk):
        lab, qc = o.query_xyz(np.ascontiguousarray(xyz[s:s + chunk]), nthreads=0, with_counts=True)
        labels[s:s + len(lab)] = lab
        tot = qc if tot is None else {k: tot[k] + qc[k] for k in tot}
    return labels, tot


def profile_json(name):
    """The JSON block of a committed rocprofv3 summary under profiles/."""
    import re
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"```json\n(.*?)\n```", open(path).read(), re.S)
    return json.loads(m.group(1)) if m else None


def measured_counters(cfg):
    """HBM bytes and executed FP64 instructions per k_query launch from the
    committed rocprofv3 --pmc summary of this workload, when its hierarchy,
    point count and query configuration (cfg) match the run; else None."""
    d = profile_json("query_fractal_1m_counters.md")
    if not d or any(d.get(k) != v for k, v in cfg.items()):
        return None
    return d


def builds(g, n):
    """Cold build (the context's first) and the median of n warm builds:
    (pre, create, classify, layout) ms each."""
    g.build()
    s = g.stats()
    cold = (s["pre_ms"], s["create_ms"], s["classify_ms"], s["layout_ms"])
    b = []
    for _ in range(n):
        g.build()
        s = g.stats()
        b.append((s["pre_ms"], s["create_ms"], s["classify_ms"], s["layout_ms"]))
    b.sort()
    return cold, b[len(b) // 2]


def table_model():
    from gahhg_amd import learned_tables
    if not os.path.exists(learned_tables.MODEL_PATH):
        return None
    params, cost, meta = learned_tables.load_model()
    return params, cost


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--queries", type=int, default=100_000_000)
    ap.add_argument("--builds", type=int, default=5)
    ap.add_argument("--cpu-sample", type=int, default=4_000_000)
    ap.add_argument("--cpu1-sample", type=int, default=200_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-small", action="store_true", help="skip the star_1k and fractal_100k lines")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-queries", type=int, default=64_000_000)
    ap.add_argument("--no-series", action="store_true")
    ap.add_argument("--no-f512", action="store_true")
    ap.add_argument("--f512-queries", type=int, default=32_000_000)
    ap.add_argument("--no-load", action="store_true")
    ap.add_argument("--load-lines", type=int, default=20_000_000)
    ap.add_argument("--no-model", action="store_true", help="library default start tables (no learned model)")
    ap.add_argument("--no-graph", action="store_true", help="launch each query pass kernel by kernel")
    ap.add_argument("--no-parity", action="store_true", help="skip the full oracle label check")
    ap.add_argument("--no-ab", action="store_true", help="skip the learned-vs-default table A/B")
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="also run one 256 MiB device-to-device copy (a known-size reference for "
                         "FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import gahhg_amd
    from gahhg_amd import datasets, learned_tables

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    t0 = time.time()
    w = datasets.fractal_1m(n_queries=args.queries)
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    del w["qlon"], w["qlat"]
    Q = xyz.shape[0]
    N = len(w["lon"]) - (len(w["ring_off"]) - 1)
    log(f"generated N={N} edges, Q={Q} in {time.time() - t0:.1f}s")
    model = None if args.no_model else table_model()
    graph = 0 if args.no_graph else 1

    # the headline context: density-rule hierarchy, learned start table
    g = gahhg_amd.GPISP(device=0, query_sub=16, query_graph=graph)
    g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"])
    if model is not None:
        # the hierarchy answers every pass of the run: the model weighs the
        # table's build against that query volume
        volume = Q * (args.steps + args.warmup + 1)
        learned_tables.apply(g, model, xyz, None, volume)
    cold, warm = builds(g, args.builds)
    st = g.stats()
    tab = g.table_layout()[0].tolist()
    log(f"build: depth {st['depth']} cold {cold[0]:.2f} ms, warm {warm[0]:.2f} ms (create {warm[1]:.2f}, "
        f"classify {warm[2]:.2f} of which layout {warm[3]:.2f}); start table {tab}")

    d_xyz = g.alloc(xyz.nbytes)
    d_lab = g.alloc(4 * Q)
    d_xyz.upload(xyz)
    if args.pmc_calibrate:
        copy_256mib(g)
    g.query_device(d_xyz, Q, d_lab)
    labels = d_lab.download(np.empty(Q, dtype=np.int32))

    for _ in range(args.warmup):
        g.query_device(d_xyz, Q, d_lab)
    gpu_ms, kernel_ms, sort_ms = [], [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        g.query_device(d_xyz, Q, d_lab)  # synchronises on return
        s = g.stats()
        gpu_ms.append(s["query_ms"])
        kernel_ms.append(s["query_kernel_ms"])
        sort_ms.append(s["query_sort_ms"])
    t_end = time.perf_counter()
    ms_per_step = (t_end - t_start) * 1e3 / args.steps
    value = Q / (ms_per_step * 1e-3)
    pass_ms = float(np.median(gpu_ms))
    log(f"query: {ms_per_step:.3f} ms/step wall, GPU pass {pass_ms:.3f} ms" +
        ("" if graph else f" = grouping {np.median(sort_ms):.3f} + k_query {np.median(kernel_ms):.3f}"))

    # the same hierarchy with the library's default start table against the
    # learned one, for 1M, 10M and all points (build + one pass)
    ab = None
    if model is not None and not args.no_ab:
        ab = {}
        for n in sorted({min(Q, 1_000_000), min(Q, 10_000_000), Q}):
            row = {}
            for mode in ("default", "learned"):
                gx = gahhg_amd.GPISP(device=0, query_sub=16, query_graph=graph)
                gx.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"])
                if mode == "learned":
                    learned_tables.apply(gx, model, xyz[:n], None, n)
                _, bw = builds(gx, 3)
                dl = gx.alloc(4 * n)
                gx.query_device(d_xyz, n, dl)
                qq = []
                for _ in range(3):
                    gx.query_device(d_xyz, n, dl)
                    qq.append(gx.stats()["query_ms"])
                lab_n = dl.download(np.empty(n, dtype=np.int32))
                row[mode] = {"table": gx.table_layout()[0].tolist(), "build_ms": round(bw[0], 3),
                             "query_gpu_ms": round(float(np.median(qq)), 3),
                             "label_differences": int(np.count_nonzero(lab_n != labels[:n]))}
                dl.free()
                gx.close()
            row["build_plus_pass_ratio"] = round(
                (row["learned"]["build_ms"] + row["learned"]["query_gpu_ms"]) /
                (row["default"]["build_ms"] + row["default"]["query_gpu_ms"]), 4)
            ab[str(n)] = row
            log(f"learned table vs library default at {n} points: {row}")

    # the plain path: the reference's kernel structure (linear level-0 scan,
    # descent, per-query leaf centre and per-ref edge planes), no query
    # layout, same density-rule hierarchy
    plain = None
    if not args.no_plain:
        gp = gahhg_amd.GPISP(device=0, query_layout=1, query_bins=0)
        gp.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"])
        pcold, pwarm = builds(gp, args.builds)
        dl2 = gp.alloc(4 * Q)
        gp.query_device(d_xyz, Q, dl2)
        pl = []
        for _ in range(3):
            gp.query_device(d_xyz, Q, dl2)
            pl.append(gp.stats()["query_ms"])
        lab2 = dl2.download(np.empty(Q, dtype=np.int32))
        plain = {"queries_per_s": round(Q / (float(np.median(pl)) * 1e-3), 1),
                 "query_gpu_ms": round(float(np.median(pl)), 3),
                 "build_ms": round(pwarm[0], 3), "build_ms_cold": round(pcold[0], 3),
                 "depth": int(gp.stats()["depth"]),
                 # same hierarchy, other query kernel: every label must agree
                 "label_differences": int(np.count_nonzero(lab2 != labels))}
        log(f"plain path: {plain}")
        dl2.free()
        gp.close()
    d_xyz.free()
    d_lab.free()

    # every label against the oracle (density-rule hierarchy, checker only),
    # with the exact FP64 operation counts of the reference query kernel
    o = oracle_lib.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"])
    tb = time.time()
    o.build()
    cpu_build_s = time.time() - tb
    parity = {"checked": 0, "mismatches": None}
    ops = None
    if not args.no_parity:
        tq = time.time()
        ref, qc = oracle_full(o, xyz)
        parity = {"checked": int(Q), "mismatches": int(np.count_nonzero(ref != labels)),
                  "against": "CPU oracle, density-rule hierarchy (the reference's)",
                  "oracle_s": round(time.time() - tq, 1)}
        ops = float(ref_flops(qc))
        del ref
        log(f"parity: {parity}")
        if parity["mismatches"]:
            log("LABEL MISMATCH against the oracle")

    # roofline of the query pass: reference-equivalent FP64 operations of all
    # points over the pass's GPU time (table pass, grouping and walk)
    cfg = {"depth": int(st["depth"]), "n_queries": int(Q), "table": tab, "graph": bool(graph)}
    cnt = measured_counters(cfg)
    roof = {"bound": "fp64", "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "traffic": None,
            "kernel": "query pass (k_query_table, grouping, k_query)", "pass_ms": round(pass_ms, 4)}
    if ops is not None:
        ach = ops / (pass_ms * 1e-3) / 1e12
        roof.update(achieved=round(ach, 3), frac=round(ach / FP64_PEAK_TFLOPS, 4), ops_per_pass=ops,
                    ops_per_query=round(ops / Q, 2),
                    ops_source="reference-equivalent FP64 ops (5 per plane test, 3 bnum + 14 per boundary-leaf "
                               "centre, 64 per edge test) counted exactly by the oracle over every point")
    roof["hbm_compulsory"] = {"bytes_per_pass": 28 * Q, "GBps": round(28.0 * Q / (pass_ms * 1e-3) / 1e9, 1),
                              "frac": round(28.0 * Q / (pass_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if cnt:
        roof["traffic"] = cnt.get("k_query_hbm_bytes")
        roof["measured"] = {k: cnt[k] for k in ("k_query_ms", "k_query_hbm_bytes", "k_query_fp64_ops_executed",
                                                "k_query_fp64_TFLOPs_executed", "pass_kernels_ms") if k in cnt}
        roof["measured"]["source"] = "profiles/query_fractal_1m_counters.md (rocprofv3 --pmc)"

    # build roofline: B_b = (M + 2) N 56 B + 2 S_grid over the warm build
    M = int(st["mark_passes"])
    Sg = int(st["gpu_space_bytes"])
    Bb = (M + 2) * N * 56 + 2 * Sg
    build_roof = {"bound": "hbm", "bytes": Bb, "M": M, "S_grid": Sg, "build_ms": round(warm[0], 3),
                  "achieved": round(Bb / (warm[0] * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(Bb / (warm[0] * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "fractal_1m: one polygon, 1M-vertex fractal ring (r=40 deg, centre lat 60 deg) "
                               "+ 3 holes x 50k vertices; 100M FP64 queries (70% uniform, 30% near-edge)",
                   "n_edges": int(N), "n_queries": int(Q),
                   "path": "single-polygon GPU, density-rule hierarchy (the reference's labels); learned start "
                           "table; query pass replayed as a HIP graph: table pass, Z-order grouping (device "
                           "radix sort), walk on node records with angular-indexed leaf tests",
                   "parallelism": "1 GPU"},
        "build_ms": round(warm[0], 3), "build_ms_cold": round(cold[0], 3), "create_ms": round(warm[1], 3),
        "classify_ms": round(warm[2], 3), "layout_ms": round(warm[3], 3), "depth": int(st["depth"]),
        "start_table": {"bins": tab[0], "sub": tab[1], "chosen_by": "learned model" if model else "library default",
                        "query_volume": int(Q * (args.steps + args.warmup + 1)) if model else None},
        "grid_bytes": Sg, "parity": parity,
        "query_gpu_ms": round(pass_ms, 4),
        "plain_path": plain,
        "speedup_vs_plain": round(value / plain["queries_per_s"], 3) if plain else None,
        "build_ratio_vs_plain": round(warm[0] / plain["build_ms"], 3) if plain else None,
        "learned_table_ab": ab,
        "roofline": roof, "build_roofline": build_roof,
    }
    if not args.no_cpu:
        nt = threads()
        ns = min(Q, args.cpu_sample)
        idx = np.random.default_rng(0).choice(Q, ns, replace=False)
        samp = np.ascontiguousarray(xyz[idx])
        tq = time.time()
        o.query_xyz(samp, nthreads=0)
        cpu_q_s = time.time() - tq
        n1 = min(Q, args.cpu1_sample)
        t1 = time.time()
        o.query_xyz(np.ascontiguousarray(samp[:n1]), nthreads=1)
        cpu1_s = time.time() - t1
        out["cpu_baseline"] = {"value": round(ns / cpu_q_s, 1), "unit": "queries/s", "cores": nt, "kind": "port",
                               "sample": f"CPU oracle (the reference's algorithm restated in C, -O2, OpenMP), same "
                                         f"polygon and hierarchy, {ns} of the {Q} queries on {nt} threads; "
                                         f"build {cpu_build_s * 1e3:.0f} ms on {nt} threads",
                               "one_thread": {"value": round(n1 / cpu1_s, 1), "sample": n1},
                               "build_ms": round(cpu_build_s * 1e3, 1)}
        log(f"cpu baseline: {out['cpu_baseline']}")
    g.close()
    del xyz, labels
    if not args.no_small:
        out["star_1k"] = bench_small("star_1k", datasets.star_1k(), oracle_lib, args)
        out["fractal_100k"] = bench_small("fractal_100k", datasets.fractal_100k(), oracle_lib, args)
    if not args.no_batch:
        out["batch_256"] = bench_batch(args, oracle_lib, model)
    if not args.no_f512:
        out["fractal_512"] = bench_fractal512(args, oracle_lib, model)
    if not args.no_series:
        out["series_mars"] = bench_series(args, oracle_lib)
    if not args.no_load:
        out["load_path"] = bench_load(args)
    print(json.dumps(out), flush=True)


def query_kw(args):
    """Context parameters of the Mars series runs (library defaults)."""
    return {}


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


def bench_small(name, w, oracle_lib, args):
    """One polygon (star_1k: 1,200 edges, 1M points; fractal_100k: 110,000
    edges, 10M points): GPU build (cold and warm) and query pass against the
    CPU oracle's build and query on all threads and on one, labels checked
    on every point."""
    import gahhg_amd
    xyz = gahhg_amd.query_lonlat_to_xyz(w["qlon"], w["qlat"])
    Q = len(xyz)
    N = len(w["lon"]) - (len(w["ring_off"]) - 1)
    g = gahhg_amd.GPISP(device=0, query_graph=0 if args.no_graph else 1)
    g.set_polygon_lonlat(w["lon"], w["lat"], w["ring_off"])
    cold, warm = builds(g, args.builds)
    dx, dl = g.alloc(xyz.nbytes), g.alloc(4 * Q)
    dx.upload(xyz)
    g.query_device(dx, Q, dl)
    qm = []
    for _ in range(5):
        g.query_device(dx, Q, dl)
        qm.append(g.stats()["query_ms"])
    lab = dl.download(np.empty(Q, np.int32))
    depth = int(g.stats()["depth"])
    dx.free()
    dl.free()
    g.close()
    nt = threads()
    o = oracle_lib.Oracle().set_lonlat(w["lon"], w["lat"], w["ring_off"])
    tb = []
    for _ in range(3):
        t = time.time()
        o.build()
        tb.append((time.time() - t) * 1e3)
    t = time.time()
    ref = o.query_xyz(xyz, nthreads=0)
    cq = time.time() - t
    omp_set_threads(1)
    t = time.time()
    o.build()
    tb1 = (time.time() - t) * 1e3
    omp_set_threads(nt)
    n1 = min(Q, 200_000)
    t = time.time()
    o.query_xyz(np.ascontiguousarray(xyz[:n1]), nthreads=1)
    cq1 = time.time() - t
    q = float(np.median(qm))
    res = {"n_edges": N, "n_queries": Q, "depth": depth,
           "gpu": {"build_ms_cold": round(cold[0], 3), "build_ms": round(warm[0], 3),
                   "query_gpu_ms": round(q, 4), "queries_per_s": round(Q / (q * 1e-3), 1)},
           "cpu_oracle": {"threads": nt, "build_ms": round(float(np.median(tb)), 3),
                          "queries_per_s": round(Q / cq, 1), "one_thread_build_ms": round(tb1, 3),
                          "one_thread_queries_per_s": round(n1 / cq1, 1)},
           "parity": {"checked": Q, "mismatches": int(np.count_nonzero(ref != lab))}}
    res["gpu_vs_cpu"] = {"build": round(res["cpu_oracle"]["build_ms"] / warm[0], 3),
                         "queries": round(res["gpu"]["queries_per_s"] / res["cpu_oracle"]["queries_per_s"], 1)}
    log(f"{name}: {res}")
    return res


def batch_run(items, xyz, qp, model, Q, bins=256, graph=1):
    """One batched build (cold + warm median of 3) and query pass (median of
    5) of the batch; returns (result dict, labels)."""
    import gahhg_amd
    from gahhg_amd import learned_tables
    g = gahhg_amd.GPISP(device=0, query_bins=bins, query_graph=graph)
    g.set_polygons_lonlat(items)
    if model is not None:
        learned_tables.apply(g, model, xyz, qp, Q)
    cold, warm = builds(g, 3)
    lay = g.table_layout()
    grid = int(sum(g.polygon_stats(k)["gpu_space_bytes"] for k in range(len(items))))
    dx, dp, dl = g.alloc(xyz.nbytes), g.alloc(4 * Q), g.alloc(4 * Q)
    dx.upload(xyz)
    dp.upload(qp)
    g.query_batch_device(dx, dp, Q, dl)
    qms = []
    tq = time.perf_counter()
    for _ in range(5):
        g.query_batch_device(dx, dp, Q, dl)
        qms.append(g.stats()["query_ms"])
    wall = (time.perf_counter() - tq) * 1e3 / 5
    lab = dl.download(np.empty(Q, dtype=np.int32))
    depths = [int(g.polygon_stats(k)["depth"]) for k in range(len(items))]
    for b in (dx, dp, dl):
        b.free()
    g.close()
    q = float(np.median(qms))
    res = {"build_ms": round(warm[0], 3), "build_ms_cold": round(cold[0], 3), "layout_ms": round(warm[3], 3),
           "query_gpu_ms": round(q, 3), "query_wall_ms": round(wall, 3), "queries_per_s": round(Q / (q * 1e-3), 1),
           "grid_bytes": grid, "table_bins": np.bincount(np.searchsorted([0, 16, 32, 64, 128, 256], lay[:, 0]),
                                                         minlength=6).tolist()}
    return res, lab, depths


def bench_batch(args, oracle_lib, model):
    """256 polygons (log-uniform 16..200k vertices), 64M Zipf-skewed queries:
    one batched build + one query pass (learned start tables, and the size
    rule for the A/B), against building and querying each polygon on its own
    with the plain path (the reference runs one polygon per process)."""
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
    graph = 0 if args.no_graph else 1
    rule, lab_rule, depths = batch_run(items, xyz, qp, None, Q, graph=graph)
    log(f"batch_256 size-rule tables: {rule}")
    learned = lab = None
    if model is not None:
        learned, lab, _ = batch_run(items, xyz, qp, model, Q, graph=graph)
        log(f"batch_256 learned tables: {learned}")
    final = lab if lab is not None else lab_rule
    # per-polygon runs: build each polygon alone, query its points (plain path)
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
            mism += int(np.count_nonzero(l2 != final[idx]))
    gp.close()
    # oracle on every point of 24 polygons (checker only)
    rng = np.random.default_rng(1)
    omism = checked = 0
    starts = np.concatenate([[0], np.cumsum(cnt)])
    for k in rng.choice(len(items), 24, replace=False):
        idx = order[starts[k]:starts[k + 1]]
        if not len(idx):
            continue
        o = oracle_lib.Oracle().set_lonlat(*items[k]).build()
        omism += int(np.count_nonzero(o.query_xyz(np.ascontiguousarray(xyz[idx])) != final[idx]))
        checked += len(idx)
    best = learned or rule
    res = {"n_polygons": len(items), "n_edges": n_edges, "n_queries": Q,
           "batched": best, "batched_size_rule_tables": rule if learned else None,
           "per_polygon_plain": {"build_ms_sum": round(pb, 3), "query_gpu_ms_sum": round(pq, 3),
                                 "queries_per_s": round(Q / (pq * 1e-3), 1) if pq else None},
           "build_ratio": round(best["build_ms"] / pb, 4) if pb else None,
           "query_speedup": round(pq / best["query_gpu_ms"], 3) if pq else None,
           "label_differences_vs_per_polygon": mism,
           "label_differences_learned_vs_size_rule": int(np.count_nonzero(lab != lab_rule)) if lab is not None else None,
           "oracle_checked": checked, "oracle_mismatches": omism}
    if learned:
        res["learned_vs_size_rule"] = {
            "build": round(learned["build_ms"] / rule["build_ms"], 4),
            "query": round(learned["query_gpu_ms"] / rule["query_gpu_ms"], 4),
            "build_plus_pass": round((learned["build_ms"] + learned["query_gpu_ms"]) /
                                     (rule["build_ms"] + rule["query_gpu_ms"]), 4)}
    log(f"batch_256: {res}")
    return res


def bench_fractal512(args, oracle_lib, model):
    """fractal_512 (512 polygons of the fractal_1m generator scaled to 4k-64k
    outer vertices, 32M points), one batch: learned start tables against the
    size rule on build ms, queries/s and bytes; the labels of the two must be
    equal (the hierarchy is the density rule's in both) and match the oracle
    on every point of a sample of polygons."""
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
    graph = 0 if args.no_graph else 1
    res = {"n_polygons": P, "n_edges": n_edges, "n_queries": Q}
    res["size_rule"], lab_rule, _ = batch_run(items, xyz, qp, None, Q, graph=graph)
    log(f"fractal_512 size rule: {res['size_rule']}")
    final = lab_rule
    if model is not None:
        res["learned"], lab, _ = batch_run(items, xyz, qp, model, Q, graph=graph)
        log(f"fractal_512 learned: {res['learned']}")
        res["label_differences_learned_vs_size_rule"] = int(np.count_nonzero(lab != lab_rule))
        r, l = res["size_rule"], res["learned"]
        res["learned_vs_size_rule"] = {"build": round(l["build_ms"] / r["build_ms"], 4),
                                       "query": round(l["query_gpu_ms"] / r["query_gpu_ms"], 4),
                                       "build_plus_pass": round((l["build_ms"] + l["query_gpu_ms"]) /
                                                                (r["build_ms"] + r["query_gpu_ms"]), 4)}
        final = lab
    rng = np.random.default_rng(7)
    om = on = 0
    for k in rng.choice(P, min(P, 24), replace=False):
        sel = np.nonzero(qp == k)[0]
        o = oracle_lib.Oracle().set_lonlat(*items[k]).build()
        om += int(np.count_nonzero(final[sel] != o.query_xyz(np.ascontiguousarray(xyz[sel]))))
        on += len(sel)
    res["oracle_checked"] = on
    res["oracle_mismatches"] = om
    log(f"fractal_512: {res}")
    return res


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



if __name__ == "__main__":
    main()
