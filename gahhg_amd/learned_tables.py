"""Learned start-table depth: the 12-64-64-1 perceptron choosing, per polygon,
how deep the query's start structure goes, without touching a label.

The hierarchy (and so every label) is the density rule's, as in the
reference (source/GPISPAHHG.cu:1595-1605).  A query first looks its point's
cube-map bin up in the polygon's start table; the entry is the deepest cell
every point of the bin reaches under the reference's first-match descent
(or the label itself when that cell is an empty leaf), so the descent that
follows ends in the same leaf whatever the table.  The table's depth is a
pure speed/memory trade: more bins (and sub-tables splitting the bins whose
entry is not final) cost build time and answer more points at once.

Candidates (bins per face side, sub-table side): gahhg_amd.TABLE_BINS /
TABLE_SUB, c = 0 .. 7; a batch considers c <= 4.  For every top-level cell r
of a polygon the model predicts y_r(c) = log2 of the query time per point
(ns) of the points falling in r; the build picks

    argmin_c  scale * sum_r w_r 2^y_r(c) + a_c + b_c N     (ns)

with w_r the points expected in r (from the query hint) and a_c + b_c N the
fitted table build time of one polygon (in a batch, whose table launches are
shared, a fitted cost per table bin).  Features are the depth model's 12 at
the polygon's last refined level, with feature 1 = log2(1 + bins x side) / 12
and feature 6 = log2(1 + w_r) / 32 (the expected query density: per-point
costs fall as points crowd the same bins and leaves).

Training (``python -m gahhg_amd.learned_tables --out F``, on the GPU): the
polygons of ``learned.training_polygons`` (seeds disjoint from the benchmark
workloads); each is built once per candidate (layout ms measured) and the
production query pass is timed on 0.5M..32M points of its family's
distribution under each candidate; forward, backward and Adam are the
project's HIP kernels (csrc/gahhg_mlp.hip).
"""
import json
import math
import os

import numpy as np

from . import TABLE_BINS, TABLE_SUB, TABLE_CAND_BATCH
from .learned import DeviceMLP, N_IN, N_PARAMS, sample_points, training_polygons

NC = len(TABLE_BINS)
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "table_model.json")
DEFAULT_SINGLE = 6  # the library's default for one polygon: 256 bins, sub-tables up to 8


def load_model(path=MODEL_PATH):
    with open(path) as f:
        d = json.load(f)
    return np.asarray(d["params"], dtype=np.float32), d["cost"], d.get("meta", {})


def cost_vector(cost, expected_points=0.0, scale=1.0):
    """The 19 constants gahhg_set_table_model takes."""
    out = [float(expected_points), float(scale)]
    for c in range(NC):
        out += [float(cost["a_ns"][c]), float(cost["b_ns_per_edge"][c])]
    return out + [float(cost.get("batch_ns_per_bin", 0.0))]


def apply(g, model, xyz, poly, n_points, sample=262_144, scale=1.0):
    """Query hint (every k-th point standing for k points) and the table model
    on a context, before its build."""
    params, cost = model[0], model[1]
    k = max(1, len(xyz) // sample)
    g.set_query_hint(xyz[::k], None if poly is None else poly[::k], float(k))
    g.set_table_model(params, cost_vector(cost, n_points, scale))


def size_rule_candidate(N, budget=8):
    """The candidate the size rule gives a polygon of N edges in a batch
    (the largest power of two of bins, 16 .. 128, with 6 B^2 <= budget N)."""
    B = 16
    while 2 * B <= 128 and 6 * (2 * B) ** 2 <= budget * N:
        B *= 2
    return TABLE_BINS.index(B)


def sample_plan(poly):
    """(uniform share, sample size, point counts timed) per sample of a
    training polygon: rings under three mixes at 1M and 4M points plus one
    32M-point sample timed at 16M and 32M; cap polygons at 0.5M, 2M, 8M."""
    if poly["kind"] == "mixed":
        return [(0.5, 4_000_000, (1_000_000, 4_000_000)), (0.9, 4_000_000, (1_000_000, 4_000_000)),
                (0.7, 32_000_000, (4_000_000, 16_000_000, 32_000_000))]
    return [(None, 8_000_000, (500_000, 2_000_000, 8_000_000))]


def collect(polys, reps=3, log=print, plan=sample_plan):
    """Per polygon (built alone, density-rule hierarchy): the layout ms of
    every candidate and, per sample and point count n (a prefix of the
    sample), the query ns per point of every candidate and the features with
    the n points as the hint (the expected points per top-level cell are a
    feature).  One record per (polygon, sample, n)."""
    from . import GPISP
    recs = []
    for i, poly in enumerate(polys):
        samples = [(sample_points(poly, size, 9000 + i + 100 * j, u if u is not None else 0.7), ns)
                   for j, (u, size, ns) in enumerate(plan(poly))]
        g = GPISP(device=0, query_sort=2, query_sub=16)
        g.set_polygon_lonlat(poly["lon"], poly["lat"], poly["ring_off"])
        N = len(poly["lon"]) - (len(poly["ring_off"]) - 1)
        dxs = []
        for xyz, _ in samples:
            dx = g.alloc(xyz.nbytes)
            dx.upload(xyz)
            dxs.append(dx)
        dl = g.alloc(4 * max(len(x) for x, _ in samples))
        lay = np.zeros(NC)
        keys = [(j, n) for j, (_, ns) in enumerate(samples) for n in ns]
        t = {k: np.zeros(NC) for k in keys}
        ref = None
        for c in range(NC):
            g.set_table_choice([c])
            lt = []
            for _ in range(reps + 1):
                g.build()
                lt.append(g.stats()["layout_ms"])
            lay[c] = float(np.median(lt[1:]))
            for (j, n) in keys:
                qt = []
                for _ in range(reps + 1):
                    g.query_device(dxs[j], n, dl)
                    qt.append(g.stats()["query_ms"])
                t[(j, n)][c] = float(np.median(qt[1:])) * 1e6 / n
                if (j, n) == keys[0]:  # labels must not depend on the table
                    lab = dl.download(np.empty(n, np.int32))
                    if ref is None:
                        ref = lab
                    elif not np.array_equal(ref, lab):
                        raise AssertionError(f"polygon {i}: candidate {c} changed labels")
        g.set_table_choice(None)
        for (j, n) in keys:
            xyz = samples[j][0][:n]
            k = max(1, n // 262_144)
            g.set_query_hint(xyz[::k], None, float(k))
            g.build()
            recs.append(dict(N=N, poly=i, n=n, lay=lay.tolist(), t=t[(j, n)].tolist(), X=g.table_features()[0]))
        log(f"[collect] polygon {i} (N={N}): layout ms " + " ".join(f"{v:.2f}" for v in lay) + "; ns/pt " +
            " | ".join(f"n={n}: " + " ".join(f"{v:.3f}" for v in t[(j, n)]) for (j, n) in keys))
        for dx in dxs:
            dx.free()
        dl.free()
        g.close()
    return recs


def batch_bin_cost(polys, reps=3, log=print):
    """ns per start-table bin in a batch: the batch polygons (kind "cap")
    built as one batch with every polygon on candidate c = 0 .. 4 in turn;
    least squares of layout ns over the batch's table bins."""
    from . import GPISP
    items = [(p["lon"], p["lat"], p["ring_off"]) for p in polys if p["kind"] == "cap"]
    if len(items) < 2:
        return 0.0
    g = GPISP(device=0)
    g.set_polygons_lonlat(items)
    xs, ys = [], []
    for c in range(TABLE_CAND_BATCH):
        g.set_table_choice(np.full(len(items), c, np.int32))
        lt = []
        for _ in range(reps + 1):
            g.build()
            lt.append(g.stats()["layout_ms"])
        xs.append(6.0 * TABLE_BINS[c] ** 2 * len(items))
        ys.append(float(np.median(lt[1:])) * 1e6)
        log(f"[collect] batch of {len(items)}, candidate {c}: layout {ys[-1] / 1e6:.3f} ms for {xs[-1]:.0f} bins")
    g.close()
    A = np.stack([np.ones(len(xs)), np.asarray(xs)], axis=1)
    coef = np.linalg.lstsq(A, np.asarray(ys), rcond=None)[0]
    return max(float(coef[1]), 0.0)


def fit_cost(recs):
    """Per candidate: layout ns = a_c + b_c N, non-negative least squares over
    the polygons (one value per polygon)."""
    from scipy.optimize import nnls
    seen = {}
    for r in recs:
        seen[r["poly"]] = r
    a, b = [], []
    for c in range(NC):
        A = np.array([[1.0, r["N"]] for r in seen.values()])
        y = np.array([r["lay"][c] * 1e6 for r in seen.values()])
        coef, _ = nnls(A, y)
        a.append(float(coef[0]))
        b.append(float(coef[1]))
    return {"a_ns": a, "b_ns_per_edge": b}


def groups(recs):
    """(features[42, 12], point shares[42], target ns/pt, record) per (record,
    candidate); the candidates of one record are consecutive."""
    G = []
    for ri, r in enumerate(recs):
        for c in range(NC):
            x = r["X"][:, c, :]
            G.append((x, x[:, 11].astype(np.float64), r["t"][c], ri))
    return G


def _block_loss_grad(y, S, T, lam=4.0, W=None):
    """Squared log2 error of each group's predicted time, plus lam times the
    squared error of each group's difference to its record's mean (the
    differences are what the argmin turns on)."""
    ng = len(T)
    y = y.astype(np.float64).reshape(ng, 42)
    ymax = y.max(axis=1, keepdims=True)
    e = S * np.exp2(y - ymax)
    P = e.sum(axis=1)
    pred = np.log2(np.maximum(P, 1e-300)) + ymax[:, 0]
    r = pred - T
    W = np.ones(ng) if W is None else W
    wsum = W.sum()
    g = 2.0 * W * r / wsum
    loss = float(np.sum(W * r * r) / wsum)
    if lam > 0:
        rb = r.reshape(-1, NC)
        d = rb - rb.mean(axis=1, keepdims=True)
        wb = W.reshape(-1, NC)
        loss += lam * float(np.sum(wb * d * d) / wsum)
        gd = 2.0 * lam * wb * d / wsum
        g += (gd - gd.mean(axis=1, keepdims=True)).reshape(-1)
    dy = g[:, None] * e / np.maximum(P, 1e-300)[:, None]
    return loss, dy


def _shares(G):
    S = np.stack([g[1] for g in G])
    return S / np.maximum(S.sum(axis=1, keepdims=True), 1e-30)


def train(ctx, G, steps=4000, lr=2e-3, seed=0, val=None, every=250, log=print):
    """Fits log2(sum_r share_r 2^y_r) to log2 ns/pt per (record, candidate)
    with the HIP forward, backward and Adam kernels; early stopping on the
    validation groups."""
    m = DeviceMLP(ctx, seed=seed)
    m.set_inputs(np.concatenate([g[0] for g in G]).astype(np.float32))
    S, T = _shares(G), np.log2(np.array([g[2] for g in G]))
    Wt = np.array([1.0 + (float(g[0][0, 10]) * 20.0 / 5.0) ** 2 for g in G])
    mv = None
    if val:
        mv = DeviceMLP(ctx, seed=seed)
        mv.set_inputs(np.concatenate([g[0] for g in val]).astype(np.float32))
        Sv, Tv = _shares(val), np.log2(np.array([g[2] for g in val]))
    best = (float("inf"), m.params(), -1)
    hist = []
    for step in range(steps + 1):
        if mv is not None and step % every == 0:
            mv.set_params(m.params())
            lv, _ = _block_loss_grad(mv.forward(), Sv, Tv)
            if lv < best[0]:
                best = (lv, m.params(), step)
        if step == steps:
            break
        loss, dy = _block_loss_grad(m.forward(), S, T, W=Wt)
        m.backward(dy.reshape(-1).astype(np.float32))
        m.adam(lr if step < steps * 0.7 else lr * 0.2)
        if step % 500 == 0 or step == steps - 1:
            log(f"[train] step {step} loss {loss:.5f}" +
                (f", best validation {math.sqrt(best[0]):.4f} at step {best[2]}" if mv is not None else ""))
            hist.append((step, loss))
    params = best[1] if mv is not None else m.params()
    m.free()
    if mv is not None:
        mv.free()
    return params, hist, (math.sqrt(best[0]) if mv is not None else None, best[2])


def predict(ctx, params, recs, ids):
    """Predicted log2 ns/pt [record][candidate] for the records ids."""
    m = DeviceMLP(ctx, params=params)
    out = {}
    for ri in ids:
        X = recs[ri]["X"]
        y = m.forward(X.reshape(-1, N_IN)).astype(np.float64).reshape(42, NC)
        S = X[:, 0, 11].astype(np.float64)
        S = S / max(S.sum(), 1e-30)
        out[ri] = np.log2(np.maximum((S[:, None] * np.exp2(y)).sum(axis=0), 1e-300))
    m.free()
    return out


def evaluate(pred, cost, recs, ids, batch=False):
    """A/B on held-out records: total time table build + Q points (ms) of the
    learned choice, of the library default (one polygon: 256 bins with
    sub-tables up to 8; batch: the size rule of query_table_budget = 8) and
    of the best candidate.  In a batch the table build is the batch's per-bin
    cost (its launches are shared); alone, the measured layout ms."""
    nc = TABLE_CAND_BATCH if batch else NC
    kb = cost.get("batch_ns_per_bin", 0.0)
    out = []
    for ri in ids:
        r = recs[ri]
        if batch:
            build = [kb * 6.0 * TABLE_BINS[c] ** 2 * 1e-6 for c in range(nc)]
            est_b = build
        else:
            build = [r["lay"][c] for c in range(nc)]
            est_b = [(cost["a_ns"][c] + cost["b_ns_per_edge"][c] * r["N"]) * 1e-6 for c in range(nc)]
        Q = r["n"]
        tot = [build[c] + Q * r["t"][c] * 1e-6 for c in range(nc)]
        est = [est_b[c] + Q * 2.0 ** pred[ri][c] * 1e-6 for c in range(nc)]
        learned = int(np.argmin(est))
        default = size_rule_candidate(r["N"]) if batch else DEFAULT_SINGLE
        best = int(np.argmin(tot))
        out.append(dict(N=r["N"], n=Q, learned=learned, default=default, best=best, learned_ms=tot[learned],
                        default_ms=tot[default], best_ms=tot[best]))
    return out


def main(argv=None):
    import argparse
    import sys
    import time
    from . import GPISP
    ap = argparse.ArgumentParser(description="Train the learned start-table depth model on the GPU.")
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--batch-polygons", type=int, default=120)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--quick", action="store_true", help="a few small polygons (pipeline check)")
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    t0 = time.time()
    if args.quick:
        polys = training_polygons(n_batch=10, rings=(30_000, 60_000))
    else:
        polys = training_polygons(n_batch=args.batch_polygons)
    log(f"[train] {len(polys)} training polygons generated in {time.time() - t0:.1f}s")
    recs = collect(polys, log=log)
    cost = fit_cost(recs)
    cost["batch_ns_per_bin"] = batch_bin_cost(polys, log=log)
    log(f"[train] batch table cost {cost['batch_ns_per_bin']:.4f} ns per bin")
    log(f"[train] table cost a_ns {[round(v) for v in cost['a_ns']]} b_ns/edge "
        f"{[round(v, 4) for v in cost['b_ns_per_edge']]}")
    G = groups(recs)
    val_ids = {ri for ri in range(len(recs)) if recs[ri]["poly"] % 5 == 3}
    hold = {ri for ri in range(len(recs)) if recs[ri]["poly"] % 5 == 4}
    Gt = [g for g in G if g[3] not in hold and g[3] not in val_ids]
    Gval = [g for g in G if g[3] in val_ids]
    ctx = GPISP(device=0)
    params, hist, (vbest, vstep) = train(ctx, Gt, steps=args.steps, val=Gval, log=log)
    log(f"[train] validation rms log2 error {vbest:.4f} at step {vstep}")
    pred = predict(ctx, params, recs, sorted(hold))
    err = [pred[ri][c] - math.log2(recs[ri]["t"][c]) for ri in hold for c in range(NC)]
    rmse = float(np.sqrt(np.mean(np.square(err)))) if err else None
    log(f"[train] held-out rms log2 error {rmse}")
    ab = {}
    for batch in (False, True):
        evs = evaluate(pred, cost, recs, sorted(hold), batch=batch)
        for Q in sorted({e["n"] for e in evs}):
            ev = [e for e in evs if e["n"] == Q]
            dt = sum(e["default_ms"] for e in ev)
            lt = sum(e["learned_ms"] for e in ev)
            bt = sum(e["best_ms"] for e in ev)
            key = ("batch" if batch else "single") + f"_{int(Q)}"
            ab[key] = {"records": len(ev), "default_ms": round(dt, 3), "learned_ms": round(lt, 3),
                       "best_ms": round(bt, 3), "learned_vs_default": round(lt / dt, 4),
                       "best_vs_default": round(bt / dt, 4),
                       "choices": np.bincount([e["learned"] for e in ev], minlength=NC).tolist()}
            log(f"[train] held-out A/B {key}: {ab[key]}")
    ctx.close()
    meta = {"training_polygons": len(polys), "records": len(recs), "groups": len(G),
            "validation_rms_log2_error": vbest, "validation_best_step": vstep, "held_out_rms_log2_error": rmse,
            "held_out_ab": ab, "loss_history": hist,
            "candidates": {"bins": list(TABLE_BINS), "sub": list(TABLE_SUB)},
            "measurements": [{"N": r["N"], "poly": r["poly"], "n": r["n"], "layout_ms": r["lay"],
                              "ns_per_point": r["t"]} for r in recs]}
    with open(args.out, "w") as f:
        json.dump({"params": [float(x) for x in params], "cost": cost, "meta": meta}, f)
    log(f"[train] model written to {args.out}")


if __name__ == "__main__":
    main()
