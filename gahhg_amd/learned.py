"""Learned refinement depth: the 12-64-64-1 perceptron on the GPU.

The model (BF16 weights, FP32 accumulation; HIP kernels in
csrc/gahhg_mlp.hip) predicts, for one top-level cell of a polygon's
hierarchy at the level being refined, the log2 of the query time per point
(ns) that points falling in that cell will cost if refinement stops now
(feature "refine" = 0) or after one more level (= 1).  The build weighs the
predicted saving over the expected number of points against a build-cost
model and refines while it pays.  The reference has only the density stop of
divGrid (source/GPISPAHHG.cu:1595-1605); this is an addition.

Features per (polygon, top-level cell r, level l, refine k), all O(1):
  0  l / 16
  1  k
  2  log2(1 + N_r) / 20          N_r: edges whose first vertex lies in r
  3  log2(1 + U_r) / 20          U_r: cells of r crossed at level l
  4  log2(1 + C_r) / 20          C_r: cells of r at level l
  5  min(U_r / max(N_r, 1), 4)
  6  log2(1 + N_r / max(U_r, 1)) / 10
  7  log2(mean edge length / cell size at l) / 10, cell size 0.55 / 2^l rad
  8  mean turning angle / pi
  9  share of r's hint points in crossed cells of level l (the boundary
     band); U_r / max(C_r, 1) when there is no hint
  10 log2(1 + N) / 20            N: edges of the polygon
  11 w_r                          share of the polygon's points in r (edge
                                 share N_r / N when no points are expected)
"""
import ctypes
import json
import math
import os

import numpy as np

from . import GahhgError, lib

N_IN = 12
N_PARAMS = 5313
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "depth_model.json")


def bf16_bits(x):
    """float -> BF16 bit patterns (round to nearest even), uint16."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    r = np.where(nan, u >> 16, r)
    return r.astype(np.uint16)


def bf16_value(bits):
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, dtype=np.float32)
    W1 = np.zeros((64, 16), dtype=np.float32)
    W1[:, :N_IN] = rng.standard_normal((64, N_IN)) * math.sqrt(2.0 / N_IN)
    p[0:1024] = W1.reshape(-1)
    p[1088:5184] = rng.standard_normal(4096) * math.sqrt(2.0 / 64)
    p[5248:5312] = rng.standard_normal(64) * math.sqrt(1.0 / 64)
    return p


class DeviceMLP:
    """Parameters of the model on the GPU of a GPISP context: FP32 master
    copy, Adam moments and the BF16 weights the kernels read."""

    def __init__(self, ctx, params=None, seed=0):
        self.g = ctx
        self.L = lib()
        self.master = ctx.alloc(4 * N_PARAMS)
        self.m1 = ctx.alloc(4 * N_PARAMS)
        self.m2 = ctx.alloc(4 * N_PARAMS)
        self.grad = ctx.alloc(4 * N_PARAMS)
        self.w = ctx.alloc(2 * N_PARAMS + 16)
        self.step_no = 0
        self.set_params(init_params(seed) if params is None else params)
        self._x = None
        self._xrows = 0

    def _check(self, rc):
        if rc:
            raise GahhgError(rc, self.L.gahhg_last_error(self.g.h).decode())

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == N_PARAMS
        self.master.upload(p)
        self.m1.upload(np.zeros(N_PARAMS, dtype=np.float32))
        self.m2.upload(np.zeros(N_PARAMS, dtype=np.float32))
        self.w.upload(np.concatenate([bf16_bits(p), np.zeros(8, dtype=np.uint16)]))
        self.step_no = 0

    def params(self):
        return self.master.download(np.empty(N_PARAMS, dtype=np.float32))

    def weights_bf16(self):
        return self.w.download(np.empty(N_PARAMS + 8, dtype=np.uint16))[:N_PARAMS]

    def set_inputs(self, X):
        """Keeps a feature matrix resident on the GPU for repeated calls."""
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, N_IN)
        if self._x is not None:
            self._x.free()
        self._x = self.g.alloc(max(X.nbytes, 4))
        self._x.upload(X)
        self._xrows = X.shape[0]
        self._y = self.g.alloc(max(4 * self._xrows, 4))
        self._dy = self.g.alloc(max(4 * self._xrows, 4))

    def forward(self, X=None):
        if X is not None:
            self.set_inputs(X)
        n = self._xrows
        self._check(self.L.gahhg_mlp_forward(self.g.h, self._x.ptr, n, self.w.ptr, self._y.ptr))
        return self._y.download(np.empty(n, dtype=np.float32))

    def backward(self, dY, X=None):
        if X is not None:
            self.set_inputs(X)
        n = self._xrows
        self._dy.upload(np.ascontiguousarray(dY, dtype=np.float32))
        self._check(self.L.gahhg_mlp_backward(self.g.h, self._x.ptr, n, self.w.ptr, self._dy.ptr, self.grad.ptr))
        return self.grad.download(np.empty(N_PARAMS, dtype=np.float32))

    def adam(self, lr):
        """One Adam step with the gradient of the last backward call."""
        self.step_no += 1
        self._check(self.L.gahhg_mlp_adam(self.g.h, self.master.ptr, self.grad.ptr, self.m1.ptr, self.m2.ptr,
                                          self.step_no, ctypes.c_float(lr), self.w.ptr))

    def free(self):
        for b in (self.master, self.m1, self.m2, self.grad, self.w, self._x,
                  getattr(self, "_y", None), getattr(self, "_dy", None)):
            if b is not None:
                b.free()


def save_model(path, params, cost, meta):
    with open(path, "w") as f:
        json.dump({"params": [float(x) for x in params], "cost": cost, "meta": meta}, f)


def load_model(path=MODEL_PATH):
    with open(path) as f:
        d = json.load(f)
    return np.asarray(d["params"], dtype=np.float32), d["cost"], d.get("meta", {})


# ---------------------------------------------------------------------------
# Training data: measured build and query times of fixed-depth hierarchies
# ---------------------------------------------------------------------------
def training_polygons(seed=101, n_batch=120, rings=(30_000, 60_000, 120_000, 250_000, 500_000, 800_000, 1_200_000,
                                                   700_000, 1_000_000, 1_400_000, 1_700_000, 2_000_000, 900_000)):
    """Training set (disjoint from the benchmark workloads by seed): polygons
    of the batch generator and fractal rings of several sizes, each with the
    point distribution of its benchmark family."""
    from . import datasets as ds
    out = []
    for p in ds.batch_polygons(n_poly=n_batch, seed=seed):
        out.append(dict(kind="cap", lon=p["lon"], lat=p["lat"], ring_off=p["ring_off"], centre=p["centre"],
                        radius=p["radius"]))
    for j, n in enumerate(rings):
        w = ds.fractal_1m(n_queries=0, seed=seed + 1 + j, n_outer=n, n_hole=max(1000, n // 20))
        out.append(dict(kind="mixed", lon=w["lon"], lat=w["lat"], ring_off=w["ring_off"]))
    return out


def sample_points(poly, n, seed, uniform=0.7):
    from . import datasets as ds, query_lonlat_to_xyz
    rng = np.random.default_rng(seed)
    if poly["kind"] == "cap":
        lon, lat = ds.cap_queries(rng, poly["centre"][0], poly["centre"][1], 1.3 * poly["radius"], n)
    else:
        lon, lat = ds.mixed_queries(rng, poly["lon"], poly["lat"], poly["ring_off"], n, uniform)
    return query_lonlat_to_xyz(lon, lat)


def collect(polys, n_points=2_000_000, span=(-3, 2), reps=3, mixes=(0.5, 0.7, 0.9), log=print):
    """Per polygon: the density-rule depth D, then for every fixed depth d in
    [D + span[0], D + span[1]] (1..15) the build time (ms) and the query time
    per point (ns, points grouped as in production), and the model features
    of every level of the deepest hierarchy with the sample as query hint.
    Rings are timed under several point mixes (uniform share), one record
    each: the hierarchy is the same, the points and hence the features and
    times differ."""
    from . import GPISP
    recs = []
    for i, poly in enumerate(polys):
        mix_list = mixes if poly["kind"] == "mixed" else (None,)
        samples = [sample_points(poly, n_points, 7000 + i + 100 * j, u if u is not None else 0.7)
                   for j, u in enumerate(mix_list)]
        g = GPISP(device=0, query_sort=2)
        g.set_polygon_lonlat(poly["lon"], poly["lat"], poly["ring_off"]).build()
        D = g.stats()["depth"]
        dxs = []
        for xyz in samples:
            dx = g.alloc(xyz.nbytes)
            dx.upload(xyz)
            dxs.append(dx)
        dl = g.alloc(4 * n_points)
        ds_ = list(range(max(1, D + span[0]), min(15, D + span[1]) + 1))
        t = [dict() for _ in samples]
        b, cells = {}, {}
        for d in ds_:
            g.set_polygon_lonlat(poly["lon"], poly["lat"], poly["ring_off"]).set_polygon_depths([d])
            bt = []
            for _ in range(reps + 1):
                g.build()
                bt.append(g.stats()["pre_ms"])
            b[d] = float(np.median(bt[1:]))
            for j, xyz in enumerate(samples):
                qt = []
                for _ in range(reps + 1):
                    g.query_device(dxs[j], len(xyz), dl)
                    qt.append(g.stats()["query_ms"])
                t[j][d] = float(np.median(qt[1:])) * 1e6 / len(xyz)
            tr, _ = g.build_trace()
            cells[d] = tr[0, :, :, 0].sum(axis=1).tolist()
        N = len(poly["lon"]) - (len(poly["ring_off"]) - 1)
        for j, xyz in enumerate(samples):
            # features of the deepest hierarchy (levels below its bottom)
            g.set_query_hint(xyz[::16], None, 16.0)
            X = (np.stack([g.depth_features(l)[0] for l in range(max(ds_) - 1)]) if max(ds_) > 1
                 else np.zeros((0, 42, 2, 12)))
            recs.append(dict(N=N, rule=D, t=t[j], b=b, cells=cells, X=X, poly=i))
            log(f"[collect] polygon {i} mix {j}: N={N} rule depth {D}; ns/pt " +
                " ".join(f"{d}:{t[j][d]:.3f}" for d in ds_) + "; build ms " + " ".join(f"{d}:{b[d]:.2f}" for d in ds_))
        for dx in dxs:
            dx.free()
        dl.free()
        g.close()
    return recs


def groups(recs):
    """(features[42, 12], point shares[42], target ns/pt) per (polygon,
    level l, k): stop at l (k = 0, depth l + 1) or refine once (depth l + 2)."""
    G = []
    for ri, r in enumerate(recs):
        for l in range(r["X"].shape[0]):
            for k in (0, 1):
                d = l + 1 + k
                if d in r["t"] and (l + 1) in r["t"] and (l + 2) in r["t"]:
                    x = r["X"][l, :, k, :]
                    G.append((x, x[:, 11].astype(np.float64), r["t"][d], ri))
    return G


def fit_build_cost(recs):
    """ns per edge, per next-level cell and per level from the build-time
    increments of consecutive depths (non-negative least squares)."""
    from scipy.optimize import nnls
    A, y = [], []
    for r in recs:
        ds_ = sorted(r["b"])
        for d0, d1 in zip(ds_, ds_[1:]):
            new_cells = r["cells"][d1][d0] if d0 < len(r["cells"][d1]) else 0
            A.append([r["N"], new_cells, 1.0])
            y.append((r["b"][d1] - r["b"][d0]) * 1e6)
    coef, _ = nnls(np.asarray(A, dtype=np.float64), np.asarray(y, dtype=np.float64))
    return [float(c) for c in coef]


def _group_loss_grad(y, S, T, lam=4.0, W=None):
    """Loss over groups (consecutive pairs are the stop / refine groups of
    one polygon and level): squared log2 error of each group, plus lam times
    the squared error of each pair's difference -- the quantity the build's
    decision turns on."""
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
    npair = ng // 2
    if npair and lam > 0:
        d = (pred[1::2] - pred[0::2]) - (T[1::2] - T[0::2])
        wp = W[0::2]
        loss += lam * float(np.sum(wp * d * d) / wp.sum())
        g[1::2] += 2.0 * lam * wp * d / wp.sum()
        g[0::2] -= 2.0 * lam * wp * d / wp.sum()
    dy = g[:, None] * e / np.maximum(P, 1e-300)[:, None]
    return loss, dy


def _shares(G):
    S = np.stack([g[1] for g in G])
    return S / np.maximum(S.sum(axis=1, keepdims=True), 1e-30)


def train(ctx, G, steps=4000, lr=2e-3, seed=0, val=None, every=250, log=print):
    """Fits the model so that, per group, log2(sum_r share_r 2^y_r) matches
    log2 of the measured ns per point.  Forward, backward and Adam run in
    the HIP kernels; the per-group loss and its upstream gradient are formed
    on the host.  With validation groups, the parameters of the step with
    the lowest validation loss are returned (early stopping)."""
    m = DeviceMLP(ctx, seed=seed)
    m.set_inputs(np.concatenate([g[0] for g in G]).astype(np.float32))
    S, T = _shares(G), np.log2(np.array([g[2] for g in G]))
    # larger polygons weigh more: their decisions move more time
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
            lv, _ = _group_loss_grad(mv.forward(), Sv, Tv)
            if lv < best[0]:
                best = (lv, m.params(), step)
        if step == steps:
            break
        loss, dy = _group_loss_grad(m.forward(), S, T, W=Wt)
        m.backward(dy.reshape(-1).astype(np.float32))
        m.adam(lr if step < steps * 0.7 else lr * 0.2)
        if step % 500 == 0 or step == steps - 1:
            log(f"[train] step {step} loss {loss:.5f} (rms log2 error {math.sqrt(loss):.4f})"
                + (f", best validation {math.sqrt(best[0]):.4f} at step {best[2]}" if mv is not None else ""))
            hist.append((step, loss))
    params = best[1] if mv is not None else m.params()
    m.free()
    if mv is not None:
        mv.free()
    return params, hist, (math.sqrt(best[0]) if mv is not None else None, best[2])


def predict_groups(ctx, params, G):
    m = DeviceMLP(ctx, params=params)
    X = np.concatenate([g[0] for g in G]).astype(np.float32)
    y = m.forward(X).astype(np.float64).reshape(len(G), 42)
    m.free()
    S = np.stack([g[1] for g in G])
    S = S / np.maximum(S.sum(axis=1, keepdims=True), 1e-30)
    return np.log2(np.maximum((S * np.exp2(y)).sum(axis=1), 1e-300))


K_FLOOR = 0.02   # the build's window: refine while crossed cells / N < K_FLOOR
K_ABOVE = 2      # at most this many levels past the density rule's stop
MARGIN = 2.0     # predicted gain factor needed to depart from the density rule


def simulate_depth(y, U, crossed, N, w_total, cells_next, cost):
    """The build's decision sequence for one polygon from per-level model
    outputs y[l][r][k], crossed cells U[l] (crossed[l][r] per top-level
    cell), cells of level l + 1 and the
    cost constants: returns the depth the build would pick."""
    rs = -1
    for l in range(len(y)):
        k = U[l] / N
        if rs < 0 and (k > 0.1 or (k < 0.1 and 0.1 - k < 0.01)):
            rs = l
        if k < K_FLOOR:
            continue
        if rs >= 0 and l >= rs + K_ABOVE:
            return l + 1
        live = crossed[l] > 0  # top-level cells with nothing crossed keep their cost
        saving = float(np.sum((w_total[l] * (np.exp2(y[l][:, 0]) - np.exp2(y[l][:, 1])))[live]))
        c = cost[1] * N + cost[2] * cells_next[l] + cost[3]
        past = k > 0.1 or (k < 0.1 and 0.1 - k < 0.01)
        refine = saving > MARGIN * c if past else saving * MARGIN > c
        if not refine:
            return l + 1
    return len(y) + 1


def evaluate(ctx, params, cost, recs, ids, Q):
    """A/B of the learned depth against the density rule on polygons the
    model did not train on: total time build + Q points per polygon (ms),
    from the measured build and query times of each depth."""
    m = DeviceMLP(ctx, params=params)
    out = []
    for ri in ids:
        r = recs[ri]
        X = r["X"]
        if X.shape[0] == 0:
            continue
        y = m.forward(X.reshape(-1, N_IN)).astype(np.float64).reshape(X.shape[0], 42, 2)
        Ur = np.rint(np.exp2(X[:, :, 0, 3].astype(np.float64) * 20.0) - 1.0)
        U = Ur.sum(axis=1)
        shares = X[:, :, 0, 11].astype(np.float64)
        deep = max(r["cells"])
        cells_next = [r["cells"][deep][l + 1] if l + 1 < len(r["cells"][deep]) else 0 for l in range(X.shape[0])]
        d = simulate_depth(y, U, Ur, r["N"], shares * Q, cells_next, [1.0, cost[0], cost[1], cost[2]])
        if d not in r["t"]:
            continue
        tot = {dd: r["b"][dd] + Q * r["t"][dd] * 1e-6 for dd in r["t"]}
        best = min(tot, key=tot.get)
        out.append(dict(N=r["N"], rule=r["rule"], learned=d, best=best, rule_ms=tot[r["rule"]],
                        learned_ms=tot[d], best_ms=tot[best]))
    m.free()
    return out


def main(argv=None):
    import argparse
    import sys
    import time
    from . import GPISP
    ap = argparse.ArgumentParser(description="Train the learned refinement-depth model on the GPU.")
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--points", type=int, default=2_000_000)
    ap.add_argument("--batch-polygons", type=int, default=120)
    ap.add_argument("--steps", type=int, default=4000)
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    t0 = time.time()
    polys = training_polygons(n_batch=args.batch_polygons)
    log(f"[train] {len(polys)} training polygons generated in {time.time() - t0:.1f}s")
    recs = collect(polys, n_points=args.points, log=log)
    cost = fit_build_cost(recs)
    log(f"[train] build cost: {cost[0]:.3f} ns/edge, {cost[1]:.2f} ns/cell, {cost[2]:.0f} ns/level")
    G = groups(recs)
    # polygons split three ways: training, validation (early stopping) and
    # held out (reported)
    val_ids = {ri for ri in range(len(recs)) if recs[ri]["poly"] % 5 == 3}
    hold = {ri for ri in range(len(recs)) if recs[ri]["poly"] % 5 == 4}
    Gt = [g for g in G if g[3] not in hold and g[3] not in val_ids]
    Gval = [g for g in G if g[3] in val_ids]
    Gv = [g for g in G if g[3] in hold]
    ctx = GPISP(device=0)
    params, hist, (vbest, vstep) = train(ctx, Gt, steps=args.steps, val=Gval, log=log)
    log(f"[train] validation rms log2 error {vbest:.4f} at step {vstep}")
    pv = predict_groups(ctx, params, Gv)
    tv = np.log2([g[2] for g in Gv])
    rmse = float(np.sqrt(np.mean((pv - tv) ** 2))) if len(Gv) else None
    log(f"[train] held-out groups {len(Gv)}: rms log2 error {rmse}")
    cost_v = [cost[0], cost[1], cost[2]]
    ab = {}
    for Q in (1e5, 1e6, 1e7, 1e8):
        ev = evaluate(ctx, params, cost_v, recs, sorted(hold), Q)
        if ev:
            rt = sum(e["rule_ms"] for e in ev)
            lt = sum(e["learned_ms"] for e in ev)
            bt = sum(e["best_ms"] for e in ev)
            ab[str(int(Q))] = {"polygons": len(ev), "rule_ms": round(rt, 3), "learned_ms": round(lt, 3),
                               "best_ms": round(bt, 3), "learned_vs_rule": round(lt / rt, 4),
                               "depth_changes": sum(e["learned"] != e["rule"] for e in ev)}
            log(f"[train] held-out A/B at {Q:.0e} points per polygon: {ab[str(int(Q))]}")
    ctx.close()
    meta = {"training_polygons": len(polys), "groups": len(G), "held_out_groups": len(Gv),
            "validation_rms_log2_error": vbest, "validation_best_step": vstep, "held_out_ab": ab,
            "held_out_rms_log2_error": rmse, "loss_history": hist, "points_per_polygon": args.points,
            "records": [{"N": r["N"], "rule": r["rule"], "ns_per_point": r["t"], "build_ms": r["b"]} for r in recs]}
    save_model(args.out, params, {"ns_edge": cost[0], "ns_cell": cost[1], "ns_level": cost[2]}, meta)
    log(f"[train] model written to {args.out}")


if __name__ == "__main__":
    main()
