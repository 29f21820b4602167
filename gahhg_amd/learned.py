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
  9  U_r / max(C_r, 1)
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
