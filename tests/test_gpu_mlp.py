"""The learned model's HIP kernels against the CPU FP64 model.

Forward and backward are compared (a) with the FP64 model that rounds to
BF16 exactly where the GPU feeds its BF16 matrix cores (differences come
only from FP32 accumulation: relative error <= 1e-4), and (b) with the plain
FP64 model of the same weights (relative error <= 1e-2, the BF16 bound).
The Adam kernel is compared with the FP64 step.
"""
import os
import sys

import numpy as np
import pytest

import gahhg_amd
from gahhg_amd import learned

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_ref as M  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def ctx():
    g = gahhg_amd.GPISP(device=0)
    yield g
    g.close()


@pytest.mark.parametrize("rows", [1, 16, 1000, 4093])
def test_forward_and_gradient(ctx, rows):
    say("mlp rows", rows)
    rng = np.random.default_rng(rows)
    p = M.init_params(rng)  # BF16-representable
    X = rng.standard_normal((rows, M.N_IN)).astype(np.float32)
    dY = rng.standard_normal(rows).astype(np.float32)
    m = learned.DeviceMLP(ctx, params=p.astype(np.float32))
    assert np.array_equal(learned.bf16_value(m.weights_bf16()), p)
    y = m.forward(X)
    ym = M.forward(p, X.astype(np.float64), mirror=True)
    yp = M.forward(p, X.astype(np.float64), mirror=False)
    assert rel(y, ym) <= 1e-4, rel(y, ym)
    # the BF16 bound is a norm-wise statement over a batch; a single output
    # can be a small difference of larger terms
    if rows >= 1000:
        assert rel(y, yp) <= 1e-2, rel(y, yp)
    g = m.backward(dY)
    gm = M.backward(p, X.astype(np.float64), dY.astype(np.float64), mirror=True)
    assert rel(g, gm) <= 1e-4, rel(g, gm)
    if rows >= 1000:
        # against the plain FP64 model: the gradient of the mean output (a
        # random-sign dY sums terms of both signs, which magnifies the BF16
        # rounding of the activations relative to the result)
        ones = np.full(rows, 1.0 / rows, dtype=np.float32)
        g1 = m.backward(ones)
        gp1 = M.backward(p, X.astype(np.float64), ones.astype(np.float64), mirror=False)
        assert rel(g1, gp1) <= 1e-2, rel(g1, gp1)
        g = m.backward(dY)
    # per block of the parameter vector too
    for a, b in [(0, 1024), (1024, 1088), (1088, 5184), (5184, 5248), (5248, 5312), (5312, 5313)]:
        assert rel(g[a:b], gm[a:b]) <= 1e-3, (a, b)
    m.free()


def test_adam_step(ctx):
    rng = np.random.default_rng(7)
    p = M.init_params(rng)
    X = rng.standard_normal((512, M.N_IN)).astype(np.float32)
    dY = rng.standard_normal(512).astype(np.float32)
    m = learned.DeviceMLP(ctx, params=p.astype(np.float32))
    mm1 = np.zeros(M.N_PARAMS)
    mm2 = np.zeros(M.N_PARAMS)
    ref = p.copy()
    for step in range(1, 4):
        g = m.backward(dY, X if step == 1 else None).astype(np.float64)
        m.adam(1e-3)
        ref, mm1, mm2 = M.adam(ref, g, mm1, mm2, step, 1e-3)
        got = m.params()
        assert np.max(np.abs(got - ref)) <= 1e-6 + 1e-5 * np.max(np.abs(ref))
        assert np.array_equal(m.weights_bf16(), learned.bf16_bits(got))
    assert not got[:1024].reshape(64, 16)[:, 12:].any()
    m.free()


def test_training_reduces_loss(ctx):
    """A few hundred Adam steps fit a smooth target."""
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (4096, M.N_IN)).astype(np.float32)
    t = (np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.5 * X[:, 3] ** 2).astype(np.float32)
    m = learned.DeviceMLP(ctx, seed=1)
    m.set_inputs(X)
    loss0 = None
    for step in range(300):
        y = m.forward()
        r = y - t
        loss = float(np.mean(r * r))
        loss0 = loss if loss0 is None else loss0
        m.backward(2 * r / len(r))
        m.adam(3e-3)
    assert loss < 0.1 * loss0, (loss0, loss)
    m.free()
