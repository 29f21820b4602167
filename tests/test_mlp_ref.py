"""The CPU FP64 restatement of the learned model, pinned by mathematics:
finite differences for its gradient, torch's BF16 conversion for its
rounding, and the Adam closed form on a one-parameter case."""
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_ref as M  # noqa: E402


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 3.4e38, 1e-40]])
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    got = M.bf16(x)
    assert np.array_equal(got, ref)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = M.init_params(rng)
    X = rng.standard_normal((37, M.N_IN))
    dY = rng.standard_normal(37)
    g = M.backward(p, X, dY, mirror=False)
    f = lambda q: float(dY @ M.forward(q, X, mirror=False))
    idx = np.concatenate([rng.choice(np.arange(0, 1024).reshape(64, 16)[:, :12].ravel(), 40, replace=False),
                          rng.choice(np.arange(1024, M.N_PARAMS), 80, replace=False), [M.N_PARAMS - 1]])
    h = 1e-6
    for k in idx:
        e = np.zeros(M.N_PARAMS)
        e[k] = h
        fd = (f(p + e) - f(p - e)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, g[k])
    # padding columns get no gradient
    assert not g[:1024].reshape(64, 16)[:, 12:].any()


def test_adam_closed_form():
    # first step: m = (1-b1) g, v = (1-b2) g^2, bias-corrected ratio g/|g|
    p = np.ones(M.N_PARAMS)
    g = np.full(M.N_PARAMS, 3.0)
    out, m1, m2 = M.adam(p, g, np.zeros_like(p), np.zeros_like(p), 1, 0.01)
    k = M.N_PARAMS - 1
    assert abs(out[k] - (1.0 - 0.01 * 3.0 / (3.0 + 1e-8))) < 1e-12
    assert out[12] == 0.0  # a padding column of W1
