"""CPU FP64 restatement of the learned refinement model -- test
infrastructure only (the product's model runs in the HIP kernels of
gahhg_amd/csrc/gahhg_mlp.hip; nothing here is shipped or measured).

The 12-64-64-1 perceptron with ReLU hidden layers and the library's
parameter layout (W1[64][16] with input columns 12..15 as zero padding,
b1[64], W2[64][64], b2[64], w3[64], b3[1]).  The model has no counterpart in
the reference; its rule for depth is divGrid's density stop
(source/GPISPAHHG.cu:1595-1605).

mirror=True rounds the inputs and the hidden-1 activations to BF16 (round to
nearest even) exactly where the GPU feeds them to its BF16 matrix cores;
mirror=False is the plain FP64 function of the same (BF16) weights.
"""
import numpy as np

N_IN = 12
OFF = {"W1": 0, "b1": 1024, "W2": 1088, "b2": 5184, "w3": 5248, "b3": 5312}
N_PARAMS = 5313


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, u & 0xFFFF0000, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return {"W1": p[0:1024].reshape(64, 16)[:, :N_IN], "b1": p[1024:1088], "W2": p[1088:5184].reshape(64, 64),
            "b2": p[5184:5248], "w3": p[5248:5312], "b3": p[5312]}


def forward(p, X, mirror=True, keep=False):
    w = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    Xr = bf16(X) if mirror else X
    Z1 = Xr @ w["W1"].T + w["b1"]
    H1 = np.maximum(Z1, 0.0)
    H1r = bf16(H1) if mirror else H1
    Z2 = H1r @ w["W2"].T + w["b2"]
    H2 = np.maximum(Z2, 0.0)
    y = H2 @ w["w3"] + w["b3"]
    if keep:
        return y, dict(Xr=Xr, Z1=Z1, H1r=H1r, Z2=Z2, H2=H2, w=w)
    return y


def backward(p, X, dY, mirror=True):
    """Gradient of sum_r dY[r] * y[r] with respect to the parameter vector
    (straight through the BF16 rounding of the activations)."""
    _, k = forward(p, X, mirror, keep=True)
    w = k["w"]
    dY = np.asarray(dY, dtype=np.float64)
    g = np.zeros(N_PARAMS)
    g[OFF["w3"]:OFF["w3"] + 64] = k["H2"].T @ dY
    g[OFF["b3"]] = dY.sum()
    D2 = (k["Z2"] > 0) * (dY[:, None] * w["w3"][None, :])
    g[OFF["W2"]:OFF["W2"] + 4096] = (D2.T @ k["H1r"]).reshape(-1)
    g[OFF["b2"]:OFF["b2"] + 64] = D2.sum(axis=0)
    D1 = (k["Z1"] > 0) * (D2 @ w["W2"])
    gW1 = np.zeros((64, 16))
    gW1[:, :N_IN] = D1.T @ k["Xr"]
    g[OFF["W1"]:OFF["W1"] + 1024] = gW1.reshape(-1)
    g[OFF["b1"]:OFF["b1"] + 64] = D1.sum(axis=0)
    return g


def adam(master, grad, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (FP64); padding columns of W1 stay zero."""
    m1 = b1 * m1 + (1 - b1) * grad
    m2 = b2 * m2 + (1 - b2) * grad * grad
    mh = m1 / (1 - b1 ** step)
    vh = m2 / (1 - b2 ** step)
    out = master - lr * mh / (np.sqrt(vh) + eps)
    pad = np.zeros(N_PARAMS, dtype=bool)
    pad[:1024] = (np.arange(1024) % 16) >= N_IN
    out[pad] = 0.0
    return out, m1, m2


def init_params(rng):
    """He-style random parameters (BF16-representable), padding zero."""
    p = np.zeros(N_PARAMS)
    W1 = np.zeros((64, 16))
    W1[:, :N_IN] = rng.standard_normal((64, N_IN)) * np.sqrt(2.0 / N_IN)
    p[0:1024] = W1.reshape(-1)
    p[1024:1088] = rng.standard_normal(64) * 0.1
    p[1088:5184] = rng.standard_normal(4096) * np.sqrt(2.0 / 64)
    p[5184:5248] = rng.standard_normal(64) * 0.1
    p[5248:5312] = rng.standard_normal(64) * np.sqrt(1.0 / 64)
    p[5312] = 0.1
    return bf16(p)
