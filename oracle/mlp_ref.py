"""FP64 CPU reference of the tuner's MLP (2 x 64 hidden ReLU, scalar output).

TEST INFRASTRUCTURE ONLY (tests/ imports it to check the GPU forward/backward
kernels of abhsf_amd/csrc/mlp.hip).  No reference counterpart exists; this is
pinned by central finite differences of its own loss (tests/test_tuner_cpu.py)
and by a closed-form check on a network with hand-set weights.

Parameter layout (same as the GPU): W1[64][F], b1[64], W2[64][64], b2[64],
W3[64], b3.  Loss = sum_b w_b (out_b - y_b)^2 / sum_b w_b.
"""
import numpy as np

H = 64


def layout(F):
    o = {}
    o["W1"] = (0, H * F)
    o["b1"] = (o["W1"][1], o["W1"][1] + H)
    o["W2"] = (o["b1"][1], o["b1"][1] + H * H)
    o["b2"] = (o["W2"][1], o["W2"][1] + H)
    o["W3"] = (o["b2"][1], o["b2"][1] + H)
    o["b3"] = (o["W3"][1], o["W3"][1] + 1)
    o["P"] = o["b3"][1]
    return o


def unpack(theta, F):
    L = layout(F)
    t = np.asarray(theta, dtype=np.float64)
    W1 = t[L["W1"][0]:L["W1"][1]].reshape(H, F)
    b1 = t[L["b1"][0]:L["b1"][1]]
    W2 = t[L["W2"][0]:L["W2"][1]].reshape(H, H)
    b2 = t[L["b2"][0]:L["b2"][1]]
    W3 = t[L["W3"][0]:L["W3"][1]]
    b3 = t[L["b3"][0]]
    return W1, b1, W2, b2, W3, b3


def forward(theta, X):
    X = np.asarray(X, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(theta, X.shape[1])
    z1 = X @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    out = h2 @ W3 + b3
    return h1, h2, out


def loss(theta, X, y, w=None):
    _, _, out = forward(theta, X)
    w = np.ones(len(y)) if w is None else np.asarray(w, dtype=np.float64)
    return float(np.sum(w * (out - y) ** 2) / np.sum(w))


def grad(theta, X, y, w=None):
    X = np.asarray(X, dtype=np.float64)
    F = X.shape[1]
    W1, b1, W2, b2, W3, b3 = unpack(theta, F)
    h1, h2, out = forward(theta, X)
    w = np.ones(len(y)) if w is None else np.asarray(w, dtype=np.float64)
    d3 = 2.0 * w * (out - np.asarray(y, dtype=np.float64)) / np.sum(w)
    d2 = np.outer(d3, W3) * (h2 > 0)
    d1 = (d2 @ W2) * (h1 > 0)
    g = np.concatenate([
        (d1.T @ X).ravel(), d1.sum(0),
        (d2.T @ h1).ravel(), d2.sum(0),
        d3 @ h2, [d3.sum()],
    ])
    return g


def init(F, seed=0):
    """He-uniform initial weights (same initialiser the tuner uses)."""
    rng = np.random.default_rng(seed)
    L = layout(F)
    t = np.zeros(L["P"])
    t[L["W1"][0]:L["W1"][1]] = rng.uniform(-1, 1, H * F) * np.sqrt(6.0 / F)
    t[L["W2"][0]:L["W2"][1]] = rng.uniform(-1, 1, H * H) * np.sqrt(6.0 / H)
    t[L["W3"][0]:L["W3"][1]] = rng.uniform(-1, 1, H) * np.sqrt(6.0 / H)
    return t
