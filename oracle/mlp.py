"""Test infrastructure: a float64 numpy restatement of the learned schedule
dispatcher's network, used only to check the HIP kernels.

The network is 8 -> 64 -> 64 -> K with ReLU hidden layers, inputs
normalized by x' = (x - mu) / sd, mean-squared-error loss over all K
outputs. Parameters use the library's flat layout:
mu[8] sd[8] W1[64][8] b1[64] W2[64][64] b2[64] W3[K][64] b3[K].

The reference has no learned component (it always takes the shortest
transform length, mul_ntt.c:103-118), so this oracle is pinned by
mathematics instead: its analytic gradient is checked against central
finite differences, and its forward pass against a direct evaluation.
"""
from __future__ import annotations

import numpy as np

NI, NH = 8, 64


def layout(K: int) -> dict:
    off = {}
    pos = 0
    for name, size in (("mu", NI), ("sd", NI), ("W1", NH * NI), ("b1", NH), ("W2", NH * NH), ("b2", NH),
                       ("W3", K * NH), ("b3", K)):
        off[name] = (pos, size)
        pos += size
    off["total"] = pos
    return off


def unpack(p: np.ndarray, K: int) -> dict:
    L = layout(K)
    p = np.asarray(p, dtype=np.float64)
    g = {k: p[v[0]:v[0] + v[1]] for k, v in L.items() if k != "total"}
    g["W1"] = g["W1"].reshape(NH, NI)
    g["W2"] = g["W2"].reshape(NH, NH)
    g["W3"] = g["W3"].reshape(K, NH)
    return g


def init_params(K: int, mu, sd, seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    L = layout(K)
    p = np.zeros(L["total"], dtype=np.float32)

    def put(name, v):
        o, s = L[name]
        p[o:o + s] = np.asarray(v, dtype=np.float32).ravel()

    put("mu", mu)
    put("sd", sd)
    for name, fan_in, shape in (("W1", NI, (NH, NI)), ("W2", NH, (NH, NH)), ("W3", NH, (K, NH))):
        lim = np.sqrt(6.0 / fan_in)
        put(name, rng.uniform(-lim, lim, size=shape))
    return p


def forward(p, x, K: int):
    g = unpack(p, K)
    xn = (np.asarray(x, dtype=np.float64) - g["mu"]) / g["sd"]
    h1 = np.maximum(xn @ g["W1"].T + g["b1"], 0.0)
    h2 = np.maximum(h1 @ g["W2"].T + g["b2"], 0.0)
    y = h2 @ g["W3"].T + g["b3"]
    return xn, h1, h2, y


def loss(p, x, t, K: int) -> float:
    y = forward(p, x, K)[3]
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def grad(p, x, t, K: int) -> np.ndarray:
    """d loss / d p by backpropagation (mu and sd get 0)."""
    g = unpack(p, K)
    xn, h1, h2, y = forward(p, x, K)
    B = xn.shape[0]
    G3 = 2.0 * (y - np.asarray(t, dtype=np.float64)) / (B * K)
    dW3 = G3.T @ h2
    db3 = G3.sum(0)
    G2 = (G3 @ g["W3"]) * (h2 > 0)
    dW2 = G2.T @ h1
    db2 = G2.sum(0)
    G1 = (G2 @ g["W2"]) * (h1 > 0)
    dW1 = G1.T @ xn
    db1 = G1.sum(0)
    out = np.zeros(layout(K)["total"])
    L = layout(K)
    for name, v in (("W1", dW1), ("b1", db1), ("W2", dW2), ("b2", db2), ("W3", dW3), ("b3", db3)):
        o, s = L[name]
        out[o:o + s] = v.ravel()
    return out


def choose(p, x, K: int, mask=None) -> np.ndarray:
    """argmin over the candidates allowed by each row's mask bits."""
    y = forward(p, x, K)[3]
    if mask is not None:
        allowed = ((np.asarray(mask, dtype=np.uint64)[:, None] >> np.arange(K, dtype=np.uint64)) & 1) == 1
        y = np.where(allowed, y, np.inf)
    return np.argmin(y, axis=1)
