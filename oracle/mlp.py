"""TEST INFRASTRUCTURE ONLY -- CPU restatement (numpy, float64) of the
precision selector's network, used to check the library's HIP forward,
backward and Adam kernels.  The network has no counterpart in the reference.

Parameter layout (flat): W1[64][IN], b1[64], W2[64][64], b2[64], W3[OUT][64], b3[OUT].
Loss: mean over samples of weight_s * softmax cross-entropy(logits_s, label_s).
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 5, 64, 8
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)),
         ("b3", (OUT,))]
PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([d[name].ravel() for name, _ in SIZES])


def init(seed: int) -> np.ndarray:
    """He-initialised parameters, zero biases."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.standard_normal((H, IN)) * np.sqrt(2.0 / IN), "b1": np.zeros(H),
         "W2": rng.standard_normal((H, H)) * np.sqrt(2.0 / H), "b2": np.zeros(H),
         "W3": rng.standard_normal((OUT, H)) * np.sqrt(2.0 / H), "b3": np.zeros(OUT)}
    return pack(d).astype(np.float32)


def forward(p, X):
    d = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ d["W1"].T + d["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ d["W2"].T + d["b2"]
    h2 = np.maximum(a2, 0.0)
    z = h2 @ d["W3"].T + d["b3"]
    return z, (X, a1, h1, a2, h2)


def loss_and_grad(p, X, y, w=None):
    d = unpack(p)
    z, (X, a1, h1, a2, h2) = forward(p, X)
    nb = X.shape[0]
    w = np.ones(nb) if w is None else np.asarray(w, dtype=np.float64)
    zm = z.max(1, keepdims=True)
    e = np.exp(z - zm)
    se = e.sum(1, keepdims=True)
    loss = float(np.mean(w * (np.log(se[:, 0]) + zm[:, 0] - z[np.arange(nb), y])))
    d3 = (e / se)
    d3[np.arange(nb), y] -= 1.0
    d3 *= w[:, None] / nb
    g = {"W3": d3.T @ h2, "b3": d3.sum(0)}
    d2 = (d3 @ d["W3"]) * (a2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ d["W2"]) * (a1 > 0)
    g["W1"] = d1.T @ X
    g["b1"] = d1.sum(0)
    return loss, pack(g)


def adam(p, g, m, v, step, lr=1e-3, wd=0.0, b1=0.9, b2=0.999, eps=1e-8):
    g = g + wd * p
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
