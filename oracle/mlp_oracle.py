"""numpy restatement of the planner MLP (3 -> 32 -> 32 ReLU -> 2), its
full-batch gradient and Adam — TEST INFRASTRUCTURE ONLY: the checker for the
HIP training kernels (gpu_msm_amd/csrc/mlp.hip). Pinned in tests by central
finite differences.

Parameter layout: W1[32][3] b1[32] W2[32][32] b2[32] W3[2][32] b3[2];
loss L = (1/n) sum_i sum_k 0.5 (y_ik - t_ik)^2.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 3, 32, 2
SHAPES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)), ("b3", (OUT,))]
NP = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def init(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            parts.append(rng.normal(0, np.sqrt(2.0 / shape[1]), size=shape).ravel())
        else:
            parts.append(np.zeros(shape).ravel())
    return np.concatenate(parts).astype(np.float32)


def forward(p: np.ndarray, x: np.ndarray):
    w = unpack(p.astype(np.float64))
    a1 = x @ w["W1"].T + w["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(a2, 0)
    y = h2 @ w["W3"].T + w["b3"]
    return y, (x, a1, h1, a2, h2)


def loss_grad(p: np.ndarray, x: np.ndarray, t: np.ndarray):
    n = len(x)
    w = unpack(p.astype(np.float64))
    y, (x, a1, h1, a2, h2) = forward(p, x)
    d = y - t
    loss = 0.5 * (d * d).sum() / n
    dy = d / n
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    dh2 = (dy @ w["W3"]) * (a2 > 0)
    g["W2"] = dh2.T @ h1
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ w["W2"]) * (a1 > 0)
    g["W1"] = dh1.T @ x
    g["b1"] = dh1.sum(0)
    return loss, np.concatenate([g[name].ravel() for name, _ in SHAPES])


def adam_train(p: np.ndarray, x: np.ndarray, t: np.ndarray, epochs: int, lr: float):
    p = p.astype(np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    curve = []
    for step in range(1, epochs + 1):
        loss, g = loss_grad(p, x, t)
        curve.append(loss)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p -= lr * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    return p.astype(np.float32), np.array(curve)
