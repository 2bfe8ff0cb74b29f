"""CPU reference of the mesh message-passing network (test infrastructure).

The network maps node features X0 (N x 16) on the node-adjacency graph of a
mesh to one output per node:

    H0   = bf16(relu(X0 We + be))
    for each of L layers:
        A  = bf16(mean over graph neighbours of H)
        Z  = H W1 + A W2 + b
        H  = bf16(H + relu(Z))
    y    = H Wd + bd

with BF16 operands (weights and activations rounded to bfloat16, round to
nearest even) and FP32 accumulation, as the device kernels compute it.
`forward` returns the intermediate values the backward pass needs;
`backward` returns the gradients of the mean squared error against targets t,
with the device kernels' rounding points (upstream gradients are rounded to
BF16 before they enter a matrix product).  The arithmetic here is float64 on
the rounded operands.  `loss_and_grad_fp64` is the same network without any
rounding, for finite-difference checks of the gradient formulas.
"""
from __future__ import annotations

import numpy as np

F_IN = 16
HIDDEN = 64


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    b = np.ascontiguousarray(np.asarray(x, dtype=np.float32)).view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return b.view(np.float32).astype(np.float64)


def init_params(layers=3, seed=0, f_in=F_IN, hidden=HIDDEN):
    rng = np.random.default_rng(seed)
    p = {"We": rng.normal(0, np.sqrt(2.0 / f_in), (f_in, hidden)), "be": np.zeros(hidden)}
    for l in range(layers):
        p[f"W1_{l}"] = rng.normal(0, np.sqrt(1.0 / hidden), (hidden, hidden))
        p[f"W2_{l}"] = rng.normal(0, np.sqrt(1.0 / hidden), (hidden, hidden))
        p[f"b_{l}"] = np.zeros(hidden)
    p["Wd"] = rng.normal(0, np.sqrt(1.0 / hidden), (hidden, 1))
    p["bd"] = np.zeros(1)
    return p


def _mean_agg(ia, ja, H):
    deg = np.diff(ia)
    out = np.zeros_like(H)
    rows = np.repeat(np.arange(len(deg)), deg)
    np.add.at(out, rows, H[ja])
    return out / np.maximum(deg, 1)[:, None]


def _mean_agg_T(ia, ja, G):
    """Adjoint of the neighbour mean: out_j = sum over i with j in N(i) of G_i / deg_i."""
    deg = np.diff(ia)
    scaled = G / np.maximum(deg, 1)[:, None]
    rows = np.repeat(np.arange(len(deg)), deg)
    out = np.zeros_like(G)
    np.add.at(out, ja, scaled[rows])
    return out


def forward(p, X, ia, ja, layers=3, rounding=True):
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    X0 = r(X)
    Ze = X0 @ r(p["We"]) + p["be"]
    H = r(np.maximum(Ze, 0.0))
    cache = {"X0": X0, "Ze": Ze, "H": [H], "A": [], "Z": []}
    for l in range(layers):
        A = r(_mean_agg(ia, ja, H))
        Z = H @ r(p[f"W1_{l}"]) + A @ r(p[f"W2_{l}"]) + p[f"b_{l}"]
        H = r(H + np.maximum(Z, 0.0))
        cache["A"].append(A)
        cache["Z"].append(Z)
        cache["H"].append(H)
    y = (H @ r(p["Wd"]))[:, 0] + p["bd"][0]
    return y, cache


def backward(p, cache, y, t, ia, ja, layers=3, rounding=True):
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    N = len(y)
    g = {}
    dy = 2.0 * (y - t) / N
    H = cache["H"][layers]
    g["Wd"] = H.T @ dy[:, None]
    g["bd"] = np.array([dy.sum()])
    dH = dy[:, None] * r(p["Wd"])[:, 0][None, :]
    for l in reversed(range(layers)):
        Hin, A, Z = cache["H"][l], cache["A"][l], cache["Z"][l]
        dZ = r(dH * (Z > 0))
        g[f"W1_{l}"] = Hin.T @ dZ
        g[f"W2_{l}"] = A.T @ dZ
        g[f"b_{l}"] = dZ.sum(axis=0)
        dA = r(dZ @ r(p[f"W2_{l}"]).T)
        dH = dH + dZ @ r(p[f"W1_{l}"]).T + _mean_agg_T(ia, ja, dA)
    dZe = r(dH * (cache["Ze"] > 0))
    g["We"] = cache["X0"].T @ dZe
    g["be"] = dZe.sum(axis=0)
    return g


def loss(y, t):
    return float(np.mean((y - t) ** 2))


def loss_and_grad_fp64(p, X, t, ia, ja, layers=3):
    y, cache = forward(p, X, ia, ja, layers, rounding=False)
    return loss(y, t), backward(p, cache, y, t, ia, ja, layers, rounding=False)
