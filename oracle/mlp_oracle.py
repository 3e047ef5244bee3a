"""FP64 numpy restatement of the parametric encoder -- TEST INFRASTRUCTURE.

Only tests/ import this module; the product is the HIP library.  The encoder
has no counterpart in the reference; its loss is the reference's pair loss
(gradient.rs:98-151): for a pair (i, j) with u = y_i - y_j, D = 1 + |u|^2,
NN w_nn D/(10 + D), MN w_mn D/(10000 + D), FP w_fp/(1 + D), self pairs
skipped, and the gradient of process_pairs (g_i += w1 u, g_j -= w1 u with
w1 = w c2/(c + D)^2 for NN/MN, g_i -= w1 u for FP with w1 = 2 w/(1 + D)^2).

``emulate=True`` rounds to bfloat16 (round to nearest even) where the GPU
path does: the input rows, W1 and W2, the first hidden activation, and the
two hidden-layer deltas.  Everything else is FP64.

Pinning: ``loss_and_grad(emulate=False)`` is checked against central finite
differences, and its dL/dY against the C oracle of gradient.rs
(``oracle.pacmap_grad``) on the same outputs.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

_SQRT1_2 = 0.7071067811865476
_INV_SQRT_2PI = 0.3989422804014327


def bf16(a):
    """Round float32 values to bfloat16 (nearest even); returns float64."""
    a32 = np.ascontiguousarray(a, np.float32)
    b = a32.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def param_count(d, H, dims):
    return H * d + H + H * H + H + dims * H + dims


def unpack(p, d, H, dims):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for shape in ((H, d), (H,), (H, H), (H,), (dims, H), (dims,)):
        size = int(np.prod(shape))
        out.append(p[o:o + size].reshape(shape))
        o += size
    return out  # W1, b1, W2, b2, W3, b3


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * _SQRT1_2))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z * _SQRT1_2)) + z * _INV_SQRT_2PI * np.exp(-0.5 * z * z)


def forward(params, x, d, H, dims, emulate=False):
    W1, b1, W2, b2, W3, b3 = unpack(params, d, H, dims)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    xb = r(x)
    w1, w2 = r(W1), r(W2)
    z1 = xb @ w1.T + b1
    h1 = r(gelu(z1).astype(np.float32)) if emulate else gelu(z1)
    z2 = h1 @ w2.T + b2
    h2 = gelu(z2)
    y = h2 @ W3.T + b3
    return y, dict(xb=xb, w1=w1, w2=w2, z1=z1, h1=h1, z2=z2, h2=h2)


def pair_terms(yi, yj, kind, same, w):
    """Per-pair loss and dL/dy_i (dL/dy_j is its negative); kind 0 NN, 1 MN, 2 FP."""
    u = yi - yj
    D = 1.0 + (u * u).sum(axis=1)
    loss = np.zeros_like(D)
    w1 = np.zeros_like(D)
    for k, (c, c2) in enumerate(((10.0, 20.0), (10000.0, 20000.0))):
        m = (kind == k) & ~same
        loss[m] = w[k] * D[m] / (c + D[m])
        w1[m] = w[k] * c2 / (c + D[m]) ** 2
    m = (kind == 2) & ~same
    loss[m] = w[2] / (1.0 + D[m])
    w1[m] = -w[2] * 2.0 / (1.0 + D[m]) ** 2
    return loss, w1[:, None] * u


def typed_pairs(pair_nn, pair_mn, pair_fp):
    lists = [np.asarray(p, np.int64).reshape(-1, 2) for p in (pair_nn, pair_mn, pair_fp)]
    pairs = np.concatenate(lists, axis=0)
    kind = np.concatenate([np.full(len(l), k) for k, l in enumerate(lists)])
    return pairs, kind


def loss_and_grad(params, x, pair_nn, pair_mn, pair_fp, w, d, H, dims, emulate=False):
    """Mean pair loss over all pairs and its gradient (flat, parameter layout)."""
    pairs, kind = typed_pairs(pair_nn, pair_mn, pair_fp)
    P = len(pairs)
    xi = np.asarray(x, np.float32)[pairs[:, 0]]
    xj = np.asarray(x, np.float32)[pairs[:, 1]]
    rows = np.concatenate([xi, xj], axis=0)
    y, c = forward(params, rows, d, H, dims, emulate)
    same = pairs[:, 0] == pairs[:, 1]
    loss, gi = pair_terms(y[:P], y[P:], kind, same, np.asarray(w, np.float64))
    dY = np.concatenate([gi, -gi], axis=0) / P
    W1, b1, W2, b2, W3, b3 = unpack(params, d, H, dims)
    r = bf16 if emulate else (lambda a: a)
    dW3 = dY.T @ c["h2"]
    db3 = dY.sum(axis=0)
    dz2 = dY @ W3 * gelu_grad(c["z2"])
    dz2 = r(dz2.astype(np.float32)) if emulate else dz2
    dW2 = dz2.T @ c["h1"]
    db2 = dz2.sum(axis=0)
    dz1 = dz2 @ c["w2"] * gelu_grad(c["z1"])
    dz1 = r(dz1.astype(np.float32)) if emulate else dz1
    dW1 = dz1.T @ c["xb"]
    db1 = dz1.sum(axis=0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])
    return loss.sum() / P, g


def transform(params, x, d, H, dims, emulate=False):
    return forward(params, x, d, H, dims, emulate)[0]
