"""ORACLE — test infrastructure, not product code.

FP64 restatement of the learned-k classifier (no counterpart in the
reference, whose neighbour count is the compile-time _K_, Config.sh /
knn.h:20).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg use it, and only as the checker of the HIP kernels.

Network: 16 features -> 64 (ReLU) -> 64 (ReLU) -> 6 logits over the k buckets
(16, 24, 32, 48, 64, 96).  Parameters are one flat float32 vector:
W1[16][64], b1[64], W2[64][64], b2[64], W3[64][6], b3[6] (z = x W + b).

The GPU computes every product on bfloat16 operands with float32
accumulation.  `forward(..., bf16=True)` rounds the same operands at the same
places (weights, features, hidden activations, output gradients), so the only
remaining difference is the accumulation order.  `bf16=False` is the plain
FP64 network, whose gradient is pinned by central finite differences.

Labels: the minimal certified k of a seed (kcert, 1 + index of the neighbour
whose clip first satisfied the security-radius test, cell.cpp:369-384)
mapped to the smallest bucket >= kcert; a seed beyond 96 (or uncertified)
gets the last bucket.
"""
from __future__ import annotations

import numpy as np

BUCKETS = (16, 24, 32, 48, 64, 96)
NF, NH, NC = 16, 64, 6
SHAPES = (("W1", (NF, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NC)),
          ("b3", (NC,)))
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    flat = np.asarray(flat)
    out, o = {}, 0
    for name, shp in SHAPES:
        m = int(np.prod(shp))
        out[name] = flat[o:o + m].reshape(shp)
        o += m
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16(a):
    """Round float32 values to bfloat16 (round to nearest even), as float64."""
    x = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = x.view(np.uint32).astype(np.uint64)
    nan = np.isnan(x)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = u.astype(np.uint32).view(np.float32).astype(np.float64)
    r[nan] = np.nan
    return r


def labels(kcert):
    """Class index of each seed's minimal certified k."""
    kc = np.asarray(kcert, dtype=np.int64)
    b = np.asarray(BUCKETS)
    cls = np.searchsorted(b, kc, side="left")
    cls = np.where((kc <= 0) | (cls >= NC), NC - 1, cls)
    return cls.astype(np.int32)


def _r(a, on):
    return bf16(a) if on else np.asarray(a, dtype=np.float64)


def forward(params, X, bf16_ops=True):
    """Logits and the intermediate values (all float64)."""
    P = unpack(params)
    x = _r(X, bf16_ops)
    W1, W2, W3 = (_r(P[n], bf16_ops) for n in ("W1", "W2", "W3"))
    b1, b2, b3 = (np.asarray(P[n], dtype=np.float64) for n in ("b1", "b2", "b3"))
    z1 = x @ W1 + b1
    h1 = _r(np.maximum(z1, 0.0).astype(np.float32) if bf16_ops else np.maximum(z1, 0.0), bf16_ops)
    z2 = h1 @ W2 + b2
    h2 = _r(np.maximum(z2, 0.0).astype(np.float32) if bf16_ops else np.maximum(z2, 0.0), bf16_ops)
    z3 = h2 @ W3 + b3
    return z3, (x, z1, h1, z2, h2, W1, W2, W3)


def softmax(z):
    z = z - z.max(axis=1, keepdims=True)
    e = np.exp(z)
    return e / e.sum(axis=1, keepdims=True)


def loss_grad(params, X, y, bf16_ops=True):
    """Mean cross-entropy over the batch and its gradient (flat, float64).

    With bf16_ops the backward products use bfloat16 operands exactly where
    the GPU kernel does: the output gradient g3 and the masked hidden
    gradients g2, g1 are rounded before they enter a product; bias gradients
    are sums of the unrounded values."""
    X = np.asarray(X)
    y = np.asarray(y)
    B = X.shape[0]
    z3, (x, z1, h1, z2, h2, W1, W2, W3) = forward(params, X, bf16_ops)
    p = softmax(z3)
    loss = float(-np.mean(np.log(p[np.arange(B), y])))
    g3 = p.copy()
    g3[np.arange(B), y] -= 1.0
    g3 /= B
    if bf16_ops:
        g3 = g3.astype(np.float32).astype(np.float64)
    g3r = _r(g3, bf16_ops)
    dW3 = h2.T @ g3r
    db3 = g3.sum(axis=0)
    g2 = (g3r @ W3.T) * (z2 > 0)
    if bf16_ops:
        g2 = g2.astype(np.float32).astype(np.float64)
    g2r = _r(g2, bf16_ops)
    dW2 = h1.T @ g2r
    db2 = g2.sum(axis=0)
    g1 = (g2r @ W2.T) * (z1 > 0)
    if bf16_ops:
        g1 = g1.astype(np.float32).astype(np.float64)
    g1r = _r(g1, bf16_ops)
    dW1 = x.T @ g1r
    db1 = g1.sum(axis=0)
    return loss, pack({"W1": dW1, "b1": db1, "W2": dW2, "b2": db2, "W3": dW3, "b3": db3})


# representative certified k of each bucket's interval (midpoints; 80 for the
# open last bucket), used by the cost model of the decision
KREP = (12.0, 20.0, 28.0, 40.0, 56.0, 80.0)


def chain_cost(k0, K, r):
    """Cost, in units of one KNN row entry, of a seed with certified k = K
    started at row length k0: each pass costs k (KNN row) + r * m (clip
    steps: m = K once certified, else the whole row), and a rejected pass
    repeats at twice the row length (capped at 256)."""
    k, cost = float(k0), 0.0
    while True:
        if K <= k:
            return cost + k + r * K
        cost += k + r * k
        k = min(2.0 * k, 256.0)


def decide(logits, r):
    """Bucket with the least expected chain cost under the softmax of the
    logits (r = clip cost per step over KNN cost per row entry)."""
    p = softmax(np.asarray(logits, dtype=np.float64))
    C = np.array([[chain_cost(BUCKETS[b], KREP[c], r) for c in range(NC)] for b in range(NC)])
    return np.argmin(p @ C.T, axis=1).astype(np.int32)


def adam(params, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (float64 restatement of the device update)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
