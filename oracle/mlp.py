"""TEST INFRASTRUCTURE: float64 numpy restatement of the seed network
(6 -> 64 -> 64 -> 1, tanh) and its gradients, with optional emulation of
the BF16 rounding the GPU applies to MFMA operands (weights of layers 1-2,
the input features, layer-1 activations feeding layer 2, and layer-2
pre-activation gradients feeding the backward products).  tanh is exact here;
the GPU evaluates it from the hardware exponential (a few FP32 ulp).  Parameter order:
W1[64][6], b1[64], W2[64][64], b2[64], w3[64], b3."""
from __future__ import annotations

import numpy as np

NPARAM = 4673


def bf16(a):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[0:384].reshape(64, 6)
    b1 = p[384:448]
    W2 = p[448:4544].reshape(64, 64)
    b2 = p[4544:4608]
    w3 = p[4608:4672]
    b3 = p[4672]
    return W1, b1, W2, b2, w3, b3


def forward(p, x, emulate_bf16=True):
    """x: (n, 6).  Returns y (n,) and the activations."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    x = np.asarray(x, dtype=np.float64)
    h1 = np.tanh(r(x) @ r(W1).T + b1)
    h2 = np.tanh(r(h1) @ r(W2).T + b2)
    y = h2 @ w3 + b3
    return y, (x, h1, h2)


def grad(p, x, t, scale, emulate_bf16=True):
    """Gradient of scale * sum (y - t)^2 and that loss."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, (x, h1, h2) = forward(p, x, emulate_bf16)
    res = y - t
    dy = 2.0 * scale * res
    g = np.zeros(NPARAM)
    g[4672] = dy.sum()
    g[4608:4672] = h2.T @ dy
    dz2 = np.outer(dy, w3) * (1 - h2 ** 2)                 # (n, 64)
    g[4544:4608] = dz2.sum(0)
    g[448:4544] = (r(dz2).T @ r(h1)).ravel()
    dh1 = r(dz2) @ r(W2)
    dz1 = dh1 * (1 - h1 ** 2)
    g[384:448] = dz1.sum(0)
    g[0:384] = (dz1.T @ x).ravel()
    return g, scale * float((res ** 2).sum())
