"""CPU oracle of the tile-cost predictor network (test infrastructure only:
the tests compare the device MFMA kernels against it; nothing in the
product imports it).

Network (parameter order of include/cpv.h, CPV_MLP_PARAMS floats):
    W1[12][64], b1[64], W2[64][64], b2[64], w3[64], b3
    z1 = x W1 + b1, h1 = relu(z1), z2 = h1 W2 + b2, h2 = relu(z2), y = h2 . w3 + b3
    loss = mean((y - t)^2)

With bf16=True every matrix-product operand is rounded to bfloat16 (round
to nearest even) exactly where the device kernels round them -- x, W1, W2,
w3, h1, h2 in the forward pass; dz2, dz1 (and the bf16 copies of W2, w3)
in the backward pass -- and every sum is accumulated in float64 and
rounded to float32, so the device's fp32 accumulation differs only by
summation order.  With bf16=False the network is plain float64 (used to
pin the backward pass against finite differences).

This network has no counterpart in the reference (it only schedules tile
work); it is pinned by finite differences, not by reference outputs.
"""
from __future__ import annotations

import numpy as np

F, H = 12, 64
O_W1 = 0
O_B1 = O_W1 + F * H
O_W2 = O_B1 + H
O_B2 = O_W2 + H * H
O_W3 = O_B2 + H
O_B3 = O_W3 + H
PARAMS = O_B3 + 1


def bf16(a) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), np.float32(np.nan), out).astype(np.float32)


def unpack(p: np.ndarray):
    p = np.asarray(p)
    return (p[O_W1:O_B1].reshape(F, H), p[O_B1:O_W2], p[O_W2:O_B2].reshape(H, H), p[O_B2:O_W3],
            p[O_W3:O_B3], p[O_B3])


def _mm(a, b, bf):
    """a @ b accumulated in float64, rounded to float32 in bf16 mode."""
    r = a.astype(np.float64) @ b.astype(np.float64)
    return r.astype(np.float32) if bf else r


def forward(p, x, bf16_mode: bool = True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    if bf16_mode:
        rnd = bf16
        xq, W1q, W2q, w3q = rnd(x), rnd(W1), rnd(W2), rnd(w3)
        z1 = (_mm(xq, W1q, True) + b1.astype(np.float32)).astype(np.float32)
        h1 = rnd(np.maximum(z1, 0))
        z2 = (_mm(h1, W2q, True) + b2.astype(np.float32)).astype(np.float32)
        h2 = rnd(np.maximum(z2, 0))
        y = (_mm(h2, w3q[:, None], True)[:, 0] + np.float32(b3)).astype(np.float32)
    else:
        x, W1, b1, W2, b2, w3, b3 = (np.asarray(v, np.float64) for v in (x, W1, b1, W2, b2, w3, b3))
        z1 = x @ W1 + b1
        h1 = np.maximum(z1, 0)
        z2 = h1 @ W2 + b2
        h2 = np.maximum(z2, 0)
        y = h2 @ w3 + b3
    return y, dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2)


def gradients(p, x, t, bf16_mode: bool = True):
    """Gradient of mean((y - t)^2) in the parameter order, and the loss."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, a = forward(p, x, bf16_mode)
    n = x.shape[0]
    if bf16_mode:
        f32 = np.float32
        d = (y - np.asarray(t, f32)).astype(f32)
        dy = (f32(2.0) * d * f32(1.0 / n)).astype(f32)
        w3q = bf16(w3)
        dz2 = np.where(a["z2"] > 0, w3q[None, :] * dy[:, None], f32(0)).astype(f32)
        dz2q = bf16(dz2)
        dh1 = _mm(dz2q, bf16(W2).T, True)
        dz1 = np.where(a["z1"] > 0, dh1, f32(0)).astype(f32)
        dz1q = bf16(dz1)
        g = np.zeros(PARAMS, np.float32)
        g[O_W1:O_B1] = _mm(bf16(x).T, dz1q, True).reshape(-1)
        g[O_B1:O_W2] = dz1.astype(np.float64).sum(0)
        g[O_W2:O_B2] = _mm(a["h1"].T, dz2q, True).reshape(-1)
        g[O_B2:O_W3] = dz2.astype(np.float64).sum(0)
        g[O_W3:O_B3] = (a["h2"].astype(np.float64) * dy[:, None].astype(np.float64)).sum(0)
        g[O_B3] = dy.astype(np.float64).sum()
        loss = float((d.astype(np.float64) ** 2).sum() / n)
        return g, loss
    x = np.asarray(x, np.float64)
    d = y - np.asarray(t, np.float64)
    dy = 2.0 * d / n
    dz2 = np.where(a["z2"] > 0, np.outer(dy, w3), 0.0)
    dz1 = np.where(a["z1"] > 0, dz2 @ np.asarray(W2, np.float64).T, 0.0)
    g = np.zeros(PARAMS, np.float64)
    g[O_W1:O_B1] = (x.T @ dz1).reshape(-1)
    g[O_B1:O_W2] = dz1.sum(0)
    g[O_W2:O_B2] = (a["h1"].T @ dz2).reshape(-1)
    g[O_B2:O_W3] = dz2.sum(0)
    g[O_W3:O_B3] = a["h2"].T @ dy
    g[O_B3] = dy.sum()
    return g, float((d ** 2).mean())


def init(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, np.float32)
    p[O_W1:O_B1] = rng.uniform(-1, 1, F * H) * np.sqrt(6.0 / F)
    p[O_W2:O_B2] = rng.uniform(-1, 1, H * H) * np.sqrt(6.0 / H)
    p[O_W3:O_B3] = rng.uniform(-1, 1, H) * np.sqrt(6.0 / H)
    p[O_B1:O_W2] = rng.uniform(-0.1, 0.1, H)
    p[O_B2:O_W3] = rng.uniform(-0.1, 0.1, H)
    p[O_B3] = 0.5
    return p
