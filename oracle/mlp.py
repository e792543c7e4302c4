"""CPU oracle of the ray-cost MLP (16-64-64-1, ReLU).

TEST INFRASTRUCTURE ONLY: the checker of the HIP kernels in
pixels_amd/csrc/px_mlp.hip; the product never imports it.  The network has
no counterpart in the reference (its traversal is the plain DDA); this file
restates the build's own definition:

  h1 = bf16(relu(x @ bf16(W1) + b1))        x: [n, 16] (features, already BF16)
  h2 = bf16(relu(h1 @ bf16(W2) + b2))
  y  = h2 @ bf16(w3) + b3
  loss = mean((y - t)^2)

with the backward pass rounding to BF16 where the kernels stage operands for
MFMA: dy as a row operand, dH2 and dH1 after masking.  `grad64` is the same
network in float64 without any rounding; it is pinned by central finite
differences in the tests, and `grad_bf16` shares its formulas.

Parameter order (px_mlp_param_count() floats):
  W1[16][64], b1[64], W2[64][64], b2[64], w3[64], b3
"""
from __future__ import annotations

import numpy as np

N_IN, N_H = 16, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * N_H
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_H
N_PARAMS = OFF_B3 + 1


def bf16(a) -> np.ndarray:
    """Round FP32 to BF16 (nearest even), returned as FP32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).reshape(a.shape)


def bf16_bits_to_f32(h: np.ndarray) -> np.ndarray:
    return (np.ascontiguousarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


def unpack(p: np.ndarray):
    p = np.asarray(p)
    return (p[OFF_W1:OFF_B1].reshape(N_IN, N_H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(N_H, N_H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def pack(W1, b1, W2, b2, w3, b3) -> np.ndarray:
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]])


def forward_bf16(p, x):
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, np.float32))
    x = np.asarray(x, np.float32)
    h1 = bf16(np.maximum(x @ bf16(W1) + b1, 0.0))
    h2 = bf16(np.maximum(h1 @ bf16(W2) + b2, 0.0))
    y = h2 @ bf16(w3) + np.float32(b3)
    return y.astype(np.float32), h1, h2


def grad_bf16(p, x, t):
    """Loss and gradient with the kernels' BF16 rounding points."""
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, np.float32))
    x = np.asarray(x, np.float32)
    t = np.asarray(t, np.float32)
    n = x.shape[0]
    y, h1, h2 = forward_bf16(p, x)
    e = y - t
    loss = float(np.sum(e.astype(np.float64) ** 2) / n)
    dy = (2.0 * e / n).astype(np.float32)
    dyb = bf16(dy)
    gw3 = h2.T @ dyb
    gb3 = np.sum(dyb)
    dh2 = bf16(np.where(h2 > 0, bf16(w3)[None, :] * dy[:, None], 0.0))
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(0)
    dh1 = bf16(np.where(h1 > 0, dh2 @ bf16(W2).T, 0.0))
    gW1 = x.T @ dh1
    gb1 = dh1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gw3, gb3).astype(np.float32)


def loss64(p, x, t):
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, np.float64))
    h1 = np.maximum(x @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    y = h2 @ w3 + b3
    return float(np.mean((y - t) ** 2))


def grad64(p, x, t):
    """Analytic gradient in float64 without rounding (same formulas)."""
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, np.float64))
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    n = x.shape[0]
    h1 = np.maximum(x @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    y = h2 @ w3 + b3
    dy = 2.0 * (y - t) / n
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    dh2 = np.where(h2 > 0, w3[None, :] * dy[:, None], 0.0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(0)
    dh1 = np.where(h1 > 0, dh2 @ W2.T, 0.0)
    gW1 = x.T @ dh1
    gb1 = dh1.sum(0)
    return pack(gW1, gb1, gW2, gb2, gw3, gb3)


def adam(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected), FP32 as the kernel."""
    p = np.asarray(p, np.float32)
    g = np.asarray(g, np.float32)
    m = np.float32(b1) * m + np.float32(1 - b1) * g
    v = np.float32(b2) * v + np.float32(1 - b2) * g * g
    c1 = np.float32(1 - b1 ** step)
    c2 = np.float32(1 - b2 ** step)
    p = p - np.float32(lr) * (m / c1) / (np.sqrt(v / c2) + np.float32(eps))
    return p.astype(np.float32), m.astype(np.float32), v.astype(np.float32)
