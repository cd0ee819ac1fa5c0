"""CPU FP64 restatement of the warm-start MLP (forward, MSE loss, backward).

TEST INFRASTRUCTURE ONLY: the checker of libpf's MLP kernels.  With
``quantize=True`` it rounds to bfloat16 (round to nearest even) at the same
points as the GPU path -- normalised input, weights, hidden activations and
the backward signals fed to the matrix products -- and accumulates in FP64;
with ``quantize=False`` it is plain FP64 (pinned by finite differences).

Parameter layout (shared with pf_warmstart_get_params): for every layer,
W [K x N] row-major, then bias [N].  GELU is the exact form
0.5 z (1 + erf(z / sqrt 2)).
"""
from __future__ import annotations

import math

import numpy as np

_erf = np.vectorize(math.erf, otypes=[np.float64])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = ((u + (((u >> 16) & 1) + 0x7FFF)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def unpack(params, shapes):
    out, o = [], 0
    for K, N in shapes:
        W = params[o:o + K * N].reshape(K, N)
        o += K * N
        b = params[o:o + N]
        o += N
        out.append((W.astype(np.float64), b.astype(np.float64)))
    return out


def loss_and_grad(params, shapes, X, T, xm, xs, ym, ys, quantize=True):
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params, shapes)
    xis = np.where(xs > 0, 1.0 / np.where(xs > 0, xs, 1.0), 1.0).astype(np.float32)
    yis = np.where(ys > 0, 1.0 / np.where(ys > 0, ys, 1.0), 1.0).astype(np.float32)
    Xn = q(((X.astype(np.float32) - xm.astype(np.float32)) * xis).astype(np.float32))
    H = [Xn]
    Z = []
    for li, (W, b) in enumerate(layers):
        z = H[-1] @ q(W) + b
        Z.append(z)
        if li < len(layers) - 1:
            H.append(q(gelu(z)))
    Y = Z[-1]
    That = (T.astype(np.float32) - ym.astype(np.float32)) * yis
    d = Y - That.astype(np.float64)
    B, n = Y.shape
    loss = float(np.mean(d * d))
    dZ = 2.0 * d / (B * n)
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        dZq = q(dZ)
        gW = H[li].T @ dZq
        gb = dZ.sum(0)
        grads.append((gW, gb))
        if li > 0:
            dH = dZq @ q(W).T
            dZ = dH * gelu_grad(Z[li - 1])
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def predict(params, shapes, X, xm, xs, ym, ys, quantize=True):
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params, shapes)
    xis = np.where(xs > 0, 1.0 / np.where(xs > 0, xs, 1.0), 1.0).astype(np.float32)
    h = q(((X.astype(np.float32) - xm.astype(np.float32)) * xis).astype(np.float32))
    for li, (W, b) in enumerate(layers):
        z = h @ q(W) + b
        h = q(gelu(z)) if li < len(layers) - 1 else z
    ysd = np.where(ys > 0, ys, 1.0)
    return h * ysd + ym
