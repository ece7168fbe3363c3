"""CPU reference of the learned-guess network (test infrastructure).

The 54 -> 256 -> 256 -> 256 -> 4 ReLU MLP with the library's flat parameter
layout (per layer W[out][in] row-major, then b[out]) and mean-squared-error
loss over all outputs.  With ``bf16=True`` every GEMM operand is rounded to
BF16 (round to nearest even) as the matrix-core kernels do, and products are
accumulated in FP64; with ``bf16=False`` everything is FP64, which the
finite-difference test uses to pin the backward pass.
"""
from __future__ import annotations

import numpy as np

DIMS = [54, 256, 256, 256, 4]


def layout():
    off, out = 0, []
    for i in range(4):
        w = off
        off += DIMS[i + 1] * DIMS[i]
        b = off
        off += DIMS[i + 1]
        out.append((w, b, DIMS[i], DIMS[i + 1]))
    return out, off


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _op(a, on):
    return bf16(a) if on else np.asarray(a, dtype=np.float64)


def unpack(P):
    lay, _ = layout()
    return [(P[w:w + o * i].reshape(o, i), P[b:b + o]) for (w, b, i, o) in lay]


def forward(P, X, bf16_ops=True):
    """Activations [a1, a2, a3, y]; hidden layers ReLU, last linear.  With
    bf16_ops the activations are stored as float32 between layers."""
    acts = []
    h = np.asarray(X, dtype=np.float32 if bf16_ops else np.float64)
    for k, (W, b) in enumerate(unpack(P)):
        z = _op(h, bf16_ops) @ _op(W, bf16_ops).T + np.asarray(b, dtype=np.float64)
        if k < 3:
            z = np.maximum(z, 0.0)
        h = z.astype(np.float32) if bf16_ops else z
        acts.append(h)
    return acts


def loss_and_grad(P, X, T, bf16_ops=True):
    acts = forward(P, X, bf16_ops)
    Y = acts[-1].astype(np.float64)
    T = np.asarray(T, dtype=np.float64)
    cnt = Y.size
    loss = float(np.sum((Y - T) ** 2) / cnt)
    d = 2.0 * (Y - T) / cnt
    if bf16_ops:
        d = d.astype(np.float32)
    params = unpack(P)
    lay, total = layout()
    G = np.zeros(total)
    for k in range(3, -1, -1):
        W, _ = params[k]
        inp = X if k == 0 else acts[k - 1]
        w, b, i, o = lay[k]
        G[w:w + o * i] = (_op(d, bf16_ops).T @ _op(inp, bf16_ops)).ravel()
        G[b:b + o] = np.sum(np.asarray(d, dtype=np.float64), axis=0)
        if k > 0:
            dh = _op(d, bf16_ops) @ _op(W, bf16_ops)
            dh = np.where(np.asarray(acts[k - 1]) > 0, dh, 0.0)
            d = dh.astype(np.float32) if bf16_ops else dh
    return loss, G
