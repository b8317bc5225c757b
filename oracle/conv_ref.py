"""CPU oracle of the learned warm-start network -- TEST INFRASTRUCTURE ONLY.

6 layers of 3x3 convolutions with circular padding, 32 channels, ReLU between
layers, MSE loss on the 2 output channels; float64 arithmetic.  With
``bf16=True`` the tensors are rounded to bfloat16 at the same points as the
GPU kernels (inputs, weights, every stored activation and activation
gradient), so the GPU's FP32-accumulated results can be compared closely.

Layout as in include/rheocell.h: x [N,H,W,3], y [N,H,W,2], w [6,9,32,32]
(layer, tap = 3*(dy+1)+(dx+1), out, in), b [6,32].
"""
from __future__ import annotations

import numpy as np

L, C = 6, 32


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _shift(a, dy, dx):
    # out[n, y, x] = a[n, y + dy, x + dx] (circular)
    return np.roll(np.roll(a, -dy, axis=1), -dx, axis=2)


def conv(a, w):
    """a [N,H,W,Cin], w [9, Cout, Cin] -> [N,H,W,Cout]"""
    out = 0.0
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        out = out + _shift(a, dy, dx) @ w[t].T
    return out


def conv_t(g, w):
    """adjoint of conv w.r.t. its input: g [N,H,W,Cout] -> [N,H,W,Cin]"""
    out = 0.0
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        out = out + _shift(g @ w[t], -dy, -dx)
    return out


def conv_w(a, g):
    """gradient w.r.t. w: sum_p g[p, co] a[p + tap, ci] -> [9, Cout, Cin]"""
    dw = np.empty((9, g.shape[-1], a.shape[-1]))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        dw[t] = np.einsum("nhwo,nhwi->oi", g, _shift(a, dy, dx))
    return dw


def forward(x, w, b, bf=True):
    r = bf16 if bf else (lambda v: v)
    a = np.zeros(x.shape[:3] + (C,))
    a[..., :3] = x
    acts = [r(a)]
    wr = r(w)
    for l in range(L):
        z = conv(acts[-1], wr[l]) + b[l]
        if l < L - 1:
            acts.append(r(np.maximum(z, 0.0)))
        else:
            acts.append(r(z))
            out = z[..., :2]
    return acts, out


def loss_and_grads(x, y, w, b, bf=True):
    """MSE over the 2 output channels; gradients of w and b."""
    r = bf16 if bf else (lambda v: v)
    acts, out = forward(x, w, b, bf)
    wr = r(w)
    npix = x.shape[0] * x.shape[1] * x.shape[2]
    e = out - y
    loss = float((e ** 2).sum() / (2 * npix))
    g = np.zeros(out.shape[:3] + (C,))
    g[..., :2] = 2 * e / (2 * npix)
    g = r(g)
    gw = np.zeros_like(w)
    gb = np.zeros_like(b)
    for l in range(L - 1, -1, -1):
        gw[l] = conv_w(acts[l], g)
        gb[l] = g.sum(axis=(0, 1, 2))
        if l > 0:
            g = r(conv_t(g, wr[l]) * (acts[l] > 0))
    mask_w, mask_b = masks()
    return loss, gw * mask_w, gb * mask_b


def masks():
    mw = np.ones((L, 9, C, C))
    mw[0, :, :, 3:] = 0
    mw[L - 1, :, 2:, :] = 0
    mb = np.ones((L, C))
    mb[L - 1, 2:] = 0
    return mw, mb


def init_params(seed=0, scale=1.0):
    """He-style initialisation with the absent channels zeroed."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((L, 9, C, C)) * np.sqrt(2.0 / (9 * C)) * scale
    w[0] = rng.standard_normal((9, C, C)) * np.sqrt(2.0 / (9 * 3)) * scale
    b = np.zeros((L, C))
    mw, mb = masks()
    return (w * mw).astype(np.float32), (b * mb).astype(np.float32)
