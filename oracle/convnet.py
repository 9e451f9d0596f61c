"""CPU oracle of the early-stop network (test infrastructure: only tests,
smoke and the bench's CPU baseline use it).

A numpy restatement of the causal dilated convolution network that
include/ems_learn.h describes: the same parameter layout, bf16 rounding of
the weights before every use, tanh activations, residual middle layers,
masked mean-squared loss, and Adam. The reference has no such network
(its run loop, src/engine/run.rs:424-540, always steps to the end
criterion), so the oracle is pinned by mathematics instead: tests compare
its gradient with central finite differences in float64 and its
convolution with numpy's own 1-D convolution.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

TAPS = 9


def bf16(w: np.ndarray) -> np.ndarray:
    """Rounds f32 values to bf16 (nearest even) and widens them back."""
    u = np.ascontiguousarray(w, dtype=np.float32).view(np.uint32)
    r = ((u >> 16) & 1) + np.uint32(0x7FFF)
    return ((u + r) & np.uint32(0xFFFF0000)).view(np.float32)


class Shape:
    def __init__(self, cin: int, cout: int, channels: int, dilation: Sequence[int]):
        self.cin, self.cout, self.C = cin, cout, channels
        self.dil = [int(d) for d in dilation]
        self.L = len(self.dil)
        self.li = [cin if l == 0 else channels for l in range(self.L)]
        self.lo = [cout if l == self.L - 1 else channels for l in range(self.L)]
        self.woff, self.boff = [], []
        n = 0
        for l in range(self.L):
            self.woff.append(n)
            n += self.lo[l] * self.li[l] * TAPS
            self.boff.append(n)
            n += self.lo[l]
        self.count = n

    def unpack(self, p: np.ndarray) -> List[Tuple[np.ndarray, np.ndarray]]:
        out = []
        for l in range(self.L):
            w = p[self.woff[l]:self.boff[l]].reshape(self.lo[l], self.li[l], TAPS)
            b = p[self.boff[l]:self.boff[l] + self.lo[l]]
            out.append((w, b))
        return out

    def pack(self, parts) -> np.ndarray:
        return np.concatenate([np.concatenate([w.reshape(-1), b]) for w, b in parts])


def conv(x: np.ndarray, w: np.ndarray, b: np.ndarray, d: int) -> np.ndarray:
    """z[b,o,t] = bias[o] + sum_c sum_k w[o,c,k] x[b,c,t - d (K-1-k)], x = 0
    before t = 0."""
    B, Ci, T = x.shape
    pad = d * (TAPS - 1)
    xp = np.concatenate([np.zeros((B, Ci, pad), x.dtype), x], axis=2)
    z = np.broadcast_to(b[None, :, None], (B, w.shape[0], T)).astype(x.dtype).copy()
    for k in range(TAPS):
        z += np.einsum("oc,bct->bot", w[:, :, k], xp[:, :, k * d:k * d + T], optimize=True)
    return z


def conv_bwd(dz: np.ndarray, x: np.ndarray, w: np.ndarray, d: int):
    B, Co, T = dz.shape
    pad = d * (TAPS - 1)
    xp = np.concatenate([np.zeros((B, x.shape[1], pad), x.dtype), x], axis=2)
    dw = np.zeros_like(w)
    dxp = np.zeros_like(xp)
    for k in range(TAPS):
        dw[:, :, k] = np.einsum("bot,bct->oc", dz, xp[:, :, k * d:k * d + T], optimize=True)
        dxp[:, :, k * d:k * d + T] += np.einsum("oc,bot->bct", w[:, :, k], dz, optimize=True)
    return dxp[:, :, pad:], dw, dz.sum(axis=(0, 2))


def forward(shape: Shape, p: np.ndarray, x: np.ndarray, keep: bool = False, round_w=True):
    parts = shape.unpack(p)
    h = x
    cache = []
    for l, (w, b) in enumerate(parts):
        wq = bf16(w) if round_w else w
        z = conv(h, wq, b, shape.dil[l])
        if l == shape.L - 1:
            cache.append((h, wq, None))
            h = z
        else:
            a = np.tanh(z)
            cache.append((h, wq, a))
            h = a if l == 0 else h + a
    return (h, cache) if keep else h


def loss_and_grad(shape: Shape, p: np.ndarray, x: np.ndarray, y: np.ndarray, t0: int,
                  round_w=True, spectral=None):
    """Mean over samples, channels and t >= t0 of (out - y)^2, plus with
    spectral = (K, s0, w) the mean over samples and channels of
    w d^T K d (d the error over [s0, s0 + len(K))), and the gradient with
    respect to the parameter vector (the bf16 rounding is treated as the
    identity, as in straight-through training)."""
    out, cache = forward(shape, p, x, keep=True, round_w=round_w)
    diff = out[:, :, t0:] - y[:, :, t0:]
    n = diff.size
    loss = (diff * diff).sum() / n
    g = np.zeros_like(out)
    g[:, :, t0:] = 2.0 * diff / n
    if spectral is not None:
        K, s0, w = spectral
        K = np.asarray(K, np.float64)
        m = K.shape[0]
        d = out[:, :, s0:s0 + m] - y[:, :, s0:s0 + m]
        Kd = np.einsum("tu,bcu->bct", K, d)
        bc = d.shape[0] * d.shape[1]
        loss = loss + w * float((d * Kd).sum()) / bc
        g[:, :, s0:s0 + m] += 2.0 * w * Kd / bc
    grads = [None] * shape.L
    for l in range(shape.L - 1, -1, -1):
        h, wq, a = cache[l]
        dz = g if a is None else g * (1.0 - a * a)
        dx, dw, db = conv_bwd(dz, h, wq, shape.dil[l])
        grads[l] = (dw, db)
        if l == shape.L - 1:
            g = dx
        elif l > 0:
            g = g + dx
    return float(loss), shape.pack(grads)


class Adam:
    """beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected; optional norm
    clipping of the gradient."""

    def __init__(self, n: int, lr: float):
        self.lr = lr
        self.m = np.zeros(n, np.float32)
        self.v = np.zeros(n, np.float32)
        self.t = 0

    def step(self, p: np.ndarray, g: np.ndarray, clip: float = 0.0) -> np.ndarray:
        """One step; the gradient scaled to 2-norm `clip` when longer."""
        n = float(np.sqrt(np.sum(np.asarray(g, np.float64) ** 2)))
        if clip > 0.0 and n > clip:
            g = (g * np.float32(clip / n)).astype(np.float32)
        self.t += 1
        b1, b2 = np.float32(0.9), np.float32(0.999)
        c1 = np.float32(1.0 - 0.9 ** self.t)
        c2 = np.float32(1.0 - 0.999 ** self.t)
        self.m = b1 * self.m + (np.float32(1) - b1) * g
        self.v = b2 * self.v + (np.float32(1) - b2) * g * g
        return p - np.float32(self.lr) * (self.m / c1) / (np.sqrt(self.v / c2) + np.float32(1e-8))
