"""CPU model of the device Fourier-feature MLP -- TEST INFRASTRUCTURE ONLY.

Checks the GPU forward / backward kernels (gf-only-sem_amd/csrc/surrogate.hip):
the same network evaluated in FP64 with every matrix-product operand rounded
to BF16 (round to nearest even) exactly where the device rounds it, so device
and model differ only by FP32-vs-FP64 accumulation.  The model's own gradient
formulas are pinned by central finite differences in plain FP64
(tests/test_surrogate.py).  The surrogate has no counterpart in the
reference; its oracle is this restatement of the mathematics.

Layout (as the device): per linear layer W [in][out] row-major then b [out];
features [sin(2 pi B x), cos(2 pi B x)] with x scaled to [-1, 1]; hidden
layers GELU(x W + b) with the exact (erf) GELU; loss = mean over rows x 18 of
((out - y) / scale)^2 where out is the scaled-output network value.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def bf16(a):
    """Round to BF16 (nearest even) and return as float64."""
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFFFFFF
    return (u & 0xFFFF0000).astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


class Model:
    def __init__(self, w, B, n_freq, width, depth, lo, hi, scale, n_out=18, emulate_bf16=True):
        self.w = np.asarray(w, np.float64)
        self.B = np.asarray(B, np.float64).reshape(n_freq, 4)
        self.n_freq, self.width, self.depth, self.n_out = n_freq, width, depth, n_out
        self.lo, self.hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
        self.scale = np.asarray(scale, np.float64)
        self.r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
        self.shapes = []
        iw = 2 * n_freq
        off = 0
        for l in range(depth + 1):
            ow = n_out if l == depth else width
            self.shapes.append((off, iw, ow))
            off += iw * ow + ow
            iw = ow
        self.n_params = off

    def W(self, l, w=None):
        w = self.w if w is None else w
        off, iw, ow = self.shapes[l]
        return w[off: off + iw * ow].reshape(iw, ow), w[off + iw * ow: off + iw * ow + ow]

    def features(self, x):
        xs = (2.0 * (np.asarray(x, np.float64) - self.lo) / (self.hi - self.lo) - 1.0).astype(np.float32)
        v = 2.0 * np.pi * (xs.astype(np.float64) @ self.B.T)
        return np.concatenate([np.sin(v), np.cos(v)], axis=1)

    def forward(self, x, w=None):
        h = self.features(x)
        cache = [h]
        for l in range(self.depth):
            Wl, bl = self.W(l, w)
            z = self.r(h) @ self.r(Wl) + bl
            h = gelu(z)
            cache += [z, h]
        Wl, bl = self.W(self.depth, w)
        out = self.r(h) @ self.r(Wl) + bl
        return out, cache

    def loss_grad(self, x, y, w=None):
        out, cache = self.forward(x, w)
        ys = np.asarray(y, np.float64) / self.scale
        n = out.size
        loss = float(np.mean((out - ys) ** 2))
        g = np.zeros(self.n_params)
        d = 2.0 * (out - ys) / n
        h = cache[-1]
        off, iw, ow = self.shapes[self.depth]
        g[off: off + iw * ow] = (self.r(h).T @ self.r(d)).ravel()
        g[off + iw * ow: off + iw * ow + ow] = d.sum(axis=0)
        Wl, _ = self.W(self.depth, w)
        dh = self.r(d) @ self.r(Wl).T
        for l in range(self.depth - 1, -1, -1):
            z = cache[1 + 2 * l]
            hin = cache[2 * l]
            dz = dh * gelu_d(z)
            off, iw, ow = self.shapes[l]
            g[off: off + iw * ow] = (self.r(hin).T @ self.r(dz)).ravel()
            g[off + iw * ow: off + iw * ow + ow] = dz.sum(axis=0)
            if l > 0:
                Wl, _ = self.W(l, w)
                dh = self.r(dz) @ self.r(Wl).T
        return loss, g

    def predict(self, x, w=None):
        return self.forward(x, w)[0] * self.scale
