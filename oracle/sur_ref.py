"""TEST INFRASTRUCTURE ONLY: numpy oracle of the surrogate network and its
crops (the checker of phase-field-gh200_amd/csrc/sur_kernels.hip). Only
tests/ may import it; product code never does.

Network: X0 (8 channels: phi, S_other, cos theta, sin theta, 4 zeros) ->
conv3x3 32 + ReLU -> conv3x3 32 + ReLU -> conv3x3 32 + ReLU -> conv3x3 1 ->
dphi; prediction = phi crop + dphi. Zero padding at the crop edge.
Rounding points mirror the kernels: inputs, weights, hidden activations and
backward activation gradients are bfloat16 (round to nearest even); sums are
evaluated here in float64 (the kernels accumulate in float32, in another
order), so comparisons carry a tolerance. With bf16=False every value stays
float64 and the backward pass is checked against finite differences.
"""
from __future__ import annotations

import numpy as np

CROP = 96


def bf16_bits(x) -> np.ndarray:
    """binary32 -> bfloat16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x007FFFFF) != 0)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r)
    return r.astype(np.uint16)


def bf16_value(bits) -> np.ndarray:
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf(x, on=True) -> np.ndarray:
    """Round to bfloat16 (via binary32) and return float64 values."""
    if not on:
        return np.asarray(x, np.float64)
    return bf16_value(bf16_bits(np.asarray(x, np.float32))).astype(np.float64)


def _pad(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))


def conv(X, W, b=None):
    """sum_{tap, ci} X[p + off(tap)][ci] W[co][tap][ci] (+ b[co]); X [n,H,W,cin],
    W [cout, 9, cin]."""
    n, H, Wd, _ = X.shape
    Xp = _pad(X)
    out = np.zeros((n, H, Wd, W.shape[0]))
    for t in range(9):
        ky, kx = divmod(t, 3)
        out += Xp[:, ky:ky + H, kx:kx + Wd, :] @ W[:, t, :].T
    if b is not None:
        out += b
    return out


def conv_wgrad(dY, X):
    """dW[co][tap][ci] = sum_p dY[p][co] X[p + off(tap)][ci]."""
    n, H, Wd, cout = dY.shape
    Xp = _pad(X)
    dW = np.zeros((cout, 9, X.shape[3]))
    for t in range(9):
        ky, kx = divmod(t, 3)
        dW[:, t, :] = np.einsum("nhwo,nhwi->oi", dY, Xp[:, ky:ky + H, kx:kx + Wd, :])
    return dW


def conv_dgrad(dY, W):
    """dX[q][ci] = sum_{tap, co} dY[q - off(tap)][co] W[co][tap][ci]."""
    Wt = np.transpose(W[:, ::-1, :], (2, 1, 0))   # [ci][8 - tap][co]
    return conv(dY, Wt)


class Params:
    """Weights as float64 values (bf16-rounded for the MFMA layers when bf16)."""

    def __init__(self, W1, b1, W2, b2, W3, b3, W4, b4, bf16=True):
        r = (lambda a: bf(a, bf16))
        self.W1, self.W2, self.W3 = r(W1), r(W2), r(W3)
        self.b1, self.b2, self.b3 = (np.asarray(v, np.float64) for v in (b1, b2, b3))
        self.W4 = np.asarray(W4, np.float64).reshape(1, 9, 32)
        self.b4 = float(b4)
        self.bf16 = bf16


def forward(X0, p: Params):
    r = (lambda a: bf(a, p.bf16))
    X0 = r(X0)
    h1 = r(np.maximum(conv(X0, p.W1, p.b1), 0.0))
    h2 = r(np.maximum(conv(h1, p.W2, p.b2), 0.0))
    h3 = r(np.maximum(conv(h2, p.W3, p.b3), 0.0))
    dphi = conv(h3, p.W4)[..., 0] + p.b4
    return {"X0": X0, "h1": h1, "h2": h2, "h3": h3, "dphi": dphi}


def backward(fw, target, p: Params):
    """Gradients of L = sum (dphi - target)^2 / count."""
    r = (lambda a: bf(a, p.bf16))
    count = target.size
    g = 2.0 * (fw["dphi"] - target) / count
    gg = g[..., None]
    out = {"loss": float(np.sum((fw["dphi"] - target) ** 2) / count)}
    out["W4"] = conv_wgrad(gg, fw["h3"])[0]
    out["b4"] = float(g.sum())
    dh3 = r(conv_dgrad(gg, p.W4) * (fw["h3"] > 0))
    out["W3"] = conv_wgrad(dh3, fw["h2"])
    out["b3"] = dh3.sum(axis=(0, 1, 2))
    dh2 = r(conv_dgrad(dh3, p.W3) * (fw["h2"] > 0))
    out["W2"] = conv_wgrad(dh2, fw["h1"])
    out["b2"] = dh2.sum(axis=(0, 1, 2))
    dh1 = r(conv_dgrad(dh2, p.W2) * (fw["h1"] > 0))
    out["W1"] = conv_wgrad(dh1, fw["X0"])
    out["b1"] = dh1.sum(axis=(0, 1, 2))
    out["dh1"], out["dh2"], out["dh3"] = dh1, dh2, dh3
    return out


def crops(phi_tiles, cells, cls, S, L, P, sincosf):
    """Network input (float values of bf16), exact phi crops and origins of
    every cell: the crop centred on the rounded centroid of the stored field.
    `sincosf(theta)` returns the project's binary32 (sin, cos)."""
    CLASSES = [(144, 144, 64, 64), (176, 144, 32, 64), (144, 176, 64, 32),
               (160, 160, 32, 32), (224, 224, 32, 32), (286, 286, 1, 1)]
    N = len(cells)
    X = np.zeros((N, CROP, CROP, 8), np.float32)
    phic = np.zeros((N, CROP, CROP), np.float32)
    origin = np.zeros((N, 2), np.int64)
    Sg = S.reshape(L, P)
    for n in range(N):
        c = cells[n]
        cx, cy = float(c["gx0"]), float(c["gy0"])
        if c["V"] > 0.0:
            cx += float(c["Cx"]) / float(c["V"])
            cy += float(c["Cy"]) / float(c["V"])
        ox = (int(np.rint(cx)) - CROP // 2) % L
        oy = (int(np.rint(cy)) - CROP // 2) % L
        origin[n] = (ox, oy)
        wx0, wy0, tx0, ty0 = CLASSES[int(cls[n])]
        tile = phi_tiles[n].reshape(288, 288)
        wx = (ox + np.arange(CROP)) % L
        wy = (oy + np.arange(CROP)) % L
        a = (wx - int(c["gx0"])) % L
        b = (wy - int(c["gy0"])) % L
        inside = (b[:, None] < wy0) & (a[None, :] < wx0)
        v = np.where(inside, tile[np.clip(ty0 + b, 0, 287)[:, None],
                                  np.clip(tx0 + a, 0, 287)[None, :]], 0.0).astype(np.float32)
        phic[n] = v
        sq = np.minimum(v * v, np.float32(4.0)).astype(np.float32)
        q = np.rint(sq * np.float32(134217728.0)).astype(np.uint64)
        s = Sg[wy[:, None], wx[None, :]].astype(np.uint64)
        other = np.where(s >= q, (s - np.minimum(q, s)).astype(np.float32) *
                         np.float32(7.450580596923828125e-9), np.float32(0.0))
        sn, cs = sincosf(float(c["theta"]))
        X[n, :, :, 0] = bf16_value(bf16_bits(v))
        X[n, :, :, 1] = bf16_value(bf16_bits(other.astype(np.float32)))
        X[n, :, :, 2] = bf16_value(bf16_bits(np.float32(cs)))
        X[n, :, :, 3] = bf16_value(bf16_bits(np.float32(sn)))
    return X, phic, origin
