"""FP64 CPU restatement of the warm-start U-Net layers.

TEST INFRASTRUCTURE ONLY (the checker of thesis_amd/csrc/unet.hip and
thesis_amd/warmstart.py); nothing in the product imports it.  There is no
reference counterpart: the learned warm start is an addition of this
project.  Each backward function here is itself pinned by centred finite
differences of the matching forward in tests/test_unet_oracle.py, and the
GPU kernels are compared with these functions on the same BF16-rounded
inputs.

Layout: channels-last arrays (N, D, H, W, C); 3x3x3 convolutions use zero
padding 1 and weights w[co, k] with k = tap*Ci + ci,
tap = 9*(dz+1) + 3*(dy+1) + (dx+1).
"""
from __future__ import annotations

import numpy as np

TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def bf16_round(a):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) >> 16 << 16, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _shift(x, dz, dy, dx):
    """y[n, z, y, x] = x[n, z+dz, y+dy, x+dx] with zeros outside."""
    out = np.zeros_like(x)
    N, D, H, W = x.shape[:4]
    zs = slice(max(dz, 0), D + min(dz, 0)); zd = slice(max(-dz, 0), D + min(-dz, 0))
    ys = slice(max(dy, 0), H + min(dy, 0)); yd = slice(max(-dy, 0), H + min(-dy, 0))
    xs = slice(max(dx, 0), W + min(dx, 0)); xd = slice(max(-dx, 0), W + min(-dx, 0))
    out[:, zd, yd, xd] = x[:, zs, ys, xs]
    return out


def conv3(x, w, b=None):
    """x (N,D,H,W,Ci), w (Co, 27*Ci) -> (N,D,H,W,Co)."""
    Ci = x.shape[-1]
    Co = w.shape[0]
    out = np.zeros(x.shape[:4] + (Co,))
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += _shift(x, dz, dy, dx) @ w[:, t * Ci:(t + 1) * Ci].T
    if b is not None:
        out += b
    return out


def conv3_bwd(x, w, gy):
    """Gradients of conv3 for upstream gy: (gx, gw, gb)."""
    Ci = x.shape[-1]
    gx = np.zeros_like(x, dtype=np.float64)
    gw = np.zeros(w.shape)
    for t, (dz, dy, dx) in enumerate(TAPS):
        xs = _shift(x, dz, dy, dx)
        gw[:, t * Ci:(t + 1) * Ci] = np.tensordot(gy, xs, axes=([0, 1, 2, 3], [0, 1, 2, 3]))
        gx += _shift(gy @ w[:, t * Ci:(t + 1) * Ci], -dz, -dy, -dx)
    gb = gy.sum(axis=(0, 1, 2, 3))
    return gx, gw, gb


def relu(x):
    return np.maximum(x, 0.0)


def pool(x):
    N, D, H, W, C = x.shape
    return x.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def pool_bwd(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_bwd(g):
    N, D, H, W, C = g.shape
    return g.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


def head(x, w, b):
    return x @ w + b


def features(u, C=8):
    """(N, n, n, n) field with filled shell -> (feat (N,n,n,n,C), base).

    Channels: transfinite blend of the faces, the shell value (zero
    inside), x, y, z in [0, 1], twice the distance to the nearest face,
    then zeros."""
    N, n = u.shape[0], u.shape[1]
    s = np.arange(n) / (n - 1)
    fx = (1 - s)[None, None, None, :] * u[:, :, :, :1] + s[None, None, None, :] * u[:, :, :, -1:]
    fy = (1 - s)[None, None, :, None] * u[:, :, :1, :] + s[None, None, :, None] * u[:, :, -1:, :]
    fz = (1 - s)[None, :, None, None] * u[:, :1, :, :] + s[None, :, None, None] * u[:, -1:, :, :]
    blend = (fx + fy + fz) / 3.0
    shell = np.ones((n, n, n), bool)
    shell[1:-1, 1:-1, 1:-1] = False
    feat = np.zeros(u.shape + (C,))
    feat[..., 0] = blend
    feat[..., 1] = np.where(shell, u, 0.0)
    # coordinates (float32 as the kernel computes them) and the distance to
    # the nearest face, scaled to [0, 1]
    sf = (np.arange(n, dtype=np.float32) * np.float32(1.0 / (n - 1))).astype(np.float64)
    cz, cy, cx = np.meshgrid(sf, sf, sf, indexing="ij")
    feat[..., 2] = cx
    feat[..., 3] = cy
    feat[..., 4] = cz
    d = np.minimum(np.minimum(np.minimum(cx, 1 - cx), np.minimum(cy, 1 - cy)),
                   np.minimum(cz, 1 - cz))
    feat[..., 5] = 2 * d
    base = np.where(shell, u, blend)
    return feat, base


def fourier_field(modes, n):
    """modes (6, M, 3) of (a, kx, ky) -> n^3 field with the faces filled."""
    u = np.zeros((n, n, n))
    s = np.arange(n) / (n - 1)

    def face(md):
        f = np.zeros((n, n))
        for a, kx, ky in md:
            f += a * np.outer(np.sin(np.pi * ky * s), np.sin(np.pi * kx * s))  # [t, s]
        return f

    # order of assignment matches the GPU kernel: x faces, then y, then z
    u[:, :, -1] = 0
    fz0, fz1 = face(modes[4]), face(modes[5])  # s = x, t = y
    u[0, :, :] = fz0
    u[-1, :, :] = fz1
    fy0, fy1 = face(modes[2]), face(modes[3])  # s = x, t = z
    u[:, 0, :] = fy0
    u[:, -1, :] = fy1
    fx0, fx1 = face(modes[0]), face(modes[1])  # s = y, t = z
    u[:, :, 0] = fx0
    u[:, :, -1] = fx1
    return u


def trilinear(c, M):
    """Interior of an M^3 field interpolated from the node-centred m^3 field c."""
    m = c.shape[0]
    g = np.arange(M) * (m - 1) / (M - 1)
    i0 = np.minimum(np.floor(g).astype(int), m - 2)
    f = g - i0
    out = np.zeros((M, M, M))
    for a in (0, 1):
        wz = (1 - f) if a == 0 else f
        for b in (0, 1):
            wy = (1 - f) if b == 0 else f
            for d in (0, 1):
                wx = (1 - f) if d == 0 else f
                out += (wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
                        * c[np.ix_(i0 + a, i0 + b, i0 + d)])
    return out


class UNetRef:
    """FP64 forward/backward of the network of thesis_amd/warmstart.py.

    With bf16=True every tensor the GPU stores in BF16 is rounded the same
    way (layer outputs, pooled tensors, the masked gradients fed to the
    backward convolutions); the remaining difference to the GPU is the FP32
    accumulation order."""

    LAYERS = [("e1a", 8, 16), ("e1b", 16, 16), ("e2a", 16, 32), ("e2b", 32, 32),
              ("bta", 32, 64), ("btb", 64, 64), ("d2a", 96, 32), ("d2b", 32, 32),
              ("d1a", 48, 16), ("d1b", 16, 16)]

    def __init__(self, params, bf16=False):
        self.p = params  # dict name -> (w, b); 'head' -> (w (16,), b (1,))
        self.r = bf16_round if bf16 else (lambda a: a)

    def forward(self, x):
        P, c, r = self.p, {}, self.r
        a = x
        for name in ("e1a", "e1b"):
            a = r(relu(conv3(a, *P[name]))); c[name] = a
        s1 = a
        a = r(pool(s1)); c["p1"] = a
        for name in ("e2a", "e2b"):
            a = r(relu(conv3(a, *P[name]))); c[name] = a
        s2 = a
        a = r(pool(s2)); c["p2"] = a
        for name in ("bta", "btb"):
            a = r(relu(conv3(a, *P[name]))); c[name] = a
        a = np.concatenate([up(a), s2], axis=-1); c["cat2"] = a
        for name in ("d2a", "d2b"):
            a = r(relu(conv3(a, *P[name]))); c[name] = a
        a = np.concatenate([up(a), s1], axis=-1); c["cat1"] = a
        for name in ("d1a", "d1b"):
            a = r(relu(conv3(a, *P[name]))); c[name] = a
        y = head(a, *P["head"])
        self.cache = (x, c)
        return y

    def backward(self, gy):
        x, c = self.cache
        P, G = self.p, {}
        hw, hb = P["head"]
        G["head"] = (np.tensordot(gy, c["d1b"], axes=([0, 1, 2, 3], [0, 1, 2, 3])), np.array([gy.sum()]))
        g = gy[..., None] * hw

        def layer(name, xin, g):
            g = self.r(g * (c[name] > 0))
            gx, gw, gb = conv3_bwd(xin, P[name][0], g)
            G[name] = (gw, gb)
            return gx

        g = layer("d1b", c["d1a"], g)
        g = layer("d1a", c["cat1"], g)
        g_up1, g_s1 = g[..., :32], g[..., 32:]
        g = up_bwd(g_up1)
        g = layer("d2b", c["d2a"], g)
        g = layer("d2a", c["cat2"], g)
        g_up2, g_s2 = g[..., :64], g[..., 64:]
        g = up_bwd(g_up2)
        g = layer("btb", c["bta"], g)
        g = layer("bta", c["p2"], g)
        g = pool_bwd(g) + g_s2
        g = layer("e2b", c["e2a"], g)
        g = layer("e2a", c["p1"], g)
        g = pool_bwd(g) + g_s1
        g = layer("e1b", c["e1a"], g)
        layer("e1a", x, g)
        return G
