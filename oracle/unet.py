"""CPU restatement (numpy, float64) of the learned warm-start network.

TEST INFRASTRUCTURE ONLY: the checker of ClarityEx_UNet's forward pass and
gradients.  The network has no counterpart in the reference; this file
restates the library's definition (include/ClarityEx.h, the U-Net block):

  in = (obs/mean, sigma_xy/4, sigma_z/8, 0 x 13)        16 channels
  e1 = c1b(c1a(in))          16      level 0
  e2 = c2b(c2a(pool(e1)))    32      level 1
  b3 = c3b(c3a(pool(e2)))    64      level 2
  d2 = c4b(c4a(up(b3) | e2)) 32      level 1
  d1 = c5b(c5a(up(d2) | e1)) 16      level 0
  out = obs/mean + head(d1)

with every conv a 3x3x3 zero-padded convolution + bias + ReLU, pool a 2x2x2
mean, up a nearest-neighbour 2x upsampling, | a channel concatenation, and
loss = mean((out - target/mean)^2).  Parameters are one float32 vector: ten
conv weights [27][cout][cin] (tap = (dz+1)*9 + (dy+1)*3 + (dx+1)), ten
biases, the head (16 weights, 1 bias).

`bf16=True` rounds to bfloat16 exactly where the library stores bfloat16
(conv weights, activations, activation gradients), so differences to the
GPU are only float32-vs-float64 accumulation; `bf16=False` is the exact
float64 function, pinned by finite differences in tests/test_unet_oracle.py.
"""
from __future__ import annotations

import numpy as np

CONV = [(16, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32),
        (48, 16), (16, 16)]
LEVEL = [0, 0, 1, 1, 2, 2, 1, 1, 0, 0]


def layout():
    """(offset, shape) of every parameter block in the flat vector."""
    blocks, o = [], 0
    for ci, co in CONV:
        blocks.append((o, (27, co, ci)))
        o += 27 * co * ci
    for _, co in CONV:
        blocks.append((o, (co,)))
        o += co
    blocks.append((o, (17,)))
    o += 17
    return blocks, o


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _r(a, on):
    return bf16(a) if on else np.asarray(a, np.float64)


def unpack(p, on):
    blocks, _ = layout()
    p = np.asarray(p, np.float32 if on else np.float64)
    W = [_r(p[o:o + np.prod(s)].reshape(s), on) for o, s in blocks[:10]]
    B = [p[o:o + s[0]].astype(np.float64) for o, s in blocks[10:20]]
    H = p[blocks[20][0]:blocks[20][0] + 17].astype(np.float64)
    return W, B, H


def _taps():
    for tap in range(27):
        yield tap, tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1


def conv(X, W, b):
    """Y[p, co] = b[co] + sum_tap sum_ci X[p + off, ci] W[tap, co, ci]."""
    D, H, Wd, _ = X.shape
    Xp = np.pad(X, ((1, 1), (1, 1), (1, 1), (0, 0)))
    Y = np.zeros((D, H, Wd, W.shape[1])) + b
    for tap, dz, dy, dx in _taps():
        sl = Xp[1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + Wd]
        Y += sl @ W[tap].T
    return Y


def conv_bwd(G, X, W):
    """(dX, dW, db) of Y = conv(X, W) given dY = G."""
    D, H, Wd, _ = X.shape
    Xp = np.pad(X, ((1, 1), (1, 1), (1, 1), (0, 0)))
    Gp = np.pad(G, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dX = np.zeros(X.shape)
    dW = np.zeros(W.shape)
    for tap, dz, dy, dx in _taps():
        sl = Xp[1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + Wd]
        dW[tap] = G.reshape(-1, G.shape[-1]).T @ sl.reshape(-1, sl.shape[-1])
        gs = Gp[1 - dz:1 - dz + D, 1 - dy:1 - dy + H, 1 - dx:1 - dx + Wd]
        dX += gs @ W[tap]
    return dX, dW, G.reshape(-1, G.shape[-1]).sum(0)


def pool(X):
    D, H, W, C = X.shape
    return X.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(1, 3, 5))


def up(X):
    return X.repeat(2, 0).repeat(2, 1).repeat(2, 2)


def forward(p, obs, sxy, sz, on=True):
    """Returns (out, cache) with out = obs/mean + head(d1) in mean units."""
    W, B, Hd = unpack(p, on)
    obs = np.asarray(obs, np.float32)
    n = obs.size
    s = np.float32(obs.astype(np.float64).sum())
    im = np.float32(np.float32(n) / s)
    x0 = np.zeros(obs.shape + (16,))
    x0[..., 0] = np.float32(obs * im)
    x0[..., 1] = np.float32(np.float32(sxy) * np.float32(0.25))
    x0[..., 2] = np.float32(np.float32(sz) * np.float32(0.125))
    x0 = _r(x0, on)
    acts, ins = [], []

    def layer(i, X):
        ins.append(X)
        Y = _r(np.maximum(conv(X, W[i], B[i]), 0.0), on)
        acts.append(Y)
        return Y

    e1 = layer(1, layer(0, x0))
    p1 = _r(pool(e1), on)
    e2 = layer(3, layer(2, p1))
    p2 = _r(pool(e2), on)
    b3 = layer(5, layer(4, p2))
    c2 = np.concatenate([up(b3), e2], axis=-1)
    d2 = layer(7, layer(6, c2))
    c1 = np.concatenate([up(d2), e1], axis=-1)
    d1 = layer(9, layer(8, c1))
    out = np.float32(obs * im).astype(np.float64) + Hd[16] + d1 @ Hd[:16]
    return out, dict(ins=ins, acts=acts, im=im, Hd=Hd, W=W, n=n)


def loss_and_grads(p, obs, target, sxy, sz, on=True):
    """(loss, flat gradient) of mean((out - target/mean)^2)."""
    out, c = forward(p, obs, sxy, sz, on)
    W, acts, ins, Hd, n = c["W"], c["acts"], c["ins"], c["Hd"], c["n"]
    t = np.asarray(target, np.float32).astype(np.float64) * np.float64(c["im"])
    d = out - t
    loss = float((d * d).mean())
    g = 2.0 * d / n
    blocks, total = layout()
    grad = np.zeros(total)
    o = blocks[20][0]
    grad[o:o + 16] = (acts[9] * g[..., None]).reshape(-1, 16).sum(0)
    grad[o + 16] = g.sum()
    gout = [None] * 10
    gout[9] = _r(g[..., None] * Hd[:16], on)
    gin = [None] * 10
    for i in range(9, -1, -1):
        dP = gout[i] * (acts[i] > 0)
        dX, dW, db = conv_bwd(dP, ins[i], W[i])
        o, s = blocks[i]
        grad[o:o + np.prod(s)] = dW.reshape(-1)
        o, s = blocks[10 + i]
        grad[o:o + s[0]] = db
        gin[i] = _r(dX, on)
        if i == 9:
            gout[8] = gin[9]
        elif i == 8:                      # c1 = up(d2) | e1
            gu, gskip1 = gin[8][..., :32], gin[8][..., 32:]
            D, H, Wd, C = gu.shape
            gout[7] = _r(gu.reshape(D // 2, 2, H // 2, 2, Wd // 2, 2, C).sum(axis=(1, 3, 5)), on)
        elif i == 7:
            gout[6] = gin[7]
        elif i == 6:                      # c2 = up(b3) | e2
            gu, gskip2 = gin[6][..., :64], gin[6][..., 64:]
            D, H, Wd, C = gu.shape
            gout[5] = _r(gu.reshape(D // 2, 2, H // 2, 2, Wd // 2, 2, C).sum(axis=(1, 3, 5)), on)
        elif i == 5:
            gout[4] = gin[5]
        elif i == 4:                      # e2 -> pool2 and skip
            gout[3] = _r(gskip2 + up(gin[4]) / 8.0, on)
        elif i == 3:
            gout[2] = gin[3]
        elif i == 2:
            gout[1] = _r(gskip1 + up(gin[2]) / 8.0, on)
        elif i == 1:
            gout[0] = gin[1]
    return loss, grad
