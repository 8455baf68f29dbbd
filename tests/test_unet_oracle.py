"""Pins the U-Net oracle (oracle/unet.py): its analytic gradient against
central finite differences of its own exact float64 loss, on a tiny volume,
and its building blocks against direct definitions."""
import numpy as np

from oracle import unet as U


def _params(seed):
    blocks, n = U.layout()
    rng = np.random.default_rng(seed)
    p = np.zeros(n)
    for i, (o, s) in enumerate(blocks):
        size = int(np.prod(s))
        if i < 10:
            fan = 27 * s[2]
            p[o:o + size] = rng.normal(0, np.sqrt(2.0 / fan), size)
        elif i < 20:
            p[o:o + size] = rng.normal(0, 0.05, size)
        else:
            p[o:o + size] = rng.normal(0, 0.3, size)
    return p


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(4, 4, 4, 3))
    W = rng.normal(size=(27, 2, 3))
    b = rng.normal(size=2)
    Y = U.conv(X, W, b)
    z, y, x = 1, 0, 3
    ref = b.copy()
    for tap in range(27):
        dz, dy, dx = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        zz, yy, xx = z + dz, y + dy, x + dx
        if 0 <= zz < 4 and 0 <= yy < 4 and 0 <= xx < 4:
            ref += W[tap] @ X[zz, yy, xx]
    assert np.allclose(Y[z, y, x], ref)


def test_pool_up_adjoint():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(4, 6, 8, 3))
    b = rng.normal(size=(2, 3, 4, 3))
    # <pool(a), b> = <a, up(b)/8>
    assert np.isclose((U.pool(a) * b).sum(), (a * U.up(b) / 8).sum())


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 3.1415926], np.float32)
    r = U.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie to even
    assert r[2] == 1.0 + 2 ** -7                 # rounds up
    assert abs(r[3] - 3.140625) < 1e-12


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = _params(4)
    obs = rng.uniform(0.5, 2.0, (4, 4, 8)).astype(np.float32)
    tgt = rng.uniform(0.5, 2.0, (4, 4, 8)).astype(np.float32)
    loss, g = U.loss_and_grads(p, obs, tgt, 1.7, 3.2, on=False)
    blocks, n = U.layout()
    picks = []
    for i, (o, s) in enumerate(blocks):
        size = int(np.prod(s))
        picks += list(o + rng.choice(size, min(3, size), replace=False))
    eps = 1e-5
    for k in picks:
        pp = p.copy()
        pp[k] += eps
        lp, _ = U.loss_and_grads(pp, obs, tgt, 1.7, 3.2, on=False)
        pp[k] -= 2 * eps
        lm, _ = U.loss_and_grads(pp, obs, tgt, 1.7, 3.2, on=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, g[k])
