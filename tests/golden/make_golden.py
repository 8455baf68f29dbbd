"""Writes tests/golden/*.npz: small fixtures whose expected outputs come from
a float64 numpy restatement that uses DIRECT circular convolution (no FFT),
independent of the C oracle and of the HIP library.

Inputs follow the reference's own test generators (tests/common/
Test_Common.h:99-161): a cube of ones over the central half of each axis
and an isotropic Gaussian kernel centred at dim/2, evaluated in double,
stored as float and normalised by a sequential float sum -- at sizes small
enough for a direct convolution.  The reference itself ships no golden
outputs (its tests print RMS only), so these are the fixtures.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def true_image(dims):
    """Test_GenerateTrueImage: ones on [d/4, d - d/4) per axis."""
    x, y, z = dims
    img = np.zeros((z, y, x), np.float32)
    img[z // 4:z - z // 4, y // 4:y - y // 4, x // 4:x - x // 4] = 1.0
    return img


def gaussian_kernel(dims, sigma):
    """Test_GenerateGaussianKernel: value in double, stored as float, float
    sequential sum, multiplied by float 1/sum."""
    x, y, z = dims
    s2 = np.float32(sigma) * np.float32(sigma)
    fz = (np.arange(z) - z // 2).astype(np.float32)
    fy = (np.arange(y) - y // 2).astype(np.float32)
    fx = (np.arange(x) - x // 2).astype(np.float32)
    r2 = (fx[None, None, :] * fx[None, None, :] + fy[None, :, None] * fy[None, :, None]
          ) + fz[:, None, None] * fz[:, None, None]
    e = np.exp((-(r2 / (np.float32(2) * s2))).astype(np.float32)).astype(np.float32)
    val = ((1.0 / (2.0 * np.pi * float(s2)) ** 1.5) * e.astype(np.float64)).astype(np.float32)
    total = np.cumsum(val.reshape(-1), dtype=np.float32)[-1]
    return (val * (np.float32(1.0) / total)).astype(np.float32)


def pad(src, wdims, shift):
    """Cyclic shift into the work domain, fill 0 (independent of the
    reference's loop form: roll a zero-padded copy)."""
    out = np.zeros(wdims, np.float64)
    sz, sy, sx = src.shape
    out[:sz, :sy, :sx] = src
    return np.roll(out, (shift[2], shift[1], shift[0]), axis=(0, 1, 2))


def circ_conv(a, b):
    """Direct circular convolution: out[d] = sum_s a[s] b[d - s]."""
    out = np.zeros_like(a)
    nz, ny, nx = a.shape
    for s in zip(*np.nonzero(a)):
        out += a[s] * np.roll(b, s, axis=(0, 1, 2))
    return out


def work_setup(img, ker):
    iz, iy, ix = img.shape
    kz, ky, kx = ker.shape
    w = (iz + kz, iy + ky, ix + kx)
    obs = pad(img, w, (0, 0, 0))
    h = pad(ker, w, (-(kx // 2), -(ky // 2), -(kz // 2)))
    return obs, h


def clip(a, shape):
    return a[:shape[0], :shape[1], :shape[2]]


def convolve(img, ker):
    obs, h = work_setup(img, ker)
    return clip(circ_conv(obs, h), img.shape)


def rl(img, ker, iters):
    obs, h = work_setup(img, ker)
    E = obs.sum()
    g = obs.copy()
    for _ in range(iters):
        s1 = circ_conv(g, h)
        r = np.where(np.abs(s1) < 1e-5, 0.0, obs / np.where(s1 == 0, 1, s1))
        s2 = circ_conv(r, h)
        g = g * s2
        g *= E / g.sum()
    return clip(g, img.shape)


def jvc(img, ker, iters):
    obs, h = work_setup(img, ker)
    A = 0.5 * float(img.max())
    inv = 1.0 / (A * A)
    i = obs.copy()
    for _ in range(max(iters, 1)):
        o = circ_conv(i, h)
        gamma = 1.0 - (o - A) ** 2 * inv
        i = np.maximum(i + gamma * (obs - o), 0.0)
    return clip(i, img.shape)


def wiener(img, ker, eps):
    obs, h = work_setup(img, ker)
    H = np.fft.rfftn(h)
    out = np.fft.irfftn(np.conj(H) / (np.abs(H) ** 2 + eps) * np.fft.rfftn(obs), s=obs.shape,
                         axes=(0, 1, 2))
    return clip(out, img.shape)


def circ_corr(r, o):
    """Direct circular correlation: out[z] = sum_y r[y] o[y - z]."""
    out = np.zeros_like(r)
    for s in zip(*np.nonzero(o)):
        out += o[s] * np.roll(r, tuple(-v for v in s), axis=(0, 1, 2))
    return out


def guard_ratio(obs, s):
    return np.where(np.abs(s) < 1e-5, 0.0, obs / np.where(s == 0, 1, s))


def blind_rl(obs_img, ker, outer, inner):
    """Alternating blind RL: per outer iteration `inner` RL steps of the
    image (kernel H in both convolutions, energy rescale), then `inner`
    kernel steps h' = h * corr(guard(obs / (h (*) o)), o), each renormalised
    to sum 1.  Returns the image and the kernel un-shifted into its dims."""
    obs, h = work_setup(obs_img, ker)
    h = h / h.sum()
    E = obs.sum()
    o = obs.copy()
    for _ in range(outer):
        for _ in range(inner):
            r = guard_ratio(obs, circ_conv(o, h))
            o = o * circ_conv(r, h)
            o *= E / o.sum()
        for _ in range(inner):
            r = guard_ratio(obs, circ_conv(h, o))
            h = h * circ_corr(r, o)
            h /= h.sum()
    kz, ky, kx = ker.shape
    hk = np.roll(h, (kz // 2, ky // 2, kx // 2), axis=(0, 1, 2))[:kz, :ky, :kx]
    return clip(o, obs_img.shape), hk


def main():
    cases = {
        # (image dims x,y,z), (kernel dims), sigma
        "ref_small": ((8, 8, 4), (4, 4, 4), 1.0),
        "ref_odd": ((7, 6, 5), (3, 5, 3), 0.9),
    }
    for name, (idims, kdims, sigma) in cases.items():
        img = true_image(idims)
        ker = gaussian_kernel(kdims, sigma)
        obs = convolve(img, ker).astype(np.float32)
        rng = np.random.default_rng(0)
        obs_noisy = (obs + rng.uniform(0.0, 0.05, obs.shape)).astype(np.float32)
        np.savez_compressed(
            os.path.join(HERE, f"{name}.npz"),
            image=img, kernel=ker, obs=obs_noisy,
            convolve=convolve(img, ker),
            rl1=rl(obs_noisy, ker, 1), rl5=rl(obs_noisy, ker, 5),
            jvc3=jvc(obs_noisy, ker, 3), wiener=wiener(obs_noisy, ker, 1e-3),
        )
        print("wrote", name)
    # alternating blind RL: a blurred box plus three off-centre beads (no
    # symmetry about the origin), noise added, kernel started 20 % too wide
    blind_cases = {
        "blind_a": ((10, 8, 6), (5, 5, 3), 1.0),
        "blind_b": ((9, 7, 8), (3, 5, 5), 0.8),
    }
    for name, (idims, kdims, sigma) in blind_cases.items():
        img = true_image(idims)
        img[1, 2, 3] += 4.0
        img[-2, 1, -3] += 2.5
        img[idims[2] // 2, -2, 1] += 3.0
        ker = gaussian_kernel(kdims, sigma)
        obs = convolve(img, ker)
        rng = np.random.default_rng(1)
        obs = (obs + rng.uniform(0.0, 0.05, obs.shape)).astype(np.float32)
        k0 = gaussian_kernel(kdims, 1.2 * sigma)
        o22, h22 = blind_rl(obs, k0, 2, 2)
        o31, h31 = blind_rl(obs, k0, 3, 1)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), obs=obs, kernel0=k0,
                            image_2x2=o22, kernel_2x2=h22, image_3x1=o31, kernel_3x1=h31)
        print("wrote", name)


if __name__ == "__main__":
    main()
