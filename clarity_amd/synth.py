"""Seeded synthetic microscopy volumes for tests and benchmarks.

Objects are point beads on a constant background plus random-walk
filaments; kernels are (anisotropic) Gaussians centred at index n//2 per
axis (the reference's test convention, tests/common/Test_Common.h:139-143),
normalised to sum 1 in float64 and stored as float32.  Observations are
Poisson(convolve(object, kernel)).  All randomness comes from one numpy
PCG64 stream per seed, so every caller that asks for the same config and
seed gets identical bytes.

Volumes are float32 arrays of shape (z, y, x).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# 26-neighbourhood unit steps for the filaments
_STEPS = np.array([(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1)
                   for dx in (-1, 0, 1) if (dz, dy, dx) != (0, 0, 0)], np.int64)


def gaussian_kernel(shape, sigma_xy: float, sigma_z: float,
                    side_lobe: float = 0.0, lobe_dz: int = 8) -> np.ndarray:
    """Separable Gaussian centred at n//2, optional axial side lobes of
    relative amplitude ``side_lobe`` at z offsets +-lobe_dz; sum 1."""
    nz, ny, nx = shape
    z = np.arange(nz, dtype=np.float64) - nz // 2
    y = np.arange(ny, dtype=np.float64) - ny // 2
    x = np.arange(nx, dtype=np.float64) - nx // 2
    gz = np.exp(-z ** 2 / (2 * sigma_z ** 2))
    if side_lobe:
        gz = gz + side_lobe * (np.exp(-(z - lobe_dz) ** 2 / (2 * sigma_z ** 2))
                               + np.exp(-(z + lobe_dz) ** 2 / (2 * sigma_z ** 2)))
    gy = np.exp(-y ** 2 / (2 * sigma_xy ** 2))
    gx = np.exp(-x ** 2 / (2 * sigma_xy ** 2))
    k = gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    k /= k.sum()
    return k.astype(np.float32)


def beads_and_filaments(shape, n_beads: int, n_filaments: int,
                        rng: np.random.Generator, background: float = 10.0,
                        bead_lo: float = 100.0, bead_hi: float = 1000.0,
                        fil_len=(40, 200), fil_intensity: float = 200.0) -> np.ndarray:
    nz, ny, nx = shape
    obj = np.full(shape, background, np.float64)
    if n_beads:
        zs = rng.integers(0, nz, n_beads)
        ys = rng.integers(0, ny, n_beads)
        xs = rng.integers(0, nx, n_beads)
        vals = rng.uniform(bead_lo, bead_hi, n_beads)
        np.add.at(obj, (zs, ys, xs), vals)
    for _ in range(n_filaments):
        length = int(rng.integers(fil_len[0], fil_len[1] + 1))
        start = np.array([rng.integers(0, nz), rng.integers(0, ny), rng.integers(0, nx)])
        steps = _STEPS[rng.integers(0, len(_STEPS), length - 1)]
        path = np.vstack([start, start + np.cumsum(steps, axis=0)])
        ok = ((path[:, 0] >= 0) & (path[:, 0] < nz) & (path[:, 1] >= 0) &
              (path[:, 1] < ny) & (path[:, 2] >= 0) & (path[:, 2] < nx))
        p = path[ok]
        np.add.at(obj, (p[:, 0], p[:, 1], p[:, 2]), fil_intensity)
    return obj.astype(np.float32)


def np_convolve(img: np.ndarray, kernel: np.ndarray) -> np.ndarray:
    """Linear convolution with Clarity_Convolve's geometry (work = image +
    kernel, kernel centred by a -k//2 cyclic shift, low-corner clip),
    computed in float64 with numpy's FFT."""
    iz, iy, ix = img.shape
    kz, ky, kx = kernel.shape
    wz, wy, wx = iz + kz, iy + ky, ix + kx
    a = np.zeros((wz, wy, wx))
    a[:iz, :iy, :ix] = img
    b = np.zeros((wz, wy, wx))
    b[:kz, :ky, :kx] = kernel
    b = np.roll(b, (-(kz // 2), -(ky // 2), -(kx // 2)), axis=(0, 1, 2))
    r = np.fft.irfftn(np.fft.rfftn(a) * np.fft.rfftn(b), s=(wz, wy, wx), axes=(0, 1, 2))
    return r[:iz, :iy, :ix]


def poisson(lam: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    return rng.poisson(np.maximum(lam, 0.0)).astype(np.float32)


@dataclass
class Case:
    obj: np.ndarray
    psf: np.ndarray
    obs: np.ndarray


WORKLOADS = ("beads-64x64x32", "filaments-256x256x64", "filaments-512x512x128")


def make_workload(name: str, seed: int | None = None, convolve=np_convolve,
                  shape=None) -> Case:
    """Synthetic inputs of the benchmark workloads (default seeds 1, 2, 3).

    beads-64x64x32: 200 beads I~U[100,1000] on background 10, Gaussian PSF
        sigma_xy 1.5, sigma_z 3.0, same size as the image.
    filaments-256x256x64: 2000 beads + 300 random-walk filaments (length
        40..200, +200 per voxel), PSF sigma_xy 2.0, sigma_z 4.5 with 10 %
        side lobes at z = +-8.
    filaments-512x512x128: the same statistics scaled by volume (x8).
    ``shape`` overrides the volume (z, y, x) with statistics scaled by volume.
    """
    if name not in WORKLOADS:
        raise ValueError(f"unknown workload {name}")
    rng = np.random.default_rng(WORKLOADS.index(name) + 1 if seed is None else seed)
    if name == "beads-64x64x32":
        base, beads, fils = (32, 64, 64), 200, 0
        sxy, sz, lobe = 1.5, 3.0, 0.0
    else:
        base, beads, fils = (64, 256, 256), 2000, 300
        sxy, sz, lobe = 2.0, 4.5, 0.1
        if name == "filaments-512x512x128":
            shape = shape or (128, 512, 512)
    shape = tuple(shape) if shape else base
    f = np.prod(shape) / np.prod(base)
    obj = beads_and_filaments(shape, int(round(beads * f)), int(round(fils * f)), rng)
    psf = gaussian_kernel(shape, sxy, sz, side_lobe=lobe)
    obs = poisson(convolve(obj, psf), rng)
    return Case(obj, psf, obs)


def make_batch(convolve=np_convolve, count: int = 48):
    """The uneven-volume batch: ``count`` volumes (z, y, x) with x, y drawn
    from {64, 96, 128, 192, 256} and z from {32, 48, 64, 96} (shapes from
    seed 4, drawn before and independently of the contents), each with
    the 256x256x64 filament statistics scaled by its volume and its own
    Gaussian kernel of its size, sigma_xy ~ U[1.2, 2.5], sigma_z ~
    U[2.5, 5.0] (contents from seed 5).  Returns (observations, kernels)."""
    srng = np.random.default_rng(4)
    shapes = [(int(srng.choice([32, 48, 64, 96])), int(srng.choice([64, 96, 128, 192, 256])),
               int(srng.choice([64, 96, 128, 192, 256]))) for _ in range(count)]
    rng = np.random.default_rng(5)
    base = (64, 256, 256)
    imgs, ks = [], []
    for s in shapes:
        f = np.prod(s) / np.prod(base)
        o = beads_and_filaments(s, int(2000 * f) + 1, int(300 * f) + 1, rng)
        k = gaussian_kernel(s, rng.uniform(1.2, 2.5), rng.uniform(2.5, 5.0))
        imgs.append(poisson(convolve(o, k), rng))
        ks.append(k)
    return imgs, ks


def blind_initial_kernel(name: str, shape) -> np.ndarray:
    """Blind-ML starting kernel: the workload's Gaussian 20 % too wide."""
    if name == "beads-64x64x32":
        return gaussian_kernel(shape, 1.5 * 1.2, 3.0 * 1.2)
    return gaussian_kernel(shape, 2.0 * 1.2, 4.5 * 1.2, side_lobe=0.1)
