"""RL's result does not depend on the FFT length once it is alias-free.

The reference transforms over n = i + k per axis
(src/MaximumLikelihoodDeconvolve.cxx:271-289).  The fused GPU engine
transforms over the smallest of its sizes N >= max(k, i + k/2) instead:
everything an RL iteration keeps is zero outside the image corner and s1, s2
are read only inside it, so the kernel (shifted by -k/2) cannot wrap into
the offsets that matter.  These tests compose RL from the oracle's own
primitives (pad-shift, R2C, modulate, divide, multiply, scale, clip, each a
restatement of the reference function it names) on chosen domains and check
that every domain meeting the bound gives the reference-domain result, while
one plane less on any axis does not (a random kernel, heavy at its edges, so
that any wrap-around shows)."""
import numpy as np
import pytest

import oracle


@pytest.fixture(autouse=True)
def _fp64_oracle():
    oracle.set_reduce_mode(oracle.REDUCE_FP64)
    oracle.set_fft_backend(oracle.FFT_OWN)
    yield


def rl_on_domain(obs, psf, iterations, wd):
    """reference Clarity_MaximumLikelihoodDeconvolveCPU
    (src/MaximumLikelihoodDeconvolve.cxx:84-153) on the work domain wd
    = (nz, ny, nx) instead of image + kernel."""
    nz, ny, nx = wd
    kz, ky, kx = psf.shape
    g = oracle.pad_shift(obs, wd, (0, 0, 0))
    obs_p = g.copy()
    h = oracle.pad_shift(psf, wd, (-(kx // 2), -(ky // 2), -(kz // 2)))
    otf = oracle.fft_r2c(h)
    energy = oracle.reduce_sum(obs_p)

    def conv(a):
        return oracle.fft_c2r(oracle.modulate(oracle.fft_r2c(a), otf, nx), nx)

    for _ in range(iterations):
        s1 = conv(g)
        r = oracle.divide(obs_p, s1).reshape(wd)
        s2 = conv(r)
        g = oracle.multiply(g, s2).reshape(wd)
        g = oracle.scale(g, np.float32(energy / oracle.reduce_sum(g))).reshape(wd)
    return oracle.clip(g, obs.shape)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def case(ishape, kshape, seed):
    rng = np.random.default_rng(seed)
    obs = rng.uniform(1.0, 20.0, ishape).astype(np.float32)
    psf = rng.uniform(0.2, 1.0, kshape).astype(np.float32)
    return obs, (psf / psf.sum()).astype(np.float32)


@pytest.mark.parametrize("ishape,kshape", [
    ((6, 10, 12), (5, 7, 9)),   # odd kernel
    ((7, 9, 8), (4, 6, 10)),    # even kernel, wider than the image along x
    ((4, 6, 6), (9, 9, 9)),     # kernel larger than the image
])
def test_alias_free_domains_match_reference(ishape, kshape):
    """Exact bound N >= i + k/2: the kernel may even overlap itself
    (N < k) as long as the offsets |d| < i stay clear; the GPU engine also
    keeps N >= k (its kernel planes then never overlap)."""
    obs, psf = case(ishape, kshape, 3)
    ref = oracle.ml(obs, psf, 4)                       # over i + k
    full = tuple(i + k for i, k in zip(ishape, kshape))
    assert rel_l2(rl_on_domain(obs, psf, 4, full), ref) < 1e-5   # composition
    exact = tuple(i + k // 2 for i, k in zip(ishape, kshape))
    engine = tuple(max(k, b) for k, b in zip(kshape, exact))
    for wd in (exact, engine, tuple(b + 3 for b in engine)):
        err = rel_l2(rl_on_domain(obs, psf, 4, wd), ref)
        assert err < 1e-5, (wd, err)
    # one plane less than the bound along any axis wraps the kernel into
    # the image corner
    for ax in range(3):
        wd = list(exact)
        wd[ax] -= 1
        err = rel_l2(rl_on_domain(obs, psf, 4, tuple(wd)), ref)
        assert err > 1e-3, (wd, err)
