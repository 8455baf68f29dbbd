"""Parity of the HIP library (called through its C ABI) with the CPU oracle.

Tolerances (relative L2, ||gpu - ref|| / ||ref||), following the stated
targets for the float32 path:
  FFT vs numpy float64                     <= 2e-6
  Convolve / Wiener vs oracle              <= 1e-6 (measured 1e-7..8.5e-7)
  RL after 1 iteration                     <= 1e-5
  RL / JVC / blind after <= 100 iterations <= 1e-4
Bit-exact: pad / clip.  The oracle runs in its fp64 reduction mode (the
library's); the float32-serial mode's deviation is reported, not asserted.
"""
import os

import numpy as np
import pytest

import oracle
from clarity_amd import synth

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    t = np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64
    a = np.asarray(a, t)
    b = np.asarray(b, t)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def small():
    return synth.make_workload("beads-64x64x32")


@pytest.fixture(autouse=True)
def _fp64_oracle():
    oracle.set_reduce_mode(oracle.REDUCE_FP64)
    oracle.set_fft_backend(oracle.FFT_OWN)
    yield


def _rand(shape, seed, lo=0.0, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, shape).astype(np.float32)


# --------------------------------------------------------------------- FFT

FFT_SHAPES = [
    (8, 8, 8), (4, 6, 10), (7, 9, 15), (64, 128, 128), (64, 160, 160),
    (3, 5, 7), (1, 1, 16), (2, 3, 1), (16, 12, 96), (32, 64, 1024),
    (11, 13, 17), (5, 49, 12), (128, 64, 512),
]


@pytest.mark.parametrize("shape", FFT_SHAPES)
def test_fft_r2c_c2r_vs_numpy(clarity_gpu, shape):
    api = clarity_gpu
    print(f"[fft] shape={shape}", flush=True)
    a = _rand(shape, 1, -1, 1)
    F = api.FFT_R2C(a)
    R = np.fft.rfftn(a.astype(np.float64))
    assert F.shape == R.shape
    assert rel_l2(F, R) <= 2e-6
    G = (np.random.default_rng(2).standard_normal(F.shape)
         + 1j * np.random.default_rng(3).standard_normal(F.shape)).astype(np.complex64)
    g = api.FFT_C2R(G, shape[2])
    ref = np.fft.irfftn(G.astype(np.complex128), s=shape, axes=(0, 1, 2)) * a.size
    assert rel_l2(g, ref) <= 2e-6
    # round trip and input preserved
    G0 = G.copy()
    api.FFT_C2R(G, shape[2])
    assert np.array_equal(G, G0)
    back = api.FFT_C2R(F, shape[2]) / a.size
    assert rel_l2(back, a) <= 2e-6


def test_fft_closed_forms(clarity_gpu):
    api = clarity_gpu
    d = np.zeros((8, 16, 32), np.float32)
    d[0, 0, 0] = 1.0
    F = api.FFT_R2C(d)
    assert np.allclose(F, 1.0, atol=1e-6)           # delta -> all ones
    c = np.full((8, 16, 32), 2.0, np.float32)
    F = api.FFT_R2C(c)
    assert abs(F[0, 0, 0] - 2.0 * c.size) < 1e-3   # constant -> DC only
    F[0, 0, 0] = 0
    assert np.abs(F).max() < 1e-3
    x = np.arange(32)
    cos = np.broadcast_to(np.cos(2 * np.pi * 3 * x / 32), (8, 16, 32)).astype(np.float32)
    F = api.FFT_R2C(cos)
    assert abs(F[0, 0, 3] - cos.size / 2) < 1e-2     # one bin in the half spectrum
    F[0, 0, 3] = 0
    assert np.abs(F).max() < 1e-2


# Axis lengths beyond the LDS passes (FFTW plans every length,
# src/FFT.cxx:46-146): primes and prime factors above 64 (Bluestein with a
# direct or four-step inner transform), lengths above 4096 (four-step when
# smooth, Bluestein otherwise), on the x (real-to-complex), y and z axes.
LARGE_SHAPES = [
    (3, 5, 67), (3, 67, 4), (67, 4, 6),          # prime 67
    (2, 97, 97), (191, 3, 5), (2, 3, 573),       # 97, 191, 573 = 3 * 191
    (4, 573, 7), (573, 2, 3),
    (2, 3, 4099), (2, 4099, 3), (4099, 2, 3),    # prime above 4096: M = 16384
    (2, 3, 6000), (6000, 2, 3),                   # smooth above 4096: four-step
    (3, 2, 8209),                                 # prime, inner M = 32768
]


@pytest.mark.parametrize("shape", LARGE_SHAPES)
def test_fft_large_lengths_vs_numpy(clarity_gpu, shape):
    api = clarity_gpu
    print(f"[fft large] shape={shape}", flush=True)
    a = _rand(shape, 21, -1, 1)
    F = api.FFT_R2C(a)
    R = np.fft.rfftn(a.astype(np.float64))
    assert F.shape == R.shape
    e_f = rel_l2(F, R)
    G = (np.random.default_rng(22).standard_normal(F.shape)
         + 1j * np.random.default_rng(23).standard_normal(F.shape)).astype(np.complex64)
    g = api.FFT_C2R(G, shape[2])
    ref = np.fft.irfftn(G.astype(np.complex128), s=shape, axes=(0, 1, 2)) * a.size
    e_b = rel_l2(g, ref)
    back = api.FFT_C2R(F, shape[2]) / a.size
    e_r = rel_l2(back, a)
    print(f"[fft large] {shape}: r2c {e_f:.2e} c2r {e_b:.2e} round trip {e_r:.2e}")
    assert e_f <= 2e-6 and e_b <= 2e-6 and e_r <= 2e-6


def test_drivers_on_a_573_axis(clarity_gpu):
    """A 512-wide image with a 61-wide kernel: the reference's work domain
    has a 573 = 3 * 191 axis.  Wiener and JVC keep that domain (generic
    passes with a Bluestein x axis), Convolve and RL may choose their own
    alias-free domain; all against the oracle.  Wiener's bar here is the
    stated 1e-5: its filter (gain up to 1/(2 sqrt(eps)) ~ 16) amplifies the
    transforms' float32 error, measured 1.2e-6 on this domain."""
    api = clarity_gpu
    rng = np.random.default_rng(31)
    obj = synth.beads_and_filaments((4, 6, 512), 40, 2, rng, fil_len=(10, 40))
    psf = synth.gaussian_kernel((3, 5, 61), 1.5, 1.0)
    obs = synth.poisson(synth.np_convolve(obj, psf), rng)
    for name, fn, ref, tol in (
            ("convolve", lambda: api.Convolve(obj, psf), oracle.convolve(obj, psf), 1e-6),
            ("wiener", lambda: api.WienerDeconvolve(obs, psf, 1e-3),
             oracle.wiener(obs, psf, 1e-3), 1e-5),
            ("jvc", lambda: api.JansenVanCittertDeconvolve(obs, psf, 5),
             oracle.jvc(obs, psf, 5), 1e-4),
            ("rl", lambda: api.MaximumLikelihoodDeconvolve(obs, psf, 5),
             oracle.ml(obs, psf, 5), 1e-4)):
        for engine in ("fused", "generic"):
            if engine == "generic":
                os.environ["CLARITY_FORCE_GENERIC"] = "1"
            try:
                got = fn()
            finally:
                os.environ.pop("CLARITY_FORCE_GENERIC", None)
            err = rel_l2(got, ref)
            print(f"[573] {name} {engine} ({api.LastEngine()}): rel L2 {err:.2e}", flush=True)
            assert err <= tol, (name, engine)


# --------------------------------------------------------------- pad / clip

@pytest.mark.parametrize("src,dst,shift", [
    ((3, 4, 5), (6, 8, 9), (0, 0, 0)),
    ((3, 4, 5), (6, 8, 9), (-2, -2, -1)),
    ((5, 5, 5), (7, 6, 9), (3, -4, 2)),
    ((4, 3, 2), (4, 3, 2), (-11, 7, 30)),
    ((1, 1, 1), (2, 3, 4), (1, 1, 1)),
])
def test_pad_shift_bit_exact(clarity_gpu, src, dst, shift):
    api = clarity_gpu
    s = _rand(src, 4)
    got = api.ImagePadSpatialShift(s, dst, shift, fill=-3.5)
    ref = oracle.pad_shift(s, dst, shift, fill=-3.5)
    assert np.array_equal(got, ref)


def test_clip_bit_exact(clarity_gpu):
    api = clarity_gpu
    s = _rand((9, 10, 11), 5)
    assert np.array_equal(api.ImageClip(s, (4, 7, 3)), oracle.clip(s, (4, 7, 3)))
    assert np.array_equal(api.ImageClip(s, (9, 10, 11)), s)


# ---------------------------------------------------------------- drivers

def _engine(engine, fn):
    """fn() on the fused passes (the default where they apply) or, with
    engine == "generic", on the generic driver (CLARITY_FORCE_GENERIC); the
    engine the library reports for the call must be the one asked for."""
    from clarity_amd import api
    if engine == "generic":
        os.environ["CLARITY_FORCE_GENERIC"] = "1"
    try:
        out = fn()
    finally:
        os.environ.pop("CLARITY_FORCE_GENERIC", None)
    assert api.LastEngine() == ("generic" if engine == "generic" else "fast")
    return out


@pytest.mark.parametrize("engine", ["fused", "generic"])
def test_convolve_vs_oracle_small(clarity_gpu, small, engine):
    """Convolve on the fused passes (RL's first convolution, over its
    alias-free domain) and on the generic driver."""
    api = clarity_gpu
    got = _engine(engine, lambda: api.Convolve(small.obj, small.psf))
    ref = oracle.convolve(small.obj, small.psf)
    assert rel_l2(got, ref) <= 1e-6
    # the oracle itself against numpy float64 (independent FFT)
    assert rel_l2(ref, synth.np_convolve(small.obj, small.psf)) <= 1e-6


@pytest.mark.parametrize("ishape,kshape", [
    ((32, 128, 128), (32, 32, 32)),   # the reference tests' geometry, 1/... scale
    ((9, 10, 11), (3, 5, 4)),
    ((16, 16, 16), (1, 1, 1)),
])
def test_convolve_shapes(clarity_gpu, ishape, kshape):
    api = clarity_gpu
    a = _rand(ishape, 6)
    k = _rand(kshape, 7)
    got = api.Convolve(a, k)
    ref = synth.np_convolve(a, k)
    assert rel_l2(got, ref) <= 1e-6


def test_convolve_delta_identity(clarity_gpu):
    api = clarity_gpu
    a = _rand((12, 14, 16), 8)
    k = np.zeros((5, 6, 7), np.float32)
    k[5 // 2, 6 // 2, 7 // 2] = 1.0
    assert rel_l2(api.Convolve(a, k), a) <= 1e-6


@pytest.mark.parametrize("engine", ["fused", "generic"])
@pytest.mark.parametrize("name", ["beads64", "96x64x48"])
def test_wiener_vs_oracle(clarity_gpu, small, name, engine):
    """Wiener on the fused passes (the filter in z_mod, over the reference
    domain: 128 x 128 x 64 and, radix 3, 192 x 128 x 96) and on the generic
    driver."""
    api = clarity_gpu
    if name == "beads64":
        obs, psf = small.obs, small.psf
    else:
        case = synth.make_workload("beads-64x64x32", seed=6, shape=(48, 64, 96))
        obs, psf = case.obs, case.psf
    for eps in (1e-3, 1e-1):
        got = _engine(engine, lambda: api.WienerDeconvolve(obs, psf, eps))
        ref = oracle.wiener(obs, psf, eps)
        err = rel_l2(got, ref)
        print(f"[wiener] {name} {engine} eps {eps}: rel L2 {err:.2e}", flush=True)
        assert err <= 1e-6, eps


def test_wiener_delta_kernel(clarity_gpu):
    api = clarity_gpu
    a = _rand((8, 10, 12), 9)
    k = np.zeros((4, 4, 4), np.float32)
    k[2, 2, 2] = 1.0
    eps = 0.25
    assert rel_l2(api.WienerDeconvolve(a, k, eps), a / (1 + eps)) <= 1e-6


@pytest.mark.parametrize("ishape,kshape", [
    ((40, 70, 150), (24, 58, 42)),   # reference domain 64 x 128 x 192
    ((33, 61, 71), (31, 67, 57)),    # 64 x 128 x 128 from odd dims
    ((20, 100, 37), (44, 28, 91)),   # kernel wider than the image in x, deeper in z
])
def test_fused_single_and_jvc_odd_shapes(clarity_gpu, ishape, kshape):
    """Wiener and JVC on the fused passes over reference domains of odd,
    unequal image and kernel shapes (general z kernel, bounds-tested x
    rows), and Convolve over its alias-free domain, against the oracle and
    the generic driver."""
    api = clarity_gpu
    rng = np.random.default_rng(61)
    obj = synth.beads_and_filaments(ishape, 60, 2, rng)
    psf = synth.gaussian_kernel(kshape, 1.6, 3.1)
    obs = synth.poisson(synth.np_convolve(obj, psf), rng)
    for name, fn, ref, tol in (
            ("convolve", lambda: api.Convolve(obj, psf), oracle.convolve(obj, psf), 1e-6),
            ("wiener", lambda: api.WienerDeconvolve(obs, psf, 1e-2),
             oracle.wiener(obs, psf, 1e-2), 1e-6),
            ("jvc", lambda: api.JansenVanCittertDeconvolve(obs, psf, 4),
             oracle.jvc(obs, psf, 4), 1e-4)):
        got = _engine("fused", fn)
        gen = _engine("generic", fn)
        print(f"[odd] {ishape} {kshape} {name}: fused {rel_l2(got, ref):.2e} "
              f"generic {rel_l2(gen, ref):.2e}", flush=True)
        assert rel_l2(got, ref) <= tol, name
        assert rel_l2(gen, ref) <= tol, name


@pytest.mark.parametrize("iters", [1, 10])
def test_ml_vs_oracle_small(clarity_gpu, small, iters):
    api = clarity_gpu
    print(f"[ml] small iterations={iters}", flush=True)
    got = api.MaximumLikelihoodDeconvolve(small.obs, small.psf, iters)
    ref = oracle.ml(small.obs, small.psf, iters)
    err = rel_l2(got, ref)
    with oracle.reduce_mode(oracle.REDUCE_SERIAL_F32):
        ref32 = oracle.ml(small.obs, small.psf, iters)
    print(f"[ml] rel_l2 vs fp64-oracle {err:.3e}; vs serial-f32 {rel_l2(got, ref32):.3e}")
    assert err <= (1e-5 if iters == 1 else 1e-4)


def test_ml_energy_and_support(clarity_gpu, small):
    api = clarity_gpu
    out = api.MaximumLikelihoodDeconvolve(small.obs, small.psf, 5)
    assert np.all(out >= 0)
    assert abs(out.astype(np.float64).sum() / small.obs.astype(np.float64).sum() - 1) < 1e-5


def test_ml_scale_invariance(clarity_gpu, small):
    api = clarity_gpu
    a = api.MaximumLikelihoodDeconvolve(small.obs, small.psf, 3)
    b = api.MaximumLikelihoodDeconvolve(small.obs * 4.0, small.psf, 3)
    assert rel_l2(b, 4.0 * a) <= 1e-5


def test_ml_zero_iterations_and_errors(clarity_gpu, small):
    api = clarity_gpu
    out = api.MaximumLikelihoodDeconvolve(small.obs, small.psf, 0)
    assert np.array_equal(out, small.obs)
    with pytest.raises(api.ClarityError) as e:
        api.MaximumLikelihoodDeconvolve(small.obs, small.psf, -1)
    assert e.value.code == api.INVALID_ARGUMENT


def test_ml_from_initial(clarity_gpu, small):
    api = clarity_gpu
    init = np.full(small.obs.shape, float(small.obs.mean()), np.float32)
    got = api.MaximumLikelihoodDeconvolveFrom(small.obs, small.psf, init, 4)
    ref = oracle.ml_ex(small.obs, small.psf, 4, initial=init)
    assert rel_l2(got, ref) <= 1e-4
    # resuming: 3 + 2 iterations == 5 iterations
    a = api.MaximumLikelihoodDeconvolve(small.obs, small.psf, 3)
    b = api.MaximumLikelihoodDeconvolveFrom(small.obs, small.psf, a, 2)
    c = api.MaximumLikelihoodDeconvolve(small.obs, small.psf, 5)
    assert rel_l2(b, c) <= 1e-5


def test_ml_mixed_radix_reference_geometry(clarity_gpu):
    """The reference tests' default geometry: 128x128x32 image with a 32^3
    kernel -> 160x160x64 work domain (radix 5)."""
    api = clarity_gpu
    img = np.zeros((32, 128, 128), np.float32)
    img[8:24, 32:96, 32:96] = 1.0      # Test_GenerateTrueImage
    z, y, x = np.meshgrid(*[np.arange(32) - 16] * 3, indexing="ij")
    k = np.exp(-(x * x + y * y + z * z) / (2 * 9.0)).astype(np.float32)
    k /= k.sum()
    obs = api.Convolve(img, k)
    got = api.MaximumLikelihoodDeconvolve(obs, k, 3)
    ref = oracle.ml(obs, k, 3)
    assert rel_l2(got, ref) <= 1e-4


@pytest.mark.parametrize("engine", ["fused", "generic"])
@pytest.mark.parametrize("name", ["beads64", "96x64x48"])
def test_jvc_vs_oracle(clarity_gpu, small, name, engine):
    """JVC on the fused passes over the reference domain (domains 128 x 128
    x 64 and, radix 3, 192 x 128 x 96) and on the generic driver
    (CLARITY_FORCE_GENERIC), against the oracle; 0 iterations run one, as
    the reference."""
    api = clarity_gpu
    if name == "beads64":
        obs, psf = small.obs, small.psf
    else:
        case = synth.make_workload("beads-64x64x32", seed=5, shape=(48, 64, 96))
        obs, psf = case.obs, case.psf
    for iters in (0, 1, 5):
        got = _engine(engine, lambda: api.JansenVanCittertDeconvolve(obs, psf, iters))
        ref = oracle.jvc(obs, psf, iters)
        err = rel_l2(got, ref)
        print(f"[jvc] {name} {engine} {iters} it: rel L2 {err:.2e}", flush=True)
        assert err <= 1e-4, iters


@pytest.mark.parametrize("engine", ["fused", "generic"])
def test_blind_as_written_vs_oracle(clarity_gpu, small, engine):
    """The as-written blind ML (h renormalised every iteration, never
    updated; no energy rescale): on the fused engine RL without rescale with
    the kernel normalised once, on the generic driver as written."""
    api = clarity_gpu
    k0 = synth.blind_initial_kernel("beads-64x64x32",small.psf.shape)
    for iters in (0, 3, 20):
        got = _engine(engine, lambda: api.BlindMaximumLikelihoodDeconvolve(small.obs, k0, iters))
        ref = oracle.blind_ml(small.obs, k0, iters)
        err = rel_l2(got, ref)
        print(f"[blind ml] {engine} {iters} it: rel L2 {err:.2e}", flush=True)
        assert err <= 1e-4, iters


def _blind_case(name):
    if name == "beads64":
        shape, case = (32, 64, 64), synth.make_workload("beads-64x64x32")
    else:  # 96x64x48: kernel-step domain 96 x 128 x 192, radix 3 along x and z
        shape = (48, 64, 96)
        case = synth.make_workload("beads-64x64x32", seed=7, shape=shape)
    return case.obs, synth.blind_initial_kernel("beads-64x64x32", shape)


@pytest.mark.parametrize("engine", ["fused", "generic"])
@pytest.mark.parametrize("sched", [(1, 2), (2, 1), (2, 2), (3, 2), (5, 5)])
@pytest.mark.parametrize("name", ["beads64", "96x64x48"])
def test_blind_rl_vs_oracle(clarity_gpu, name, sched, engine):
    """Alternating blind RL (outer x inner) against the float64-FFT oracle,
    on the fused engines (same-size even shapes: image steps on the RL
    engine, kernel steps on its kernel-step form) and on the generic driver.
    The oracle's float32 FFT (FFTW3 API) differs from its float64 one by
    1e-7..1e-6 on these schedules (measured on the CPU), so the float32 bar
    of 1e-4 applies."""
    api = clarity_gpu
    obs, k0 = _blind_case(name)
    o, h = _engine(engine, lambda: api.BlindRL(obs, k0, *sched))
    key = (name, sched)
    if key not in _blind_ref:
        _blind_ref[key] = oracle.blind_rl(obs, k0, *sched)
    ro, rh = _blind_ref[key]
    eo, eh = rel_l2(o, ro), rel_l2(h, rh)
    print(f"[blind] {name} {sched[0]}x{sched[1]} {engine}: err o {eo:.2e} h {eh:.2e}")
    assert eo <= 1e-4 and eh <= 1e-4
    assert abs(float(h.astype(np.float64).sum()) - 1.0) < 1e-5
    assert h.min() >= -1e-6 * h.max()


_blind_ref = {}


def _blind_scene():
    """Noise-free scene with no symmetry about the origin (beads and
    filaments inside a 16x20x24 volume, a 5x7x7 kernel), so that the blur
    stays inside the image."""
    rng = np.random.default_rng(5)
    obj = np.zeros((16, 20, 24), np.float32)
    obj[4:12, 4:16, 4:20] = synth.beads_and_filaments((8, 12, 16), 25, 2, rng,
                                                      fil_len=(6, 14))
    k_true = synth.gaussian_kernel((5, 7, 7), 1.0, 1.4)
    obs = synth.np_convolve(obj, k_true).astype(np.float32)
    return obj, k_true, obs


def test_blind_rl_fixed_point_and_recovery(clarity_gpu):
    """ClarityEx_BlindRLFrom from the true object and kernel on noise-free
    data keeps both (the kernel step is a correlation); from the
    observation with the kernel 20 % too wide, the kernel steps move the
    kernel toward the true one, as the oracle's."""
    api = clarity_gpu
    obj, k_true, obs = _blind_scene()
    kt = k_true / k_true.astype(np.float64).sum()
    o, h = _engine("generic", lambda: api.BlindRL(obs, k_true, 2, 2, initial=obj))
    assert rel_l2(h, kt) <= 1e-5 and rel_l2(o, obj) <= 1e-5
    k0 = synth.gaussian_kernel((5, 7, 7), 1.2, 1.68)
    e0 = rel_l2(k0, kt)
    for outer in (1, 2):
        o, h = _engine("generic", lambda: api.BlindRL(obs, k0, outer, 3))
        ro, rh = oracle.blind_rl(obs, k0, outer, 3)
        assert rel_l2(h, rh) <= 1e-4 and rel_l2(o, ro) <= 1e-4
        print(f"[blind] recovery {outer} x 3: kernel rel L2 {rel_l2(h, kt):.3f} (start {e0:.3f})")
        assert rel_l2(h, kt) < 0.5 * e0


def test_blind_rl_from_initial_fused(clarity_gpu):
    """ClarityEx_BlindRLFrom on the fused engines: a given first image
    estimate, against the oracle from the same estimate."""
    api = clarity_gpu
    obs, k0 = _blind_case("beads64")
    init = np.full(obs.shape, float(obs.mean()), np.float32)
    o, h = _engine("fused", lambda: api.BlindRL(obs, k0, 2, 2, initial=init))
    ro, rh = oracle.blind_rl(obs, k0, 2, 2, initial=init)
    assert rel_l2(o, ro) <= 1e-4 and rel_l2(h, rh) <= 1e-4


def test_divide_guard_and_quotient(clarity_gpu):
    """The fused passes' guarded ratio: the guard exactly at the reference's
    double compare |b| < 1e-5 (b = 0, +-1e-5f -> value; +-nextafterf(1e-5f,
    inf) -> a / b), and the quotient within 1 ulp of IEEE float division
    over a sweep of magnitudes (reference src/ComputePrimitives.cxx:143-148)."""
    api = clarity_gpu
    f32 = np.float32
    up = np.nextafter(f32(1e-5), f32(np.inf))
    b = np.array([0.0, -0.0, 1e-5, -1e-5, up, -up, 1e-6, 3.0], np.float32)
    a = np.full(b.shape, 7.0, np.float32)
    r = api.DivideArraysComponentWise(a, b, -2.5)
    assert np.array_equal(r[:4], np.full(4, -2.5, np.float32))
    assert r[6] == f32(-2.5)
    exact = (a[[4, 5, 7]] / b[[4, 5, 7]]).astype(np.float32)
    ulp = np.abs(r[[4, 5, 7]].view(np.int32).astype(np.int64)
                 - exact.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1
    # the guard's outcome is the oracle's (the reference's double compare)
    assert np.array_equal(r == f32(-2.5), oracle.divide(a, b, -2.5) == f32(-2.5))
    # numerators 1e-10..1e10, denominators 1e-5..1e20 of either sign:
    # quotients 1e-30..1e15, well inside the normal range
    rng = np.random.default_rng(17)
    n = 1 << 20
    sign = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    bb = (sign * 10.0 ** rng.uniform(-5, 20, n)).astype(np.float32)
    bb = np.where(np.abs(bb) < up, np.copysign(up, bb), bb).astype(np.float32)
    aa = (10.0 ** rng.uniform(-10, 10, n) * np.where(rng.random(n) < 0.5, -1, 1)
          ).astype(np.float32)
    got = api.DivideArraysComponentWise(aa, bb, 0.0)
    want = (aa / bb).astype(np.float32)
    fin = np.isfinite(want) & (np.abs(want) >= np.finfo(np.float32).tiny)
    d = np.abs(got[fin].view(np.int32).astype(np.int64) - want[fin].view(np.int32).astype(np.int64))
    print(f"[divide] {fin.sum()} quotients: max ulp {d.max()}, exact {np.mean(d == 0):.4f}")
    assert d.max() <= 1


def test_idivergence_vs_oracle(clarity_gpu):
    api = clarity_gpu
    g = _rand((16, 16, 16), 10, 0, 5)
    g[0, 0, :4] = 0
    gh = _rand((16, 16, 16), 11, 0.5, 5)
    assert abs(api.IDivergence(g, gh) - oracle.idivergence(g, gh)) <= 1e-9 * abs(
        oracle.idivergence(g, gh)) + 1e-9
    assert abs(api.IDivergence(gh, gh)) < 1e-6


def test_stubs_write_zeros(clarity_gpu):
    api = clarity_gpu
    a = _rand((4, 4, 4), 12)
    out = np.ones_like(a)
    rc = api.lib().Clarity_IDivergenceDeconvolve(api._p(a), api.dim3(a), api._p(a),
                                                 api.dim3(a), api._p(out))
    assert rc == api.SUCCESS and not out.any()


def test_device_oom_injection(clarity_gpu, small):
    api = clarity_gpu
    for every in (1, 3):
        api.lib().ClarityEx_InjectDeviceOOM(every)
        try:
            with pytest.raises(api.ClarityError) as e:
                api.MaximumLikelihoodDeconvolve(small.obs, small.psf, 2)
            assert e.value.code == api.DEVICE_OUT_OF_MEMORY
        finally:
            api.lib().ClarityEx_InjectDeviceOOM(0)
    # the library recovers: same call succeeds afterwards
    out = api.MaximumLikelihoodDeconvolve(small.obs, small.psf, 2)
    assert rel_l2(out, oracle.ml(small.obs, small.psf, 2)) <= 1e-4


def test_register_refcount(clarity_gpu):
    api = clarity_gpu
    api.Register()
    api.Register()
    api.UnRegister()
    api.UnRegister()
    # still registered by the session fixture
    a = _rand((4, 4, 4), 13)
    assert api.FFT_R2C(a).shape == (4, 4, 3)


def test_calls_without_register(tmp_path):
    """A process that never calls Clarity_Register still gets results (the
    reference's functions work unregistered; its raw-image app relies on
    it): the first entry point registers implicitly."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import numpy as np\n"
        "import oracle\n"
        "from clarity_amd import api, synth\n"
        "c = synth.make_workload('beads-64x64x32')\n"
        "out = api.MaximumLikelihoodDeconvolve(c.obs, c.psf, 3)\n"
        "ref = oracle.ml(c.obs, c.psf, 3)\n"
        "e = np.linalg.norm(out.astype(np.float64) - ref) / np.linalg.norm(ref)\n"
        "print('rel', e)\n"
        "sys.exit(0 if e <= 1e-4 else 3)\n" % root)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0
