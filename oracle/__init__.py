"""ctypes view of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py use this package; the product (clarity_amd/) never imports it.

Volumes are numpy float32 arrays of shape (z, y, x) -- x fastest, as the
reference stores them (src/ImageClip.cxx:41).  Half spectra are complex64
arrays of shape (nz, ny, nx//2 + 1).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclarity_oracle.so")

SUCCESS = 5
INVALID_ARGUMENT = 4
REDUCE_SERIAL_F32 = 0
REDUCE_FP64 = 1
FFT_OWN = 0
FFT_FFTW_API = 1

_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        i64 = ctypes.c_int64
        ci = ctypes.c_int
        L.oracle_fft_r2c.argtypes = [ci, ci, ci, f32p, f32p]
        L.oracle_fft_c2r.argtypes = [ci, ci, ci, f32p, f32p]
        L.oracle_reduce_sum.argtypes = [f32p, i64]
        L.oracle_reduce_sum.restype = ctypes.c_float
        L.oracle_multiply.argtypes = [f32p, f32p, f32p, i64]
        L.oracle_divide.argtypes = [f32p, f32p, f32p, ctypes.c_float, i64]
        L.oracle_scale.argtypes = [f32p, f32p, i64, ctypes.c_float]
        L.oracle_modulate.argtypes = [ci, ci, ci, f32p, f32p, f32p]
        L.oracle_image_max.argtypes = [f32p, i64]
        L.oracle_image_max.restype = ctypes.c_float
        L.oracle_jvc_update.argtypes = [i64, f32p, ctypes.c_float, ctypes.c_float,
                                        f32p, f32p, f32p]
        L.oracle_wiener_kernel.argtypes = [ci, ci, ci, f32p, f32p, f32p,
                                           ctypes.c_float]
        L.oracle_pad_shift.argtypes = [f32p, ci, ci, ci, f32p, ci, ci, ci,
                                       ctypes.POINTER(ci), ctypes.c_float]
        L.oracle_clip.argtypes = [f32p, ci, ci, ci, f32p, ci, ci, ci]
        drv = [f32p, ci, ci, ci, f32p, ci, ci, ci, f32p]
        L.oracle_convolve.argtypes = drv
        L.oracle_wiener.argtypes = drv + [ctypes.c_float]
        L.oracle_jvc.argtypes = drv + [ci]
        L.oracle_ml.argtypes = drv + [ci]
        L.oracle_blind_ml.argtypes = drv + [ci]
        L.oracle_ml_ex.argtypes = [f32p, ci, ci, ci, f32p, ci, ci, ci,
                                   ctypes.c_void_p, f32p, ci, ctypes.c_void_p]
        L.oracle_blind_rl.argtypes = [f32p, ci, ci, ci, f32p, ci, ci, ci,
                                      ctypes.c_void_p, f32p, ctypes.c_void_p, ci, ci]
        L.oracle_idivergence.argtypes = [f32p, f32p, i64]
        L.oracle_idivergence.restype = ctypes.c_double
        L.oracle_fft_backend_name.restype = ctypes.c_char_p
        L.oracle_last_loop_seconds.restype = ctypes.c_double
        L.oracle_last_deconvolve_seconds.restype = ctypes.c_double
        L.oracle_set_fft_backend.argtypes = [ci]
        L.oracle_set_reduce_mode.argtypes = [ci]
        L.oracle_set_threads.argtypes = [ci]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _dims(a):
    z, y, x = a.shape
    return x, y, z


def _check(rc, what):
    if rc != SUCCESS:
        raise RuntimeError(f"oracle {what} returned {rc}")


def set_reduce_mode(mode: int) -> None:
    lib().oracle_set_reduce_mode(mode)


def get_reduce_mode() -> int:
    return lib().oracle_get_reduce_mode()


def set_fft_backend(backend: int) -> bool:
    return lib().oracle_set_fft_backend(backend) == SUCCESS


def fft_backend_name() -> str:
    return lib().oracle_fft_backend_name().decode()


def set_threads(n: int) -> None:
    lib().oracle_set_threads(n)


def get_threads() -> int:
    return lib().oracle_get_threads()


class reduce_mode:
    """Context manager switching the reduction mode temporarily."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        self.old = get_reduce_mode()
        set_reduce_mode(self.mode)

    def __exit__(self, *exc):
        set_reduce_mode(self.old)


def fft_r2c(vol):
    vol = _f32(vol)
    nx, ny, nz = _dims(vol)
    out = np.empty((nz, ny, nx // 2 + 1), np.complex64)
    o = out.view(np.float32)
    _check(lib().oracle_fft_r2c(nx, ny, nz, vol, o), "fft_r2c")
    return out


def fft_c2r(spec, nx):
    spec = np.ascontiguousarray(spec, dtype=np.complex64)
    nz, ny, hx = spec.shape
    assert hx == nx // 2 + 1
    out = np.empty((nz, ny, nx), np.float32)
    _check(lib().oracle_fft_c2r(nx, ny, nz, spec.view(np.float32).reshape(-1), out),
           "fft_c2r")
    return out


def reduce_sum(a) -> float:
    a = _f32(a).reshape(-1)
    return float(lib().oracle_reduce_sum(a, a.size))


def divide(a, b, value=0.0):
    a = _f32(a).reshape(-1)
    b = _f32(b).reshape(-1)
    r = np.empty_like(a)
    lib().oracle_divide(r, a, b, value, a.size)
    return r


def multiply(a, b):
    a = _f32(a).reshape(-1)
    b = _f32(b).reshape(-1)
    r = np.empty_like(a)
    lib().oracle_multiply(r, a, b, a.size)
    return r


def scale(a, s):
    a = _f32(a).reshape(-1)
    r = np.empty_like(a)
    lib().oracle_scale(r, a, a.size, s)
    return r


def modulate(a, b, nx):
    a = np.ascontiguousarray(a, np.complex64)
    b = np.ascontiguousarray(b, np.complex64)
    nz, ny, hx = a.shape
    out = np.empty_like(a)
    lib().oracle_modulate(nx, ny, nz, a.view(np.float32).reshape(-1),
                          b.view(np.float32).reshape(-1),
                          out.view(np.float32).reshape(-1))
    return out


def wiener_kernel(in_ft, psf_ft, nx, eps):
    a = np.ascontiguousarray(in_ft, np.complex64)
    b = np.ascontiguousarray(psf_ft, np.complex64)
    nz, ny, hx = a.shape
    out = np.empty_like(a)
    lib().oracle_wiener_kernel(nx, ny, nz, a.view(np.float32).reshape(-1),
                               b.view(np.float32).reshape(-1),
                               out.view(np.float32).reshape(-1), eps)
    return out


def image_max(a) -> float:
    a = _f32(a).reshape(-1)
    return float(lib().oracle_image_max(a, a.size))


def jvc_update(inp, a, inv_a_sq, ik, ok):
    inp = _f32(inp).reshape(-1)
    ik = _f32(ik).reshape(-1)
    ok = _f32(ok).reshape(-1)
    out = np.empty_like(inp)
    lib().oracle_jvc_update(inp.size, inp, a, inv_a_sq, ik, ok, out)
    return out


def pad_shift(src, dst_shape, shift, fill=0.0):
    src = _f32(src)
    sz, sy, sx = src.shape
    dz, dy, dx = dst_shape
    dst = np.empty(dst_shape, np.float32)
    sh = (ctypes.c_int * 3)(*shift)
    _check(lib().oracle_pad_shift(dst, dx, dy, dz, src, sx, sy, sz, sh, fill),
           "pad_shift")
    return dst


def clip(src, dst_shape):
    src = _f32(src)
    sz, sy, sx = src.shape
    dz, dy, dx = dst_shape
    dst = np.empty(dst_shape, np.float32)
    _check(lib().oracle_clip(dst, dx, dy, dz, src, sx, sy, sz), "clip")
    return dst


def _driver(name, img, psf, *extra):
    img = _f32(img)
    psf = _f32(psf)
    ix, iy, iz = _dims(img)
    kx, ky, kz = _dims(psf)
    out = np.empty_like(img)
    rc = getattr(lib(), name)(img, ix, iy, iz, psf, kx, ky, kz, out, *extra)
    _check(rc, name)
    return out


def convolve(img, psf):
    return _driver("oracle_convolve", img, psf)


def wiener(img, psf, eps):
    return _driver("oracle_wiener", img, psf, eps)


def jvc(img, psf, iterations):
    return _driver("oracle_jvc", img, psf, iterations)


def ml(img, psf, iterations):
    return _driver("oracle_ml", img, psf, iterations)


def blind_ml(img, psf, iterations):
    return _driver("oracle_blind_ml", img, psf, iterations)


def ml_ex(img, psf, iterations, initial=None, want_trace=False):
    img = _f32(img)
    psf = _f32(psf)
    ix, iy, iz = _dims(img)
    kx, ky, kz = _dims(psf)
    out = np.empty_like(img)
    init = None if initial is None else _f32(initial)
    trace = np.zeros(max(iterations, 1), np.float32) if want_trace else None
    rc = lib().oracle_ml_ex(img, ix, iy, iz, psf, kx, ky, kz,
                            None if init is None else init.ctypes.data,
                            out, iterations,
                            None if trace is None else trace.ctypes.data)
    _check(rc, "ml_ex")
    return (out, trace[:iterations]) if want_trace else out


def blind_rl(img, psf, outer, inner, initial=None):
    """Alternating blind RL; returns (image, kernel).  ``initial`` is the
    first image estimate (None: the observation)."""
    img = _f32(img)
    psf = _f32(psf)
    ix, iy, iz = _dims(img)
    kx, ky, kz = _dims(psf)
    out = np.empty_like(img)
    psf_out = np.empty_like(psf)
    init = None if initial is None else _f32(initial)
    assert init is None or init.shape == img.shape
    rc = lib().oracle_blind_rl(img, ix, iy, iz, psf, kx, ky, kz,
                               None if init is None else init.ctypes.data, out,
                               psf_out.ctypes.data, outer, inner)
    _check(rc, "blind_rl")
    return out, psf_out


def last_loop_seconds() -> float:
    return float(lib().oracle_last_loop_seconds())


def last_deconvolve_seconds() -> float:
    """The last RL call's deconvolution scope (the reference's
    TIME_DECONVOLVE: after padding, before clipping)."""
    return float(lib().oracle_last_deconvolve_seconds())


def idivergence(g, ghat) -> float:
    g = _f32(g).reshape(-1)
    ghat = _f32(ghat).reshape(-1)
    return float(lib().oracle_idivergence(g, ghat, g.size))
