"""ctypes mirror of the Clarity C API (include/Clarity.h, include/ClarityEx.h).

Function names follow the reference's C API (reference Clarity.h.in:123-383)
without the ``Clarity_`` prefix; argument meaning, result codes and error
behaviour are the C library's.  Volumes are numpy float32 arrays of shape
(z, y, x) -- x fastest, as the reference stores them.  A result code other
than CLARITY_SUCCESS raises :class:`ClarityError` carrying the code.

The library must be built (``python -m clarity_amd.build``); there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclarity.so")

# Clarity.h.in:90-101
FFT_FAILED = 0
OUT_OF_MEMORY = 1
DEVICE_OUT_OF_MEMORY = 2
INVALID_OPERATION = 3
INVALID_ARGUMENT = 4
SUCCESS = 5
RESULT_NAMES = {
    0: "CLARITY_FFT_FAILED",
    1: "CLARITY_OUT_OF_MEMORY",
    2: "CLARITY_DEVICE_OUT_OF_MEMORY",
    3: "CLARITY_INVALID_OPERATION",
    4: "CLARITY_INVALID_ARGUMENT",
    5: "CLARITY_SUCCESS",
}


class Dim3(ctypes.Structure):
    """Clarity_Dim3 (Clarity.h.in:110-114): x, y, z ints by value."""

    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("z", ctypes.c_int)]


class ClarityError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        super().__init__(f"{what}: {RESULT_NAMES.get(code, code)}")


_lib = None

# (name, restype, argtypes)
_F = ctypes.POINTER(ctypes.c_float)
_R = ctypes.c_int
_SIGS = [
    ("Clarity_Dim3FromArray", Dim3, [ctypes.POINTER(ctypes.c_int)]),
    ("Clarity_Register", _R, []),
    ("Clarity_UnRegister", _R, []),
    ("Clarity_Init", _R, []),
    ("Clarity_Shutdown", _R, []),
    ("Clarity_SetNumberOfThreads", _R, [ctypes.c_uint]),
    ("Clarity_ImagePadSpatialShift", _R,
     [_F, Dim3, _F, Dim3, ctypes.POINTER(ctypes.c_int), ctypes.c_float]),
    ("Clarity_ImageClip", _R, [_F, Dim3, _F, Dim3]),
    ("Clarity_ResultString", ctypes.c_char_p, [ctypes.c_int]),
    ("Clarity_WienerDeconvolve", _R, [_F, Dim3, _F, Dim3, _F, ctypes.c_float]),
    ("Clarity_JansenVanCittertDeconvolve", _R, [_F, Dim3, _F, Dim3, _F, ctypes.c_int]),
    ("Clarity_SmoothedJansenVanCittertDeconvolve", _R,
     [_F, Dim3, _F, Dim3, _F, ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p]),
    ("Clarity_IDivergenceDeconvolve", _R, [_F, Dim3, _F, Dim3, _F]),
    ("Clarity_MaximumLikelihoodDeconvolve", _R, [_F, Dim3, _F, Dim3, _F, ctypes.c_int]),
    ("Clarity_BlindMaximumLikelihoodDeconvolve", _R,
     [_F, Dim3, _F, Dim3, _F, ctypes.c_int]),
    ("Clarity_Convolve", _R, [_F, Dim3, _F, Dim3, _F]),
    ("Clarity_FFT_R2C_float", _R, [ctypes.c_int] * 3 + [_F, _F]),
    ("Clarity_FFT_C2R_float", _R, [ctypes.c_int] * 3 + [_F, _F]),
    ("ClarityEx_MaximumLikelihoodDeconvolveFrom", _R,
     [_F, Dim3, _F, Dim3, _F, _F, ctypes.c_int]),
    ("ClarityEx_BlindRL", _R, [_F, Dim3, _F, Dim3, _F, _F, ctypes.c_int, ctypes.c_int]),
    ("ClarityEx_BlindRLFrom", _R,
     [_F, Dim3, _F, Dim3, _F, _F, _F, ctypes.c_int, ctypes.c_int]),
    ("ClarityEx_DivideArraysComponentWise", _R,
     [_F, _F, _F, ctypes.c_float, ctypes.c_int64]),
    ("ClarityEx_LastEngine", ctypes.c_char_p, []),
    ("ClarityEx_ConvolveNoAlloc", _R, [_F, Dim3, _F, Dim3, _F, ctypes.c_void_p]),
    ("ClarityEx_IDivergence", _R,
     [_F, _F, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]),
    ("ClarityEx_RLPlanCreate", _R, [ctypes.POINTER(ctypes.c_void_p), Dim3, Dim3]),
    ("ClarityEx_RLPlanSetInputs", _R, [ctypes.c_void_p, _F, _F, ctypes.c_int]),
    ("ClarityEx_RLPlanReset", _R, [ctypes.c_void_p, _F, ctypes.c_int]),
    ("ClarityEx_RLPlanIterate", _R, [ctypes.c_void_p, ctypes.c_int]),
    ("ClarityEx_RLPlanGetEstimate", _R, [ctypes.c_void_p, _F, ctypes.c_int]),
    ("ClarityEx_RLPlanGetScale", _R, [ctypes.c_void_p, _F]),
    ("ClarityEx_RLPlanDestroy", _R, [ctypes.c_void_p]),
    ("ClarityEx_RLPlanDescribe", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double),
      ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
    ("ClarityEx_RLPlanSetTiming", _R, [ctypes.c_void_p, ctypes.c_int]),
    ("ClarityEx_RLPlanAutotune", _R, [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p,
                                      ctypes.c_int]),
    ("ClarityEx_RLPlanTimings", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
      ctypes.c_int, ctypes.c_int]),
    ("ClarityEx_RLPlanEngine", ctypes.c_char_p, [ctypes.c_void_p]),
    ("ClarityEx_RLPlanWorkDims", _R, [ctypes.c_void_p, ctypes.POINTER(Dim3)]),
    ("ClarityEx_Synchronize", _R, []),
    ("ClarityEx_DeviceName", ctypes.c_char_p, []),
    ("ClarityEx_InjectDeviceOOM", _R, [ctypes.c_int64]),
    ("ClarityEx_MaximumLikelihoodDeconvolveBatch", _R,
     [ctypes.c_int, ctypes.POINTER(_F), ctypes.POINTER(Dim3), ctypes.POINTER(_F),
      ctypes.POINTER(Dim3), ctypes.POINTER(_F), ctypes.c_int]),
    ("ClarityEx_UNetCreate", _R, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]),
    ("ClarityEx_UNetDestroy", _R, [ctypes.c_void_p]),
    ("ClarityEx_UNetNumParams", ctypes.c_int64, [ctypes.c_void_p]),
    ("ClarityEx_UNetGetParams", _R, [ctypes.c_void_p, _F]),
    ("ClarityEx_UNetSetParams", _R, [ctypes.c_void_p, _F]),
    ("ClarityEx_UNetPredict", _R,
     [ctypes.c_void_p, _F, Dim3, ctypes.c_float, ctypes.c_float, _F, ctypes.c_int]),
    ("ClarityEx_UNetTrainStep", _R,
     [ctypes.c_void_p, _F, _F, Dim3, ctypes.c_float, ctypes.c_float, ctypes.c_float, _F, _F,
      ctypes.c_int]),
]

EXPORTED = [name for name, _, _ in _SIGS]


def lib():
    """Loads libclarity.so (raises if it has not been built).

    PyTorch-ROCm wheels ship their own HIP and HSA runtimes under the same
    sonames as /opt/rocm's (libamdhip64.so.7, libhsa-runtime64.so.1) but
    load them by file name, so a process that loads libclarity first and
    torch afterwards ends up with two runtimes and torch finds no GPU.
    torch, when installed, is therefore imported first: libclarity then
    binds to the runtime already in the process."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m clarity_amd.build`")
        try:
            import torch  # noqa: F401  (one HIP runtime per process)
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != SUCCESS:
        raise ClarityError(rc, what)


def _vol(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 3:
        raise ValueError("volumes are 3-D (z, y, x) float32 arrays")
    return a


def _p(a):
    return a.ctypes.data_as(_F)


def dim3(a) -> Dim3:
    z, y, x = a.shape
    return Dim3(x, y, z)


def ResultString(code: int) -> str:
    return lib().Clarity_ResultString(code).decode()


def Register() -> None:
    _check(lib().Clarity_Register(), "Clarity_Register")


def UnRegister() -> None:
    _check(lib().Clarity_UnRegister(), "Clarity_UnRegister")


def SetNumberOfThreads(n: int) -> None:
    _check(lib().Clarity_SetNumberOfThreads(n), "Clarity_SetNumberOfThreads")


def DeviceName() -> str:
    return lib().ClarityEx_DeviceName().decode()


def _driver(fn, img, psf, *extra):
    img = _vol(img)
    psf = _vol(psf)
    out = np.empty_like(img)
    _check(getattr(lib(), fn)(_p(img), dim3(img), _p(psf), dim3(psf), _p(out), *extra), fn)
    return out


def Convolve(img, kernel):
    """Clarity_Convolve: linear convolution, output of img's size."""
    return _driver("Clarity_Convolve", img, kernel)


def WienerDeconvolve(img, kernel, epsilon: float):
    return _driver("Clarity_WienerDeconvolve", img, kernel, ctypes.c_float(epsilon))


def JansenVanCittertDeconvolve(img, kernel, iterations: int):
    return _driver("Clarity_JansenVanCittertDeconvolve", img, kernel, iterations)


def MaximumLikelihoodDeconvolve(img, kernel, iterations: int):
    return _driver("Clarity_MaximumLikelihoodDeconvolve", img, kernel, iterations)


def BlindMaximumLikelihoodDeconvolve(img, kernel, iterations: int):
    return _driver("Clarity_BlindMaximumLikelihoodDeconvolve", img, kernel, iterations)


def MaximumLikelihoodDeconvolveFrom(img, kernel, initial, iterations: int):
    img = _vol(img)
    psf = _vol(kernel)
    init = _vol(initial)
    if init.shape != img.shape:
        raise ValueError("initial estimate must have the image's shape")
    out = np.empty_like(img)
    _check(lib().ClarityEx_MaximumLikelihoodDeconvolveFrom(
        _p(img), dim3(img), _p(psf), dim3(psf), _p(init), _p(out), iterations),
        "ClarityEx_MaximumLikelihoodDeconvolveFrom")
    return out


def MaximumLikelihoodDeconvolveBatch(images, kernels, iterations: int):
    """ClarityEx_MaximumLikelihoodDeconvolveBatch over lists of volumes."""
    imgs = [_vol(a) for a in images]
    ks = [_vol(k) for k in kernels]
    if len(imgs) != len(ks):
        raise ValueError("one kernel per image")
    outs = [np.empty_like(a) for a in imgs]
    n = len(imgs)
    P = _F * n
    D = Dim3 * n
    _check(lib().ClarityEx_MaximumLikelihoodDeconvolveBatch(
        n, P(*[_p(a) for a in imgs]), D(*[dim3(a) for a in imgs]),
        P(*[_p(k) for k in ks]), D(*[dim3(k) for k in ks]),
        P(*[_p(o) for o in outs]), iterations), "ClarityEx_MaximumLikelihoodDeconvolveBatch")
    return outs


def BlindRL(img, kernel, outer: int, inner: int, initial=None):
    """ClarityEx_BlindRL (ClarityEx_BlindRLFrom with ``initial``, the first
    image estimate) -> (image estimate, kernel estimate)."""
    img = _vol(img)
    psf = _vol(kernel)
    out = np.empty_like(img)
    kout = np.empty_like(psf)
    if initial is None:
        _check(lib().ClarityEx_BlindRL(_p(img), dim3(img), _p(psf), dim3(psf), _p(out),
                                       _p(kout), outer, inner), "ClarityEx_BlindRL")
    else:
        init = _vol(initial)
        if init.shape != img.shape:
            raise ValueError("initial estimate must have the image's shape")
        _check(lib().ClarityEx_BlindRLFrom(_p(img), dim3(img), _p(psf), dim3(psf),
                                           _p(init), _p(out), _p(kout), outer, inner),
               "ClarityEx_BlindRLFrom")
    return out, kout


def DivideArraysComponentWise(a, b, value: float = 0.0):
    """ClarityEx_DivideArraysComponentWise: the fused passes' guarded ratio
    over flat float32 arrays."""
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    b = np.ascontiguousarray(b, np.float32).reshape(-1)
    if a.shape != b.shape:
        raise ValueError("a and b must have the same size")
    r = np.empty_like(a)
    _check(lib().ClarityEx_DivideArraysComponentWise(_p(r), _p(a), _p(b), value, a.size),
           "ClarityEx_DivideArraysComponentWise")
    return r


def LastEngine() -> str:
    """Engine of the last single-volume driver call: "fast" or "generic"."""
    return lib().ClarityEx_LastEngine().decode()


def IDivergence(g, ghat) -> float:
    g = np.ascontiguousarray(g, np.float32).reshape(-1)
    gh = np.ascontiguousarray(ghat, np.float32).reshape(-1)
    r = ctypes.c_double(0.0)
    _check(lib().ClarityEx_IDivergence(_p(g), _p(gh), g.size, ctypes.byref(r)),
           "ClarityEx_IDivergence")
    return r.value


def FFT_R2C(vol):
    """Clarity_FFT_R2C_float -> complex64 (nz, ny, nx//2+1)."""
    vol = _vol(vol)
    nz, ny, nx = vol.shape
    out = np.empty((nz, ny, nx // 2 + 1), np.complex64)
    _check(lib().Clarity_FFT_R2C_float(nx, ny, nz, _p(vol),
                                       out.view(np.float32).ctypes.data_as(_F)),
           "Clarity_FFT_R2C_float")
    return out


def FFT_C2R(spec, nx: int):
    spec = np.ascontiguousarray(spec, np.complex64)
    nz, ny, hx = spec.shape
    if hx != nx // 2 + 1:
        raise ValueError("spectrum width must be nx//2 + 1")
    out = np.empty((nz, ny, nx), np.float32)
    _check(lib().Clarity_FFT_C2R_float(nx, ny, nz,
                                       spec.view(np.float32).ctypes.data_as(_F), _p(out)),
           "Clarity_FFT_C2R_float")
    return out


def ImagePadSpatialShift(src, dst_shape, shift, fill: float = 0.0):
    src = _vol(src)
    dst = np.empty(tuple(dst_shape), np.float32)
    sh = (ctypes.c_int * 3)(*shift)
    _check(lib().Clarity_ImagePadSpatialShift(_p(dst), dim3(dst), _p(src), dim3(src),
                                              sh, fill), "Clarity_ImagePadSpatialShift")
    return dst


def ImageClip(src, dst_shape):
    src = _vol(src)
    dst = np.empty(tuple(dst_shape), np.float32)
    _check(lib().Clarity_ImageClip(_p(dst), dim3(dst), _p(src), dim3(src)),
           "Clarity_ImageClip")
    return dst


class RLPlan:
    """Resident Richardson-Lucy plan (ClarityEx_RLPlan*)."""

    def __init__(self, image_shape, kernel_shape):
        self.image_shape = tuple(image_shape)
        self.kernel_shape = tuple(kernel_shape)
        h = ctypes.c_void_p()
        iz, iy, ix = self.image_shape
        kz, ky, kx = self.kernel_shape
        _check(lib().ClarityEx_RLPlanCreate(ctypes.byref(h), Dim3(ix, iy, iz),
                                            Dim3(kx, ky, kz)), "ClarityEx_RLPlanCreate")
        self.h = h

    @property
    def engine(self) -> str:
        return lib().ClarityEx_RLPlanEngine(self.h).decode()

    @property
    def work_shape(self):
        """(nz, ny, nx) of the FFT domain the engine transforms over
        (ClarityEx_RLPlanWorkDims)."""
        d = Dim3()
        _check(lib().ClarityEx_RLPlanWorkDims(self.h, ctypes.byref(d)),
               "ClarityEx_RLPlanWorkDims")
        return (d.z, d.y, d.x)

    def set_inputs(self, img, kernel):
        img = _vol(img)
        k = _vol(kernel)
        _check(lib().ClarityEx_RLPlanSetInputs(self.h, _p(img), _p(k), 0),
               "ClarityEx_RLPlanSetInputs")

    def set_inputs_device(self, img_ptr: int, kernel_ptr: int):
        _check(lib().ClarityEx_RLPlanSetInputs(self.h, ctypes.cast(img_ptr, _F),
                                               ctypes.cast(kernel_ptr, _F), 1),
               "ClarityEx_RLPlanSetInputs")

    def reset(self, initial=None):
        if initial is None:
            _check(lib().ClarityEx_RLPlanReset(self.h, None, 0), "ClarityEx_RLPlanReset")
        else:
            init = _vol(initial)
            _check(lib().ClarityEx_RLPlanReset(self.h, _p(init), 0), "ClarityEx_RLPlanReset")

    def reset_device(self, initial_ptr: int):
        """Sets the estimate from a device buffer of the image's shape."""
        _check(lib().ClarityEx_RLPlanReset(self.h, ctypes.cast(initial_ptr, _F), 1),
               "ClarityEx_RLPlanReset")

    def iterate(self, n: int):
        _check(lib().ClarityEx_RLPlanIterate(self.h, n), "ClarityEx_RLPlanIterate")

    def estimate(self):
        out = np.empty(self.image_shape, np.float32)
        _check(lib().ClarityEx_RLPlanGetEstimate(self.h, _p(out), 0),
               "ClarityEx_RLPlanGetEstimate")
        return out

    def estimate_device(self, out_ptr: int):
        _check(lib().ClarityEx_RLPlanGetEstimate(self.h, ctypes.cast(out_ptr, _F), 1),
               "ClarityEx_RLPlanGetEstimate")

    def scale(self) -> float:
        s = ctypes.c_float()
        _check(lib().ClarityEx_RLPlanGetScale(self.h, ctypes.byref(s)),
               "ClarityEx_RLPlanGetScale")
        return s.value

    def describe(self):
        cap = 32
        names = (ctypes.c_char_p * cap)()
        byts = (ctypes.c_double * cap)()
        launches = (ctypes.c_int * cap)()
        n = lib().ClarityEx_RLPlanDescribe(self.h, names, byts, launches, cap)
        return [(names[i].decode(), byts[i], launches[i]) for i in range(n)]

    def set_timing(self, on: bool):
        _check(lib().ClarityEx_RLPlanSetTiming(self.h, 1 if on else 0),
               "ClarityEx_RLPlanSetTiming")

    def autotune(self, iterations: int = 3) -> str:
        """Times launch-shape candidates on this plan's inputs and keeps the
        fastest (ClarityEx_RLPlanAutotune); returns the choice in
        CLARITY_FAST_TUNE syntax ('' for the generic engine)."""
        buf = ctypes.create_string_buffer(256)
        _check(lib().ClarityEx_RLPlanAutotune(self.h, iterations, buf, 256),
               "ClarityEx_RLPlanAutotune")
        return buf.value.decode()

    def timings(self, reset: bool = False):
        cap = 32
        cnt = (ctypes.c_int * cap)()
        ms = (ctypes.c_double * cap)()
        n = lib().ClarityEx_RLPlanTimings(self.h, cnt, ms, cap, 1 if reset else 0)
        return [(cnt[i], ms[i]) for i in range(n)]

    def close(self):
        if self.h:
            lib().ClarityEx_RLPlanDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def Synchronize():
    _check(lib().ClarityEx_Synchronize(), "ClarityEx_Synchronize")


class UNet:
    """Learned warm-start network (ClarityEx_UNet*).  Parameters are a flat
    float32 vector: 10 conv weights [27][cout][cin], 10 biases, head (17)."""

    def __init__(self, seed: int = 0):
        h = ctypes.c_void_p()
        _check(lib().ClarityEx_UNetCreate(ctypes.byref(h), seed), "ClarityEx_UNetCreate")
        self.h = h

    @property
    def nparams(self) -> int:
        return int(lib().ClarityEx_UNetNumParams(self.h))

    def get_params(self):
        p = np.empty(self.nparams, np.float32)
        _check(lib().ClarityEx_UNetGetParams(self.h, _p(p)), "ClarityEx_UNetGetParams")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.nparams:
            raise ValueError("wrong parameter count")
        _check(lib().ClarityEx_UNetSetParams(self.h, _p(p)), "ClarityEx_UNetSetParams")

    def predict(self, obs, sigma_xy: float, sigma_z: float):
        obs = _vol(obs)
        out = np.empty_like(obs)
        _check(lib().ClarityEx_UNetPredict(self.h, _p(obs), dim3(obs), sigma_xy, sigma_z,
                                           _p(out), 0), "ClarityEx_UNetPredict")
        return out

    def predict_device(self, obs_ptr: int, shape, sigma_xy, sigma_z, out_ptr: int):
        z, y, x = shape
        _check(lib().ClarityEx_UNetPredict(self.h, ctypes.cast(obs_ptr, _F), Dim3(x, y, z),
                                           sigma_xy, sigma_z, ctypes.cast(out_ptr, _F), 1),
               "ClarityEx_UNetPredict")

    def train_step(self, obs, target, sigma_xy: float, sigma_z: float, lr: float,
                   want_grads: bool = False):
        obs = _vol(obs)
        tgt = _vol(target)
        loss = ctypes.c_float()
        g = np.empty(self.nparams, np.float32) if want_grads else None
        _check(lib().ClarityEx_UNetTrainStep(self.h, _p(obs), _p(tgt), dim3(obs), sigma_xy,
                                             sigma_z, lr, ctypes.byref(loss),
                                             _p(g) if g is not None else None, 0),
               "ClarityEx_UNetTrainStep")
        return (loss.value, g) if want_grads else loss.value

    def train_step_device(self, obs_ptr: int, tgt_ptr: int, shape, sigma_xy, sigma_z, lr):
        z, y, x = shape
        loss = ctypes.c_float()
        _check(lib().ClarityEx_UNetTrainStep(self.h, ctypes.cast(obs_ptr, _F),
                                             ctypes.cast(tgt_ptr, _F), Dim3(x, y, z),
                                             sigma_xy, sigma_z, lr, ctypes.byref(loss),
                                             None, 1), "ClarityEx_UNetTrainStep")
        return loss.value

    def close(self):
        if self.h:
            lib().ClarityEx_UNetDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
