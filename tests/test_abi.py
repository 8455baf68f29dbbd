"""C-ABI checks that need no GPU: the library loads, exports exactly what
include/*.h declares, keeps the reference's enum values and struct layout,
and rejects bad arguments before touching a device."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from clarity_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("Clarity.h", "ClarityEx.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        for m in re.finditer(r"C_FUNC_DEF\s+[^;(]*?\b(Clarity\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def exported_symbols():
    out = subprocess.check_output(["nm", "-D", "--defined-only", api.LIB_PATH], text=True)
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_library_loads():
    assert api.lib() is not None


def test_exports_match_headers():
    decl = declared_symbols()
    assert len(decl) >= 30
    exp = {s for s in exported_symbols() if s.startswith("Clarity")}
    assert decl == exp, (decl ^ exp)
    assert set(api.EXPORTED) == decl


def test_reference_api_names_present():
    # the 14 public functions of the reference header (Clarity.h.in:123-383)
    ref = {"Clarity_Dim3FromArray", "Clarity_Register", "Clarity_UnRegister",
           "Clarity_SetNumberOfThreads", "Clarity_ImagePadSpatialShift",
           "Clarity_ImageClip", "Clarity_ResultString", "Clarity_WienerDeconvolve",
           "Clarity_JansenVanCittertDeconvolve",
           "Clarity_SmoothedJansenVanCittertDeconvolve",
           "Clarity_IDivergenceDeconvolve", "Clarity_MaximumLikelihoodDeconvolve",
           "Clarity_BlindMaximumLikelihoodDeconvolve", "Clarity_Convolve"}
    assert ref <= declared_symbols()


def test_enum_values_and_strings():
    assert (api.FFT_FAILED, api.OUT_OF_MEMORY, api.DEVICE_OUT_OF_MEMORY,
            api.INVALID_OPERATION, api.INVALID_ARGUMENT, api.SUCCESS) == (0, 1, 2, 3, 4, 5)
    for code, name in api.RESULT_NAMES.items():
        assert api.ResultString(code) == name
    assert api.ResultString(99) == "<unknown Clarity result>"


def test_dim3_layout():
    assert ctypes.sizeof(api.Dim3) == 12
    arr = (ctypes.c_int * 3)(7, 8, 9)
    d = api.lib().Clarity_Dim3FromArray(arr)
    assert (d.x, d.y, d.z) == (7, 8, 9)


def test_null_and_bad_arguments_rejected_without_device():
    L = api.lib()
    d = api.Dim3(4, 4, 4)
    buf = np.zeros(64, np.float32)
    p = api._p(buf)
    null = ctypes.cast(None, api._F)
    assert L.Clarity_Convolve(null, d, p, d, p) == api.INVALID_ARGUMENT
    assert L.Clarity_MaximumLikelihoodDeconvolve(p, d, p, d, p, -1) == api.INVALID_ARGUMENT
    assert L.Clarity_JansenVanCittertDeconvolve(p, d, p, d, p, -2) == api.INVALID_ARGUMENT
    assert L.Clarity_WienerDeconvolve(p, api.Dim3(0, 4, 4), p, d, p, 0.1) == api.INVALID_ARGUMENT
    assert L.Clarity_ImageClip(null, d, p, d) == api.INVALID_ARGUMENT
    sh = (ctypes.c_int * 3)(0, 0, 0)
    assert L.Clarity_ImagePadSpatialShift(p, d, null, d, sh, 0.0) == api.INVALID_ARGUMENT
    assert L.Clarity_FFT_R2C_float(0, 4, 4, p, p) == api.INVALID_ARGUMENT
    assert L.ClarityEx_BlindRL(p, d, p, d, p, p, -1, 1) == api.INVALID_ARGUMENT
    assert L.Clarity_SetNumberOfThreads(4) == api.SUCCESS


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device path")
def test_no_device_fails_loudly():
    # there is no CPU code path: without a HIP device the library refuses
    assert api.lib().Clarity_Register() == api.INVALID_OPERATION
    d = api.Dim3(4, 4, 4)
    buf = np.zeros(64, np.float32)
    p = api._p(buf)
    assert api.lib().Clarity_Convolve(p, d, p, d, p) == api.INVALID_OPERATION
    with pytest.raises(api.ClarityError):
        api.FFT_R2C(np.zeros((4, 4, 4), np.float32))
