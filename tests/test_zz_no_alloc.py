"""ClarityEx_ConvolveNoAlloc (device buffers, per-device static scratch, no
allocation) against the existing entry point Clarity_Convolve on the same
inputs, two calls in a row with the second smaller, through the C ABI.
Bar: relative L2 <= 1e-6, Convolve's bar against the oracle (the two entry
points may take different engines and work domains; each is within ~2e-7
of the oracle).  Device buffers come from torch (plumbing only)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _fp(t, api):
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), api._F)


def test_convolve_no_alloc_matches_convolve(clarity_gpu):
    import torch
    api = clarity_gpu
    L = api.lib()
    cases = [((32, 64, 64), (16, 16, 16), 1, False),   # work 48 x 80 x 80
             ((12, 20, 24), (5, 7, 3), 2, True)]       # smaller, on torch's stream
    for ishape, kshape, seed, caller_stream in cases:
        rng = np.random.default_rng(seed)
        a = rng.uniform(0.0, 10.0, ishape).astype(np.float32)
        k = rng.uniform(0.0, 1.0, kshape).astype(np.float32)
        k /= k.sum()
        ref = api.Convolve(a, k)
        da = torch.from_numpy(a).cuda()
        dk = torch.from_numpy(k).cuda()
        dout = torch.full(ishape, float("nan"), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        # a stream of the caller's own (default priority); torch's current
        # stream is the null stream, whose handle 0 means the library's
        side = torch.cuda.Stream() if caller_stream else None
        stream = ctypes.c_void_p(side.cuda_stream) if caller_stream else None
        print(f"[no-alloc] {ishape} * {kshape} caller stream {caller_stream}", flush=True)
        rc = L.ClarityEx_ConvolveNoAlloc(_fp(da, api), api.dim3(a), _fp(dk, api), api.dim3(k),
                                         _fp(dout, api), stream)
        assert rc == api.SUCCESS
        if caller_stream:
            side.synchronize()
        else:
            api.Synchronize()
        got = dout.cpu().numpy()
        err = rel_l2(got, ref)
        print(f"[no-alloc] rel L2 vs Clarity_Convolve {err:.2e}", flush=True)
        assert np.all(np.isfinite(got))
        assert err <= 1e-6


def test_convolve_no_alloc_rejects_before_launch(clarity_gpu):
    """Shapes beyond the scratch array or the generic axis lengths are
    refused on the host, before anything is enqueued."""
    import torch
    api = clarity_gpu
    L = api.lib()
    buf = torch.zeros(8, dtype=torch.float32, device="cuda")
    p = _fp(buf, api)
    big = api.Dim3(100, 100, 100)        # work 200^3: 32 MB of real volume alone
    assert L.ClarityEx_ConvolveNoAlloc(p, big, p, big, p, None) == api.INVALID_ARGUMENT
    # work z = 60 + 7 = 67, a prime above the generic radices
    assert L.ClarityEx_ConvolveNoAlloc(p, api.Dim3(4, 4, 60), p, api.Dim3(1, 1, 7), p,
                                       None) == api.FFT_FAILED
    assert L.ClarityEx_ConvolveNoAlloc(None, api.Dim3(4, 4, 4), p, api.Dim3(1, 1, 1), p,
                                       None) == api.INVALID_ARGUMENT
