"""The raw-image app (clarity_amd/bin/DeconvolveRawImages): the reference's
apps/DeconvolveRawImages/DeconvolveRawImages.cxx command line and uint16
raw format, checked against the oracle's RL on the same float inputs."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from clarity_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "clarity_amd", "bin", "DeconvolveRawImages")


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def _run(*args):
    if not os.path.exists(APP):
        pytest.fail(f"{APP} is missing: build with `python -m clarity_amd.build`")
    return subprocess.run([APP, *map(str, args)], capture_output=True, text=True, timeout=120)


def cast_like_reference(a):
    """static_cast<unsigned short>(float) as x86-64 executes it: truncate to
    int32 (0x80000000 when out of range or NaN), keep the low 16 bits."""
    a = np.asarray(a, np.float64)
    bad = ~np.isfinite(a) | (a >= 2.0 ** 31) | (a < -2.0 ** 31)
    i = np.where(bad, -2 ** 31, np.trunc(np.where(bad, 0, a))).astype(np.int64)
    return (i & 0xFFFF).astype(np.uint16)


def test_cast_helper():
    v = np.array([0.0, 0.99, 1.0, 65535.9, 65536.0, 70000.5, -1.0, -0.5, 3e9, np.nan])
    assert cast_like_reference(v).tolist() == [0, 0, 1, 65535, 0, 4464, 65535, 0, 0, 0]


def _write_inputs(tmp_path, scale):
    case = synth.make_workload("beads-64x64x32")
    img = np.round(case.obs * scale)
    img = np.clip(img, 0, 65535).astype(np.uint16)
    k = synth.gaussian_kernel((16, 16, 16), 1.5, 3.0)
    psf = np.round(k / k.max() * 1000).astype(np.uint16)
    fi, fp = tmp_path / "img.raw", tmp_path / "psf.raw"
    img.astype("<u2").tofile(fi)
    psf.astype("<u2").tofile(fp)
    return img, psf, fi, fp


def test_usage_and_bad_input(tmp_path):
    r = _run("a", 1, 2)
    assert r.returncode == 1 and "Usage: DeconvolveRawImages" in r.stdout
    r = _run(tmp_path / "missing.raw", 4, 4, 4, tmp_path / "missing.raw", 2, 2, 2,
             tmp_path / "out.raw")
    assert r.returncode == 1 and "Could not read" in r.stdout
    # a file shorter than the stated dimensions is an error, not a partial read
    (tmp_path / "short.raw").write_bytes(b"\0" * 10)
    r = _run(tmp_path / "short.raw", 4, 4, 4, tmp_path / "short.raw", 2, 2, 2,
             tmp_path / "out.raw")
    assert r.returncode == 1 and not (tmp_path / "out.raw").exists()


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device path")
def test_no_device_exits_with_error(tmp_path):
    img, psf, fi, fp = _write_inputs(tmp_path, 1.0)
    z, y, x = img.shape
    r = _run(fi, x, y, z, fp, 16, 16, 16, tmp_path / "out.raw")
    assert r.returncode == 1 and "deconvolution failed" in r.stdout
    assert not (tmp_path / "out.raw").exists()


def _compare(got, ref_f):
    ref = cast_like_reference(ref_f)
    diff = (got.astype(np.int64) - ref.astype(np.int64)) % 65536
    diff = np.minimum(diff, 65536 - diff)      # distance on the 16-bit circle
    # float results agree to ~1e-6 relative (0.05 at 5e4); a truncation can
    # land on either side of an integer only where the value is that close
    # to one
    assert diff.max() <= 1
    assert np.count_nonzero(diff) <= max(3, ref.size // 20)


@pytest.mark.gpu
@pytest.mark.parametrize("scale", [1.0, 1500.0])
def test_app_matches_oracle(tmp_path, scale):
    """Scale 1500 pushes deconvolved bead peaks past 65535, where the
    reference's unclamped cast wraps."""
    img, psf, fi, fp = _write_inputs(tmp_path, scale)
    z, y, x = img.shape
    out = tmp_path / "out.raw"
    print(f"[app] scale {scale} image {img.shape}", flush=True)
    r = _run(fi, x, y, z, fp, 16, 16, 16, out)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"Input image dimensions: {x} x {y} x {z}" in r.stdout
    got = np.fromfile(out, "<u2").reshape(img.shape)
    ref_f = oracle.ml(img.astype(np.float32), psf.astype(np.float32), 5)
    if scale > 1:
        assert ref_f.max() > 65535             # the wrap case is exercised
    _compare(got, ref_f)
    r = _run(fi, x, y, z, fp, 16, 16, 16, out, "--clamp", "--iterations", 3)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(out, "<u2").reshape(img.shape)
    ref3 = oracle.ml(img.astype(np.float32), psf.astype(np.float32), 3)
    want = np.clip(np.trunc(ref3), 0, 65535).astype(np.int64)
    d = np.abs(got.astype(np.int64) - want)
    assert d.max() <= 1 and np.count_nonzero(d) <= max(3, want.size // 20)
