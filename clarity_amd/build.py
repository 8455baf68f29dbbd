"""Builds clarity_amd/libclarity.so from clarity_amd/csrc/*.hip with hipcc.

Every source is compiled for gfx950 only.  The raw-image app
(clarity_amd/apps/DeconvolveRawImages.cpp, host code over the C API) is then
linked against the library into clarity_amd/bin/.  Packed FP32 instructions are
disabled (-packed-fp32-ops): the compiler's v_pk_* float forms are kept out
of every kernel.  Objects go to clarity_amd/_build/, the library next to this
file, where the ctypes loader finds it.

    python -m clarity_amd.build [-j N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libclarity.so")
APPS = os.path.join(HERE, "apps")
BIN = os.path.join(HERE, "bin")
CXX = os.environ.get("CXX", "g++")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    "-Wall",
    "-Wno-unused-command-line-argument",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]


def sources():
    return sorted(os.path.join(SRC, f) for f in os.listdir(SRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    # The host pass does not know the packed-fp32 target feature and says
    # so; that warning is expected and dropped.
    err = "\n".join(l for l in p.stderr.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l
                    and l.strip() != "1 warning generated." and l.strip() != "2 warnings generated.")
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{p.stderr}")
    return obj, err.strip()


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, srcs))
    for obj, err in results:
        if err and verbose:
            print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    # -z defs: an unresolved symbol (e.g. a kernel whose host stub the
    # compiler did not emit) fails the link instead of the first dlopen
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-z,defs", "-o", tmp,
           *[o for o, _ in results]]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_apps()
    return LIB


def build_apps() -> list:
    """Links every apps/*.cpp against libclarity.so (found at run time
    through the binary's $ORIGIN/.. rpath)."""
    os.makedirs(BIN, exist_ok=True)
    out = []
    for f in sorted(os.listdir(APPS)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(BIN, f[:-4])
        cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "..", "include"),
               os.path.join(APPS, f), "-o", exe + ".tmp", "-L", HERE, "-lclarity",
               "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib"]
        subprocess.check_call(cmd)
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, verbose=a.v))
