"""MI355X-native Clarity: FFT-based 3D deconvolution on HIP (gfx950).

The product is ``libclarity.so`` (built from ``csrc/`` by ``build.py``),
a C-ABI library with the reference's C API (``include/Clarity.h``) and
extensions (``include/ClarityEx.h``).  ``clarity_amd.api`` is a thin ctypes
mirror of that API for Python callers and tests; importing it fails loudly
when the library has not been built -- there is no other code path.
"""

from .api import (  # noqa: F401
    ClarityError,
    lib,
    LIB_PATH,
)
