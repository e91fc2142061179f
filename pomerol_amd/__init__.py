"""MI355X (gfx950) full-ED path behind pomerol's C++ API.

The compute path is ``lib/libpomx.so`` (hand-written HIP kernels behind the C ABI of
``include/pomx_c.h``) under ``lib/libpomerol_amd.so`` (the pomerol-compatible C++ API).
``_pomerol`` is a thin Python binding of that C++ API.  Importing this package fails loudly
when the native libraries are missing: there is no CPU fallback.
"""
from __future__ import annotations

import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
POMX_LIB = os.path.join(LIB_DIR, "libpomx.so")
API_LIB = os.path.join(LIB_DIR, "libpomerol_amd.so")

try:
    from . import _pomerol  # noqa: F401
except ImportError as exc:  # pragma: no cover - exercised only on broken installs
    raise ImportError(
        "pomerol_amd: native extension not built (run __graft_entry__.build() or `make -C pomerol_amd`): "
        f"{exc}"
    ) from exc

from ._pomerol import ED, Batch, G2Container, device_ok, eval_timing, kernel_stats  # noqa: E402

__all__ = ["ED", "Batch", "G2Container", "device_ok", "eval_timing", "kernel_stats", "POMX_LIB", "API_LIB", "LIB_DIR"]
