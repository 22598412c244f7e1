"""MI355X-native forward 3D Gaussian Splatting rasterizer.

A drop-in for the reference's ``polysplat.ops`` hot-path operators and for
``Rasterizer.forward(scene, cam, bg, render_variant=...)``::

    import polysplat_amd as polysplat
    polysplat.ops.preprocess(...)

The operators run hand-written HIP kernels for gfx950 through the C ABI in
``include/polysplat_amd.h``.  There is no CPU fallback: importing the package
raises if the native extension has not been built (``python
polysplat_amd/_build.py``).
"""
from __future__ import annotations

import importlib
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIBRARY_PATH = _PKG / "libpolysplat_amd.so"

try:
    import torch  # noqa: F401  (loads libc10/libtorch before the extension)
    _C = importlib.import_module("polysplat_amd._C")
except ImportError as exc:  # pragma: no cover - exercised only on broken installs
    raise ImportError(
        "polysplat_amd: the native extension is not built or cannot be loaded "
        f"({exc}); run `python polysplat_amd/_build.py` first"
    ) from exc

ops = _C.ops
version = _C.version

from .rasterizer import Camera, Rasterizer, Scene  # noqa: E402

__all__ = ["ops", "version", "Scene", "Camera", "Rasterizer", "LIBRARY_PATH"]
