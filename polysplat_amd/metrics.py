"""Image drift metrics on uint8 images (scripts/correctness/check.py:118-145).

PSNR is computed over the uint8 MSE of all channels; max/mean absolute
difference and the fraction of pixels with any channel different."""
from __future__ import annotations

import math

import numpy as np


def psnr_from_mse(mse: float) -> float:
    if mse <= 0.0:
        return math.inf
    return 20.0 * math.log10(255.0 / math.sqrt(mse))


def diff_metrics(a: np.ndarray, b: np.ndarray) -> dict:
    a32 = np.asarray(a).astype(np.int32)
    b32 = np.asarray(b).astype(np.int32)
    if a32.shape != b32.shape:
        raise ValueError(f"shape mismatch {a32.shape} vs {b32.shape}")
    d = np.abs(a32 - b32)
    mse = float((d.astype(np.float64) ** 2).mean()) if d.size else 0.0
    pix = int((d.sum(axis=-1) > 0).sum()) if d.size else 0
    total = a32.shape[0] * a32.shape[1] if a32.ndim >= 2 else a32.size
    return {
        "psnr_db": psnr_from_mse(mse),
        "max_abs_diff": int(d.max()) if d.size else 0,
        "mean_abs_diff": float(d.mean()) if d.size else 0.0,
        "diff_pixels": pix,
        "total_pixels": int(total),
        "diff_pct": (100.0 * pix / total) if total else 0.0,
    }
