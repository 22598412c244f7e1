"""Image drift metrics on uint8 images (scripts/correctness/check.py:118-156).

PSNR is computed over the uint8 MSE of all channels; max/mean absolute
difference and the fraction of pixels with any channel different; SSIM
(Wang et al. 2004) with a 7x7 uniform window, K1 = 0.01, K2 = 0.03, sample
covariances and the window-radius border cropped, averaged over channels;
the check's OK / WARN / FAIL verdict."""
from __future__ import annotations

import math

import numpy as np


def psnr_from_mse(mse: float) -> float:
    if mse <= 0.0:
        return math.inf
    return 20.0 * math.log10(255.0 / math.sqrt(mse))


def ssim_uint8(a: np.ndarray, b: np.ndarray, win: int = 7) -> float:
    """Mean structural similarity of two uint8 images [H, W, C] (or [H, W]),
    data range 255."""
    from scipy.ndimage import uniform_filter

    x = np.asarray(a, np.float64)
    y = np.asarray(b, np.float64)
    if x.shape != y.shape:
        raise ValueError(f"shape mismatch {x.shape} vs {y.shape}")
    if x.ndim == 2:
        x, y = x[..., None], y[..., None]
    if min(x.shape[0], x.shape[1]) < win:
        raise ValueError(f"images smaller than the {win}x{win} window")
    c1, c2 = (0.01 * 255.0) ** 2, (0.03 * 255.0) ** 2
    n = win * win
    cov_norm = n / (n - 1.0)
    r = (win - 1) // 2
    vals = []
    for ch in range(x.shape[2]):
        X, Y = x[..., ch], y[..., ch]
        f = lambda t: uniform_filter(t, size=win, mode="reflect")  # noqa: E731
        ux, uy = f(X), f(Y)
        vx = cov_norm * (f(X * X) - ux * ux)
        vy = cov_norm * (f(Y * Y) - uy * uy)
        vxy = cov_norm * (f(X * Y) - ux * uy)
        s = ((2 * ux * uy + c1) * (2 * vxy + c2)) / ((ux * ux + uy * uy + c1) * (vx + vy + c2))
        vals.append(float(s[r:s.shape[0] - r, r:s.shape[1] - r].mean()))
    return float(np.mean(vals))


def verdict(m: dict) -> str:
    """OK: PSNR >= 30 dB and SSIM >= 0.95; WARN: PSNR >= 25 and SSIM >= 0.90;
    FAIL otherwise (a missing SSIM passes its half of the test)."""
    s = m.get("ssim")
    if m["psnr_db"] >= 30 and (s is None or s >= 0.95):
        return "OK"
    if m["psnr_db"] >= 25 and (s is None or s >= 0.90):
        return "WARN"
    return "FAIL"


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
