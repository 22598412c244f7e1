"""Scene-directory harness on the CPU: camera selection, the synthetic scene
writer read back through ops.loadPly, SSIM against a brute-force evaluation,
the verdict thresholds, pairing and the CSV (scripts/correctness/check.py,
_runner_polysplat.py)."""
import json

import numpy as np
import pytest
import torch

import polysplat_amd as polysplat
from polysplat_amd import harness as H
from polysplat_amd import metrics


def test_camera_selection():
    cams = list(range(10))
    assert H.select_cameras(cams, "first") == [0]
    assert H.select_cameras(cams, "all") == cams
    assert H.select_cameras(cams, 3) == [0, 1, 2]
    assert H.select_cameras(cams, "4") == [0, 1, 2, 3]
    assert H.select_cameras(cams, "uniform-4") == [0, 3, 6, 9]
    assert H.select_cameras(cams, "stride:2") == [0, 9]
    assert H.select_cameras(cams, "uniform-20") == cams
    with pytest.raises(ValueError):
        H.select_cameras(cams, "bogus")
    assert H.apply_llffhold(cams, 8) == [0, 8]
    assert H.apply_llffhold(cams, None) == cams and H.apply_llffhold(cams, 1) == cams


def test_synthetic_scene_dir_round_trip(tmp_path):
    cams = H.orbit_cameras_json(3, 64, 48, radius=12.0)
    d = H.write_synthetic_scene_dir(tmp_path / "s", 500, 8.0, [(0.7, -3.0, 0.3), (0.3, -1.5, 0.3)],
                                    cams, seed=5)
    assert H.find_ply(d) == str(d / "point_cloud.ply")
    n, pos, shs, op, cov = polysplat.ops.loadPly(str(d / "point_cloud.ply"))
    assert n == 500 and pos.shape == (500, 3) and shs.shape[0] == 500
    assert torch.isfinite(cov).all() and ((op > 0) & (op < 1)).all()
    # every camera looks at the origin from the orbit radius
    back = json.load(open(d / "cameras.json"))
    for c in back:
        R = np.array(c["rotation"])
        p = np.array(c["position"])
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-6)
        np.testing.assert_allclose(np.linalg.norm(p), 12.0, rtol=1e-6)
        np.testing.assert_allclose(R[:, 2], -p / np.linalg.norm(p), atol=1e-6)
    cams_obj = [polysplat.Camera(c) for c in back]
    assert [c.img_name for c in cams_obj] == ["view_000", "view_001", "view_002"]


def _ssim_brute(a, b, win=7):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    c1, c2 = (0.01 * 255) ** 2, (0.03 * 255) ** 2
    r = win // 2
    vals = []
    for ch in range(a.shape[2]):
        s = []
        for i in range(r, a.shape[0] - r):
            for j in range(r, a.shape[1] - r):
                x = a[i - r:i + r + 1, j - r:j + r + 1, ch].ravel()
                y = b[i - r:i + r + 1, j - r:j + r + 1, ch].ravel()
                mx, my = x.mean(), y.mean()
                vx, vy = x.var(ddof=1), y.var(ddof=1)
                vxy = ((x - mx) * (y - my)).sum() / (x.size - 1)
                s.append((2 * mx * my + c1) * (2 * vxy + c2) /
                         ((mx * mx + my * my + c1) * (vx + vy + c2)))
        vals.append(np.mean(s))
    return float(np.mean(vals))


def test_ssim_against_brute_force():
    rng = np.random.default_rng(1)
    a = rng.integers(0, 256, (20, 23, 3), dtype=np.uint8)
    b = np.clip(a.astype(int) + rng.integers(-30, 31, a.shape), 0, 255).astype(np.uint8)
    assert metrics.ssim_uint8(a, a) == pytest.approx(1.0)
    assert metrics.ssim_uint8(a, b) == pytest.approx(_ssim_brute(a, b), rel=1e-9)
    c = np.zeros_like(a)
    assert metrics.ssim_uint8(a, c) == pytest.approx(_ssim_brute(a, c), rel=1e-9)


def test_verdicts():
    assert metrics.verdict({"psnr_db": 31.0, "ssim": 0.96}) == "OK"
    assert metrics.verdict({"psnr_db": 31.0, "ssim": 0.93}) == "WARN"
    assert metrics.verdict({"psnr_db": 26.0, "ssim": None}) == "WARN"
    assert metrics.verdict({"psnr_db": 24.0, "ssim": 0.99}) == "FAIL"
    assert metrics.verdict({"psnr_db": float("inf"), "ssim": 1.0}) == "OK"


def test_pairing_table_and_csv(tmp_path):
    rng = np.random.default_rng(2)
    recs_a, recs_b = [], []
    for k, noise in enumerate((0, 3, 60)):
        img = rng.integers(0, 256, (16, 16, 3), dtype=np.uint8)
        other = np.clip(img.astype(int) + rng.integers(-noise, noise + 1, img.shape), 0,
                        255).astype(np.uint8)
        for side, arr, recs in (("a", img, recs_a), ("b", other, recs_b)):
            (tmp_path / side).mkdir(exist_ok=True)
            p = tmp_path / side / f"s__v{k}.npy"
            np.save(p, arr)
            recs.append({"scene": "s", "img_name": f"v{k}", "width": 16, "height": 16,
                         "path": str(p)})
    recs_b.append({"scene": "s", "img_name": "only_b", "width": 16, "height": 16,
                   "path": str(tmp_path / "b" / "s__v0.npy")})
    rows, missing = H.compare_pairs(recs_a, recs_b)
    assert missing == [("s", "only_b")]
    assert [r["verdict"] for r in rows] == ["OK", "OK", "FAIL"]
    assert rows[0]["psnr_db"] == float("inf") and rows[0]["max_abs_diff"] == 0
    t = H.format_table(rows)
    assert "Summary: 2 OK, 0 WARN, 1 FAIL" in t
    p = H.save_csv(tmp_path / "out", "x", rows)
    assert open(p).readline().startswith("scene,img_name,width,height")
    assert H.dumps_in(tmp_path / "b", recs_a)[1]["path"] == str(tmp_path / "b" / "s__v1.npy")
