"""Scene directories, the pixel check and the two-pass benchmark.

A scene directory holds what a trained 3D Gaussian Splatting model leaves
behind: ``point_cloud.ply`` (or ``point_cloud/iteration_<n>/point_cloud.ply``)
in the INRIA layout and ``cameras.json``.  This module loads one into the
Rasterizer facade and runs, over its cameras:

* the pixel check of scripts/correctness/check.py:118-256 with the camera
  selection and dump format of scripts/correctness/_runner_polysplat.py:
  44-190 — every selected camera rendered to a uint8 ``(H, W, 3)`` ``.npy``
  named ``<scene>__<img_name>.npy``, paired by (scene, img_name) with a
  second set (another render variant or feature precision of this build,
  or a directory of dumps written by any other renderer in that format),
  PSNR / SSIM / max / mean / diff% per image, an OK / WARN / FAIL verdict
  and a CSV;
* the benchmark protocol of scripts/benchmark/bench_polysplat.py:81-130:
  pass 1 times the end-to-end variant with one event pair per frame, pass 2
  the staged variant with per-stage events, each after its own warm-up.

``write_synthetic_scene_dir`` writes a scene directory from the seeded
generator (no trained scenes ship with this build).
"""
from __future__ import annotations

import csv
import json
import math
import os
import statistics
from pathlib import Path

import numpy as np

PLY_PROPS = (["x", "y", "z", "nx", "ny", "nz"] + [f"f_dc_{i}" for i in range(3)] +
             [f"f_rest_{i}" for i in range(45)] + ["opacity"] +
             [f"scale_{i}" for i in range(3)] + [f"rot_{i}" for i in range(4)])

E2E_VARIANT = "preranges_smem_persistent_lite_e2e"
PERKERNEL_VARIANT = "preranges_smem_persistent_lite"


# ----------------------------------------------------------------- scene dirs
def find_ply(model_path, iteration: int = 30000) -> str:
    """point_cloud.ply of a scene directory (_runner_polysplat.py:32-41)."""
    for it in (iteration, 30000, 7000, 15000):
        p = os.path.join(model_path, "point_cloud", f"iteration_{it}", "point_cloud.ply")
        if os.path.exists(p):
            return p
    p = os.path.join(model_path, "point_cloud.ply")
    if os.path.exists(p):
        return p
    raise FileNotFoundError(f"no point_cloud.ply under {model_path}")


def select_cameras(cams: list, which):
    """'first', 'all', N, 'N', 'uniform-N' or 'stride:N' (N evenly spaced)."""
    if which == "first":
        return cams[:1]
    if which == "all":
        return cams
    if isinstance(which, int):
        return cams[:which]
    if isinstance(which, str):
        if which.isdigit():
            return cams[: int(which)]
        if which.startswith("uniform-"):
            n = int(which.split("-", 1)[1])
        elif which.startswith("stride:"):
            n = int(which.split(":", 1)[1])
        else:
            raise ValueError(f"bad cameras spec: {which!r}")
        N = len(cams)
        if n <= 0 or n >= N:
            return cams
        step = (N - 1) / (n - 1) if n > 1 else 0
        return [cams[i] for i in sorted({int(round(i * step)) for i in range(n)})]
    raise ValueError(f"bad cameras spec: {which!r}")


def apply_llffhold(cams: list, hold):
    """Every hold-th camera (the usual test split); None/'all'/<= 1 keeps all."""
    if hold in (None, False, "", "none", "None", "all"):
        return cams
    hold = int(hold)
    if hold <= 1:
        return cams
    return [c for i, c in enumerate(cams) if i % hold == 0]


def load_scene_dir(model_path, device, iteration: int = 30000):
    """(Scene, [Camera], ply path) of a scene directory."""
    import contextlib
    import io

    from .rasterizer import Camera, Scene

    ply = find_ply(model_path, iteration)
    scene = Scene(device)
    with contextlib.redirect_stdout(io.StringIO()):
        scene.loadPly(ply)
    with open(os.path.join(model_path, "cameras.json")) as f:
        cams = [Camera(c) for c in json.load(f)]
    return scene, cams, ply


def safe_dump_name(name) -> str:
    return str(name).replace("/", "__").replace("\\", "__")


def write_synthetic_scene_dir(path, P: int, box: float, mix, cameras: list, seed: int = 1):
    """Writes point_cloud.ply (INRIA layout: SH DC and rest, opacity logit,
    log-scales, w-first quaternion) and cameras.json for P seeded Gaussians:
    means uniform in a cube of side `box`, per-axis log-scales from the
    (weight, mean, std) mixture `mix`.  `cameras` are dicts in cameras.json
    form (id, img_name, width, height, position, rotation rows, fx, fy)."""
    rng = np.random.default_rng(seed)
    rec = np.zeros((P, len(PLY_PROPS)), np.float32)
    rec[:, 0:3] = (rng.random((P, 3)) - 0.5) * box
    rec[:, 6:54] = rng.normal(0.0, 0.3, (P, 48))
    rec[:, 54] = rng.normal(0.0, 1.5, P)  # opacity logit
    w = np.array([m[0] for m in mix], np.float64)
    comp = rng.choice(len(mix), size=P, p=w / w.sum())
    mean = np.array([m[1] for m in mix])[comp]
    std = np.array([m[2] for m in mix])[comp]
    rec[:, 55:58] = rng.normal(0.0, 1.0, (P, 3)) * std[:, None] + mean[:, None]
    rec[:, 58:62] = rng.normal(0.0, 1.0, (P, 4))
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    head = ["ply", "format binary_little_endian 1.0", f"element vertex {P}"]
    head += [f"property float {p}" for p in PLY_PROPS] + ["end_header"]
    with open(path / "point_cloud.ply", "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        f.write(rec.astype("<f4").tobytes())
    with open(path / "cameras.json", "w") as f:
        json.dump(cameras, f)
    return path


def orbit_cameras_json(n: int, width: int, height: int, radius: float, fov_deg: float = 60.0):
    """n cameras on a circle of `radius` around the origin in the x-z plane,
    looking at the origin, in cameras.json form."""
    from .scenes import look_at_camera

    out = []
    for i in range(n):
        a = 2.0 * math.pi * i / n
        pos = (radius * math.sin(a), 0.0, -radius * math.cos(a))
        c = look_at_camera(width, height, pos, (0.0, 0.0, 0.0), fov_deg)
        out.append({"id": i, "img_name": f"view_{i:03d}", "width": width, "height": height,
                    "position": [float(v) for v in c.position],
                    "rotation": c.rotation.double().tolist(),
                    "fx": c.focal_x, "fy": c.focal_y})
    return out


# ----------------------------------------------------------------- the check
def render_dumps(scene_dirs, out_dir, variant: str = E2E_VARIANT, feature_dtype: str = "fp32",
                 cameras="first", llffhold=8, iteration: int = 30000,
                 max_num_rendered: int = 2 ** 27, max_num_tiles: int = 2 ** 20,
                 device="cuda:0", emit=None):
    """Renders the selected cameras of each scene directory into
    <out_dir>/<scene>__<img_name>.npy; returns one record per image (and
    passes each record, and errors, to `emit` when given)."""
    import torch

    from .rasterizer import Rasterizer

    dev = torch.device(device)
    os.makedirs(out_dir, exist_ok=True)
    bg = torch.zeros(3, dtype=torch.float32)
    records = []
    for scene_path in scene_dirs:
        name = os.path.basename(str(scene_path).rstrip("/"))
        try:
            scene, cams, _ = load_scene_dir(scene_path, dev, iteration)
        except Exception as e:  # reported per scene, as the reference runner does
            if emit:
                emit({"type": "error", "scene": name, "error": str(e)})
            continue
        rast = Rasterizer(scene, max_num_rendered, max_num_tiles, feature_dtype=feature_dtype)
        for cam in select_cameras(apply_llffhold(cams, llffhold), cameras):
            out = rast.forward(scene, cam, bg, render_variant=variant)
            torch.cuda.synchronize()
            path = os.path.join(out_dir, f"{name}__{safe_dump_name(cam.img_name)}.npy")
            np.save(path, out.cpu().numpy())
            rec = {"type": "image", "scene": name, "img_name": cam.img_name,
                   "camera_id": getattr(cam, "id", ""), "width": cam.width,
                   "height": cam.height, "num_vertex": scene.num_vertex, "path": path}
            records.append(rec)
            if emit:
                emit(rec)
        del rast, scene
        torch.cuda.empty_cache()
    return records


def dumps_in(dir_path, like_records):
    """Records for the dumps in `dir_path` that pair with `like_records`
    (same file names, written by any renderer)."""
    out = []
    for r in like_records:
        p = os.path.join(dir_path, os.path.basename(r["path"]))
        if os.path.exists(p):
            out.append({**r, "path": p})
    return out


def compare_pairs(recs_a, recs_b, ssim: bool = True):
    """Metrics and verdict per (scene, img_name) present on both sides."""
    from .metrics import diff_metrics, ssim_uint8, verdict

    a_by = {(r["scene"], r["img_name"]): r for r in recs_a}
    b_by = {(r["scene"], r["img_name"]): r for r in recs_b}
    rows = []
    for key in sorted(set(a_by) & set(b_by)):
        a_img, b_img = np.load(a_by[key]["path"]), np.load(b_by[key]["path"])
        if a_img.shape != b_img.shape:
            continue
        m = diff_metrics(a_img, b_img)
        m["ssim"] = ssim_uint8(a_img, b_img) if ssim else None
        a = a_by[key]
        rows.append({"scene": a["scene"], "img_name": a["img_name"], "width": a["width"],
                     "height": a["height"], "num_vertex": a.get("num_vertex"), **m,
                     "verdict": verdict(m)})
    missing = sorted((set(a_by) ^ set(b_by)))
    return rows, missing


def format_table(rows) -> str:
    if not rows:
        return "(no rows to report)"
    fp = lambda v: "  inf" if v == math.inf else f"{v:6.2f}"  # noqa: E731
    fs = lambda v: "   n/a" if v is None else f"{v:6.4f}"  # noqa: E731
    hdr = (f"{'scene':<14} {'img_name':<20} {'WxH':>12}  {'PSNR(dB)':>9}  {'SSIM':>7}  "
           f"{'max':>4}  {'mean':>6}  {'diff%':>8}  {'verdict':>7}")
    lines = [hdr, "-" * len(hdr)]
    for r in rows:
        lines.append(f"{r['scene']:<14} {r['img_name']:<20} {r['width']:>5}x{r['height']:<5}  "
                     f"{fp(r['psnr_db'])}  {fs(r['ssim'])}  {r['max_abs_diff']:>4}  "
                     f"{r['mean_abs_diff']:>6.3f}  {r['diff_pct']:>7.4f}%  {r['verdict']:>7}")
    n = {v: sum(r["verdict"] == v for r in rows) for v in ("OK", "WARN", "FAIL")}
    lines.append("-" * len(hdr))
    lines.append(f"Summary: {n['OK']} OK, {n['WARN']} WARN, {n['FAIL']} FAIL  "
                 f"(out of {len(rows)})")
    return "\n".join(lines)


def save_csv(out_dir, stamp: str, rows) -> str | None:
    if not rows:
        return None
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, f"correctness_{stamp}.csv")
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)
    return path


# ----------------------------------------------------------------- the benchmark
def pass_e2e(rast, scene, cameras, bg, warmup: int, runs: int, variant: str = E2E_VARIANT):
    """(camera, mean ms, median ms) per camera: one event pair around each
    whole forward() call."""
    import torch

    out = []
    for cam in cameras:
        for _ in range(warmup):
            rast.forward(scene, cam, bg, render_variant=variant)
        torch.cuda.synchronize()
        times = []
        for _ in range(runs):
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
            rast.forward(scene, cam, bg, render_variant=variant)
            t1.record()
            t1.synchronize()
            times.append(t0.elapsed_time(t1))
        out.append((cam, statistics.fmean(times), statistics.median(times)))
    return out


def pass_perkernel(rast, scene, cameras, bg, warmup: int, runs: int,
                   variant: str = PERKERNEL_VARIANT):
    """(camera, preprocess, sort, render mean ms, num_rendered) per camera,
    from the staged variant's per-stage events."""
    import torch

    out = []
    for cam in cameras:
        kw = dict(render_variant=variant, measure_preprocess=True, measure_sort=True,
                  measure_render=True, return_stats=True)
        for _ in range(warmup):
            rast.forward(scene, cam, bg, **kw)
        torch.cuda.synchronize()
        pp, so, rd, nr = [], [], [], 0
        for _ in range(runs):
            _, st = rast.forward(scene, cam, bg, **kw)
            pp.append(st["preprocess_ms"])
            so.append(st["sort_ms"])
            rd.append(st["render_ms"])
            nr = st["num_rendered"]
        out.append((cam, statistics.fmean(pp), statistics.fmean(so), statistics.fmean(rd), nr))
    return out
