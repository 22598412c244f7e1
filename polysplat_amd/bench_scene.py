"""Two-pass benchmark of a scene directory (scripts/benchmark/bench_polysplat.py).

    python -m polysplat_amd.bench_scene <model_path> [--cameras first|all|N]
        [--warmup 10] [--runs 50] [--skip-e2e] [--skip-perkernel]
        [--feature-dtype fp32|fp16|bf16]

Pass 1 times the end-to-end variant with one event pair per frame and no
stage events; pass 2 (its own warm-up) the staged variant with per-stage
events (instrumentation perturbs the end-to-end time, so the two are kept
apart, bench_polysplat.py:5-14).  Prints both tables and, last, one JSON
line with every number.
"""
from __future__ import annotations

import argparse
import json
import statistics
import sys
import time

from . import harness as H


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("model_path")
    ap.add_argument("--cameras", default="first")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--runs", type=int, default=50)
    ap.add_argument("--skip-e2e", action="store_true")
    ap.add_argument("--skip-perkernel", action="store_true")
    ap.add_argument("--feature-dtype", default="fp32", choices=("fp32", "fp16", "bf16"))
    ap.add_argument("--max-num-rendered", type=int, default=2 ** 27)
    ap.add_argument("--max-num-tiles", type=int, default=2 ** 20)
    a = ap.parse_args(argv)

    import torch

    from .rasterizer import Rasterizer

    dev = torch.device("cuda:0")
    bg = torch.zeros(3, dtype=torch.float32)
    scene, cams, ply = H.load_scene_dir(a.model_path, dev)
    sel = H.select_cameras(cams, a.cameras)
    print(f"scene:       {a.model_path}")
    print(f"ply:         {ply}")
    print(f"gaussians:   {scene.num_vertex}")
    print(f"cameras:     {len(sel)} / {len(cams)} total")
    print(f"warmup/runs: {a.warmup}/{a.runs}\n")
    rast = Rasterizer(scene, a.max_num_rendered, a.max_num_tiles, feature_dtype=a.feature_dtype)
    summary = {"model_path": a.model_path, "gaussians": scene.num_vertex,
               "feature_dtype": a.feature_dtype, "warmup": a.warmup, "runs": a.runs}
    if not a.skip_e2e:
        print(f"[pass 1] E2E wall-clock  (variant = {H.E2E_VARIANT})")
        print(f"{'img_name':<20} {'WxH':>12} {'mean(ms)':>10} {'median(ms)':>12} {'fps':>8}")
        print("-" * 68)
        t = time.time()
        e2e = H.pass_e2e(rast, scene, sel, bg, a.warmup, a.runs)
        for cam, mean, med in e2e:
            print(f"{cam.img_name:<20} {cam.width:>5}x{cam.height:<5} {mean:>9.3f}  "
                  f"{med:>11.3f}  {1000.0 / mean:>7.1f}")
        print(f"(pass 1 wall: {time.time() - t:.1f}s)\n")
        summary["e2e"] = [{"img_name": c.img_name, "mean_ms": m, "median_ms": md}
                          for c, m, md in e2e]
        summary["e2e_mean_ms"] = statistics.fmean(m for _, m, _ in e2e)
    if not a.skip_perkernel:
        print(f"[pass 2] Per-kernel      (variant = {H.PERKERNEL_VARIANT})")
        print(f"{'img_name':<20} {'N_rendered':>11} {'preproc(ms)':>12} {'sort(ms)':>10} "
              f"{'render(ms)':>11} {'sum(ms)':>9}")
        print("-" * 80)
        t = time.time()
        pk = H.pass_perkernel(rast, scene, sel, bg, a.warmup, a.runs)
        for cam, pp, so, rd, nr in pk:
            print(f"{cam.img_name:<20} {nr:>11} {pp:>11.3f}  {so:>9.3f}  {rd:>10.3f}  "
                  f"{pp + so + rd:>8.3f}")
        print(f"(pass 2 wall: {time.time() - t:.1f}s)\n")
        summary["per_kernel"] = [{"img_name": c.img_name, "num_rendered": nr,
                                  "preprocess_ms": pp, "sort_ms": so, "render_ms": rd}
                                 for c, pp, so, rd, nr in pk]
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
