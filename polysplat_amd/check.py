"""Pixel check of this build's renders against a second set of renders.

    python -m polysplat_amd.check --config cfg.json [--keep-images]

The configuration follows scripts/correctness/config.smoke.json: "scenes"
(scene directories), "iteration", "cameras" ("first", "all", N,
"uniform-N"), "llffhold", "polysplat_variant", "max_num_rendered",
"max_num_tiles", "output_dir".  The second side is either another
configuration of this build ("compare_variant" and/or
"compare_feature_dtype") or "reference_dumps_dir": a directory of
``<scene>__<img_name>.npy`` uint8 images in the reference runners' dump
format (_runner_polysplat.py:176-183), e.g. rendered by the reference on
its own hardware.  Prints the table and verdicts of check.py:198-241 and
writes a CSV under output_dir; exits 1 when any image FAILs.
"""
from __future__ import annotations

import argparse
import datetime
import json
import shutil
import sys
import tempfile

from . import harness as H


def log(msg):
    print(f"[check] {msg}", flush=True)


def run(cfg: dict, keep_images: bool = False) -> list:
    variant = cfg.get("polysplat_variant", H.E2E_VARIANT)
    common = dict(cameras=cfg.get("cameras", "first"), llffhold=cfg.get("llffhold", 8),
                  iteration=int(cfg.get("iteration", 30000)),
                  max_num_rendered=int(cfg.get("max_num_rendered", 2 ** 27)),
                  max_num_tiles=int(cfg.get("max_num_tiles", 2 ** 20)),
                  device=cfg.get("device", "cuda:0"))
    work = tempfile.mkdtemp(prefix="polysplat_check_")
    try:
        recs_a = H.render_dumps(cfg["scenes"], f"{work}/a", variant=variant,
                                feature_dtype=cfg.get("feature_dtype", "fp32"),
                                emit=lambda e: log(json.dumps(e)) if e["type"] == "error"
                                else None, **common)
        if cfg.get("reference_dumps_dir"):
            recs_b = H.dumps_in(cfg["reference_dumps_dir"], recs_a)
        else:
            recs_b = H.render_dumps(cfg["scenes"], f"{work}/b",
                                    variant=cfg.get("compare_variant", variant),
                                    feature_dtype=cfg.get("compare_feature_dtype",
                                                          cfg.get("feature_dtype", "fp32")),
                                    **common)
        rows, missing = H.compare_pairs(recs_a, recs_b)
        for k in missing:
            log(f"MISSING on one side: {k}")
        print(H.format_table(rows), flush=True)
        stamp = datetime.datetime.now().strftime("%Y%m%d_%H%M%S")
        out_dir = cfg.get("output_dir", "./bench_results")
        p = H.save_csv(out_dir, stamp, rows)
        if p:
            log(f"wrote {p}")
        if keep_images:
            dst = f"{out_dir}/images_{stamp}"
            shutil.copytree(work, dst)
            log(f"images kept in {dst}")
        return rows
    finally:
        shutil.rmtree(work, ignore_errors=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--config", required=True)
    ap.add_argument("--keep-images", action="store_true")
    a = ap.parse_args(argv)
    with open(a.config) as f:
        cfg = json.load(f)
    rows = run(cfg, a.keep_images)
    return 1 if any(r["verdict"] == "FAIL" for r in rows) else 0


if __name__ == "__main__":
    sys.exit(main())
