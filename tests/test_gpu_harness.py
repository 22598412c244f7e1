"""The pixel check and the two-pass benchmark over a scene directory on the
GPU (scripts/correctness/check.py, _runner_polysplat.py,
scripts/benchmark/bench_polysplat.py): dumps of the end-to-end FP16 variant
against the plain FP32 variant and against the oracle, a check against a
directory of dumps, and both benchmark passes."""
import json

import numpy as np
import pytest
import torch

import polysplat_amd as polysplat
from polysplat_amd import bench_scene, check
from polysplat_amd import harness as H
from polysplat_amd import scenes
from polysplat_amd.metrics import diff_metrics
from tests.gpu_helpers import oracle_frame, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene_dir(tmp_path_factory, gpu):
    d = tmp_path_factory.mktemp("scenes") / "orbit"
    cams = H.orbit_cameras_json(4, 320, 240, radius=14.0)
    H.write_synthetic_scene_dir(d, 20_000, 8.0, [(0.85, -4.0, 0.5), (0.15, -2.2, 0.4)], cams,
                                seed=9)
    return d


def test_dumps_match_the_oracle_and_other_variants(scene_dir, tmp_path):
    recs_a = H.render_dumps([scene_dir], tmp_path / "a", variant="default", cameras="all",
                            llffhold=None)
    recs_b = H.render_dumps([scene_dir], tmp_path / "b", variant=H.E2E_VARIANT,
                            feature_dtype="fp16", cameras="all", llffhold=None)
    assert len(recs_a) == len(recs_b) == 4
    rows, missing = H.compare_pairs(recs_a, recs_b)
    say("check rows", [(r["img_name"], round(r["psnr_db"], 2), r["verdict"]) for r in rows])
    assert not missing and all(r["verdict"] == "OK" for r in rows)
    # the plain FP32 dump of one view against the oracle on the same scene
    scene, cams, _ = H.load_scene_dir(scene_dir, torch.device("cpu"))
    arrays = scenes.SceneArrays(scene.position, scene.shs, scene.opacity, scene.cov3d)
    cam = cams[1]
    spec = scenes.CameraSpec(cam.width, cam.height, cam.position, cam.rotation, cam.focal_x,
                             cam.focal_y)
    ref = oracle_frame(arrays, spec, check_every=2, fp64=False)
    m = diff_metrics(np.load(recs_a[1]["path"]), ref["image"])
    assert m["max_abs_diff"] <= 1, m


def test_check_cli_against_a_dump_directory(scene_dir, tmp_path):
    recs = H.render_dumps([scene_dir], tmp_path / "ref", variant=H.E2E_VARIANT,
                          feature_dtype="fp16", cameras="uniform-2", llffhold=None)
    cfg = {"scenes": [str(scene_dir)], "cameras": "uniform-2", "llffhold": None,
           "polysplat_variant": H.E2E_VARIANT, "feature_dtype": "fp16",
           "reference_dumps_dir": str(tmp_path / "ref"), "output_dir": str(tmp_path / "out"),
           "max_num_rendered": 1 << 22, "max_num_tiles": 1 << 12}
    p = tmp_path / "cfg.json"
    p.write_text(json.dumps(cfg))
    assert check.main(["--config", str(p)]) == 0
    csvs = list((tmp_path / "out").glob("correctness_*.csv"))
    assert len(csvs) == 1 and len(open(csvs[0]).readlines()) == 1 + len(recs)


def test_bench_scene_two_passes(scene_dir, capsys):
    assert bench_scene.main([str(scene_dir), "--cameras", "2", "--warmup", "1", "--runs", "3",
                             "--max-num-rendered", str(1 << 22),
                             "--max-num-tiles", str(1 << 12)]) == 0
    out = capsys.readouterr().out
    summary = json.loads(out.strip().splitlines()[-1])
    assert "[pass 1] E2E wall-clock" in out and "[pass 2] Per-kernel" in out
    assert len(summary["e2e"]) == 2 and len(summary["per_kernel"]) == 2
    for r in summary["per_kernel"]:
        assert r["num_rendered"] > 0 and r["preprocess_ms"] > 0 and r["render_ms"] > 0
