"""Writes the golden fixtures under tests/golden/ (run from the repository
root: ``python tests/golden/make_fixtures.py``).

Each fixture is a small seeded scene with the benchmark's distributions,
a camera, and what the oracle computes for it: the key count, the
canonical unsorted keys and values, the stable-sorted keys and values, the
tile ranges in both conventions (binary search: empty tile (lo, lo);
boundary scan: (0, 0)), the FP32 image at the persistent renderers'
cadence (8) and at the plain renderer's (2), the FP16-SH / FP16-feature
image at cadence 8, and the FP64 evaluation of the FP32 frame.  Inputs and
outputs are stored together (numpy .npz, no pickled objects), so the
fixtures stay valid if the scene generator changes.

The reference ships no golden vectors for this path; these pin the
oracle's restatement against regressions and give the GPU and emulated
kernels a fixed target.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

import oracle as O  # noqa: E402
from polysplat_amd import scenes  # noqa: E402

BG = (0.1, 0.2, 0.3)

FIXTURES = {
    # C1 distributions (log-scales N(-4.5, 0.5)) at a density that covers
    # the frame: 500 Gaussians in a 2-unit cube.
    "c1_small": dict(P=500, box=2.0, mix=[(1.0, -4.5, 0.5)], seed=101, W=160, H=120,
                     distance=2.6),
    # The 90/9/1 scale mixture of the 4K workload, all three emission routes,
    # an image size that is not a multiple of 16.
    "mixed_small": dict(P=400, box=2.5, mix=[(0.90, -5.0, 0.4), (0.09, -3.0, 0.4),
                                             (0.01, -1.0, 0.3)], seed=102, W=190, H=130,
                        distance=3.0),
}


def make(name: str, f: dict) -> dict:
    a = scenes.make_scene(f["P"], f["box"], f["mix"], seed=f["seed"])
    spec = scenes.make_camera(f["W"], f["H"], f["box"], distance=f["distance"])
    pos, sh, op, cov = (a.position.numpy(), a.shs.numpy(), a.opacity.numpy(), a.cov3d.numpy())
    cam = O.camera_from_spec(spec)
    r8 = O.frame(pos, sh, op, cov, cam, BG, check_every=8, fp64=True)
    r2 = O.frame(pos, sh, op, cov, cam, BG, check_every=2)
    sh16 = a.shs.half().numpy().view(np.uint16)
    r16 = O.frame(pos, sh16, op, cov, cam, BG, sh_type="fp16", feat_prec="fp16", check_every=8)
    T = O.tiles_of(f["W"], f["H"])
    return dict(
        positions=pos, shs=sh, opacities=op, cov3d=cov,
        width=np.int32(f["W"]), height=np.int32(f["H"]),
        cam_position=spec.position.numpy(), cam_rotation=spec.rotation.numpy(),
        focal=np.array([spec.focal_x, spec.focal_y], np.float32),
        z_far_near=np.array([spec.zFar, spec.zNear], np.float32),
        bg=np.array(BG, np.float32),
        num_rendered=np.int64(r8["pre"]["M"]), counts=r8["pre"]["counts"],
        keys=r8["pre"]["keys"], values=r8["pre"]["vals"],
        keys_sorted=r8["keys_sorted"], values_sorted=r8["vals_sorted"],
        ranges=r8["ranges"], ranges_scan=O.tile_ranges(r8["keys_sorted"], T, 1),
        image_check8=r8["image"], image_check2=r2["image"], image_fp16_check8=r16["image"],
        image_fp64=r8["image64"])


def main():
    for name, f in FIXTURES.items():
        d = make(name, f)
        np.savez_compressed(HERE / f"{name}.npz", **d)
        print(name, "pairs", int(d["num_rendered"]), "bytes", (HERE / f"{name}.npz").stat().st_size)


if __name__ == "__main__":
    main()
