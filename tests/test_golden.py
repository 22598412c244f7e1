"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_fixtures.py): the oracle must reproduce every stored
output bit for bit, and the HIP kernels, run on the CPU through the host
emulation, must reproduce the keys, order and ranges bit for bit and the
images within one 8-bit level."""
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle as O
from polysplat_amd import scenes
from tests import capi
from tests.emu import frames as HF

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = ["c1_small", "mixed_small"]


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def spec_of(d):
    return scenes.CameraSpec(int(d["width"]), int(d["height"]), torch.from_numpy(d["cam_position"]),
                             torch.from_numpy(d["cam_rotation"]), float(d["focal"][0]),
                             float(d["focal"][1]), float(d["z_far_near"][0]),
                             float(d["z_far_near"][1]))


def arrays_of(d):
    return scenes.SceneArrays(torch.from_numpy(d["positions"]), torch.from_numpy(d["shs"]),
                              torch.from_numpy(d["opacities"]), torch.from_numpy(d["cov3d"]))


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_the_fixture(name):
    d = load(name)
    cam = O.camera_from_spec(spec_of(d))
    bg = tuple(d["bg"])
    args = (d["positions"], d["shs"], d["opacities"], d["cov3d"], cam, bg)
    r8 = O.frame(*args, check_every=8, fp64=True)
    assert r8["pre"]["M"] == int(d["num_rendered"])
    np.testing.assert_array_equal(r8["pre"]["counts"], d["counts"])
    for k, v in (("keys", r8["pre"]["keys"]), ("values", r8["pre"]["vals"]),
                 ("keys_sorted", r8["keys_sorted"]), ("values_sorted", r8["vals_sorted"]),
                 ("ranges", r8["ranges"]), ("image_check8", r8["image"]),
                 ("image_fp64", r8["image64"])):
        np.testing.assert_array_equal(v, d[k], err_msg=k)
    np.testing.assert_array_equal(O.tile_ranges(r8["keys_sorted"], len(d["ranges"]), 1),
                                  d["ranges_scan"])
    r2 = O.frame(*args, check_every=2)
    np.testing.assert_array_equal(r2["image"], d["image_check2"])
    sh16 = torch.from_numpy(d["shs"]).half().numpy().view(np.uint16)
    r16 = O.frame(d["positions"], sh16, d["opacities"], d["cov3d"], cam, bg, sh_type="fp16",
                  feat_prec="fp16", check_every=8)
    np.testing.assert_array_equal(r16["image"], d["image_fp16_check8"])


@pytest.fixture(scope="module")
def emu():
    return HF.load()


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("prec,variant,image", [("fp32", capi.V_DEFAULT, "image_check2"),
                                                ("fp32", capi.V_LITE, "image_check8"),
                                                ("fp16", capi.V_LITE, "image_fp16_check8")])
def test_emulated_kernels_reproduce_the_fixture(emu, name, prec, variant, image):
    d = load(name)
    M = int(d["num_rendered"])
    f = HF.HostFrame(emu, arrays_of(d), spec_of(d), sh_type=prec, feat_prec=prec,
                     capacity=M + 50, garbage=True, seed=3)
    st, n, img = f.forward(tuple(d["bg"]), variant)
    capi.check_with(emu, st)
    assert n == M
    np.testing.assert_array_equal(f.ku[:M], d["keys"])
    np.testing.assert_array_equal(f.ks[:M], d["keys_sorted"])
    np.testing.assert_array_equal(f.vs[:M], d["values_sorted"])
    if variant == capi.V_LITE:
        np.testing.assert_array_equal(f.ranges, d["ranges_scan"])
    assert np.abs(img.astype(int) - d[image].astype(int)).max() <= 1
