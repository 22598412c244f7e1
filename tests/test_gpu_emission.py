"""Key emission on the GPU, through the C ABI: the record form (count pass
leaves hit masks, emit pass writes from them) and the three-launch form
(count, scan, re-projecting emit) against the oracle.

The scene mixes small splats with a few that cover hundreds of tiles, so
both the hit-mask route and the re-sweep of large rects run.  Bars: keys,
values and features bit-exact; the pair count and status word as the oracle
implies; repeated launches on one workspace give the same keys.
"""
import numpy as np
import pytest
import torch

import oracle as O
from polysplat_amd import scenes
from tests import capi
from tests.gpu_helpers import Gpu, say

pytestmark = pytest.mark.gpu


def _scene(P=20_000):
    sc = scenes.make_scene(P, 6.0, [(0.9, -3.5, 0.5), (0.097, -2.0, 0.4), (0.003, 0.2, 0.3)],
                           seed=41)
    return sc, scenes.make_camera(640, 480, 6.0, distance=7.0)


@pytest.fixture(scope="module")
def case(gpu):
    sc, spec = _scene()
    cam = O.camera_from_spec(spec)
    ref = O.preprocess(sc.position.numpy(), sc.shs.numpy(), sc.opacity.numpy(),
                       sc.cov3d.numpy(), cam)
    say("emission case", ref["M"], int(ref["counts"].max()))
    return sc, spec, ref


@pytest.mark.parametrize("form", ["records", "3pass"])
def test_emission_bit_exact(case, form, monkeypatch):
    if form == "3pass":
        monkeypatch.setenv("POLYSPLAT_EMIT", "3pass")
    sc, spec, ref = case
    M = ref["M"]
    g = Gpu(sc, spec, capacity=M + 4096)
    g.ku.fill_(-7)
    g.ws.random_(0, 256)  # the workspace needs no clearing
    say("emission", form)
    for rep in range(3):
        assert g.preprocess() == M
        assert g.status_word() == 0
        np.testing.assert_array_equal(g.ku[:M].cpu().numpy().view(np.uint64), ref["keys"],
                                      err_msg=f"{form} launch {rep}")
        np.testing.assert_array_equal(g.vu[:M].cpu().numpy().view(np.uint32), ref["vals"])
        assert (g.ku[M:] == -7).all().item()
    vis = ref["counts"] > 0
    np.testing.assert_array_equal(g.xy.cpu().numpy()[vis], ref["xy"][vis])
    np.testing.assert_array_equal(g.conic.cpu().numpy()[vis], ref["conic"][vis])
    np.testing.assert_array_equal(g.rgbd.cpu().numpy()[vis], ref["rgbd"][vis])


def test_emission_large_scene_against_three_pass(gpu, monkeypatch):
    # 1.5M Gaussians at 1920x1080 (5860 workgroups, 23 chunk sums).
    sc = scenes.make_scene(1_500_000, 30.0,
                           [(0.9, -3.8, 0.5), (0.097, -2.2, 0.4), (0.003, -0.5, 0.3)], seed=5)
    spec = scenes.make_camera(1920, 1080, 30.0)
    g = Gpu(sc, spec, sh_type="fp16", feat_prec="fp16", capacity=12_000_000)
    say("large emission", g.P)
    M = g.preprocess()
    assert g.status_word() == 0
    rec_k = g.ku[:M].clone()
    rec_v = g.vu[:M].clone()
    rec_feat = g.records.clone()
    M2 = g.preprocess()
    assert M2 == M and torch.equal(g.ku[:M], rec_k) and torch.equal(g.vu[:M], rec_v)
    monkeypatch.setenv("POLYSPLAT_EMIT", "3pass")
    g.ku.zero_()
    g.records.zero_()
    assert g.preprocess() == M and g.status_word() == 0
    assert torch.equal(g.ku[:M], rec_k) and torch.equal(g.vu[:M], rec_v)
    assert torch.equal(g.records, rec_feat)
    say("large emission pairs", M)
