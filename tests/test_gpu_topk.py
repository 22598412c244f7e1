"""Top-k cache renderers on the GPU: the drop-in ops with the reference's
positional signatures (pybind.cpp:366-437) and the Rasterizer's top-k
variants (example.py:168-197, :676-697).

Bars: the plain and persistent top-k renders over packed values are
byte-identical to the plain render over the unpacked values and within one
8-bit level of the oracle (saturation test every 2 Gaussians); the packed
values decode to the sorted Gaussian ids through the cache's ids.
"""
import numpy as np
import pytest
import torch

import polysplat_amd as polysplat
from polysplat_amd import scenes
from polysplat_amd.metrics import diff_metrics
from tests.gpu_helpers import oracle_frame, say
from tests.test_topk import pack_topk

pytestmark = pytest.mark.gpu

ops = polysplat.ops
BG = (0.2, 0.1, 0.3)
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def case(gpu):
    arrays = scenes.make_scene(60_000, 8.0, [(0.85, -4.0, 0.5), (0.12, -2.2, 0.4),
                                             (0.03, -0.7, 0.3)], seed=61)
    spec = scenes.make_camera(1000, 700, 8.0)
    ref = oracle_frame(arrays, spec, bg=BG, check_every=2, fp64=False)
    scene = polysplat.Scene.from_arrays(arrays, DEV)
    say("topk case", ref["pre"]["M"])
    return arrays, spec, scene, ref


def test_topk_ops(case):
    arrays, spec, scene, ref = case
    W, H = spec.width, spec.height
    M = ref["pre"]["M"]
    cap = M + 4096
    z = lambda *s, dt=torch.int32: torch.zeros(*s, device=DEV, dtype=dt)  # noqa: E731
    ku, ks = z(cap, dt=torch.int64), z(cap, dt=torch.int64)
    vu, vs = z(cap), z(cap)
    curr = z(1)
    P = scene.num_vertex
    xy, rgbd, con = z(P, 2, dt=torch.float32), z(P, 4, dt=torch.float32), z(P, 4, dt=torch.float32)
    cam = (W, H, 16, 16, spec.position, spec.rotation, spec.focal_x, spec.focal_y, spec.zFar,
           spec.zNear)
    ops.preprocess(scene.position, scene.shs, scene.opacity, scene.cov3d, *cam, xy, rgbd, con,
                   ku, vu, curr)
    n = int(curr.item())
    assert n == M
    ops.sort_gaussian(n, W, H, 16, 16, z(ops.get_sort_buffer_size(cap), dt=torch.int8), ku, vu,
                      ks, vs)
    out = z(H, W, 3, dt=torch.uint8)
    bg = torch.tensor(BG, dtype=torch.float32)
    ops.render_16x16(n, W, H, xy, rgbd, con, ks, vs, bg, out)
    plain = out.cpu().numpy().copy()
    m = diff_metrics(plain, ref["image"])
    assert m["max_abs_diff"] <= 1, m
    vals = vs[:n].cpu().numpy().view(np.uint32).copy()
    e = torch.empty(0, dtype=torch.float32, device=DEV)
    fns = (ops.render_16x16_topk_smem, ops.render_16x16_topk_smem_persistent,
           ops.render_16x16_topk_smem_persistent_v2)
    # each renderer without a cache (what the reference's persistent_v2
    # variant renders with): the plain one is the plain render, the
    # persistent ones within a level of the oracle (on a frame with fewer
    # tiles than waves they work in 8x8 quadrants, each stopping on its own
    # pixels' transmittance)
    unpacked = {}
    for fn in fns:
        out.fill_(7)
        fn(n, W, H, xy, rgbd, con, ks, vs, e, e, e, 0, bg, out)
        unpacked[fn] = out.cpu().numpy().copy()
        m = diff_metrics(unpacked[fn], ref["image"])
        assert m["max_abs_diff"] <= 1, (fn, m)
    np.testing.assert_array_equal(unpacked[fns[0]], plain)
    for k in (1365, 1, 50_000):
        ids, packed = pack_topk(vals, P, k)
        idt = torch.from_numpy(ids).to(DEV)
        txy, trgbd, tcon = (torch.index_select(a, 0, idt).contiguous() for a in (xy, rgbd, con))
        vp = vs.clone()
        vp[:n] = torch.from_numpy(packed.view(np.int32)).to(DEV)
        for fn in fns:
            out.fill_(7)
            fn(n, W, H, xy, rgbd, con, ks, vp, txy, trgbd, tcon, len(ids), bg, out)
            # the cache changes where features come from, never the image
            np.testing.assert_array_equal(out.cpu().numpy(), unpacked[fn],
                                          err_msg=f"{fn} k={k}")


@pytest.mark.parametrize("variant", ["topk_smem", "topk_smem_persistent",
                                     "topk_smem_persistent_v2"])
def test_rasterizer_topk_variants(case, variant):
    arrays, spec, scene, ref = case
    M = ref["pre"]["M"]
    T = ((spec.width + 15) // 16) * ((spec.height + 15) // 16)
    rast = polysplat.Rasterizer(scene, MAX_NUM_RENDERED=M + 4096, MAX_NUM_TILES=T)
    bg = torch.tensor(BG, dtype=torch.float32)
    img, stats = rast.forward(scene, spec, bg, render_variant=variant, return_stats=True)
    got = img.cpu().numpy().copy()
    m = diff_metrics(got, ref["image"])
    say(variant, m)
    assert m["max_abs_diff"] <= 1, m
    assert stats["num_rendered"] == M
    vals = rast.gaussian_values_sorted[:M].cpu().numpy().view(np.uint32)
    if variant != "topk_smem_persistent_v2":
        # packed in place, as the reference leaves them; decoding gives the ids
        hit = vals >= 0x80000000
        assert hit.any()
        slot = (vals[hit] & 0x7FFFFFFF).astype(np.int64)
        lookup = rast.smem_topk_lookup.cpu().numpy().astype(np.int64)
        cached = np.full(1365, -1, np.int64)
        cached[lookup[lookup >= 0]] = np.nonzero(lookup >= 0)[0]
        dec = vals.astype(np.int64).copy()
        dec[hit] = cached[slot]
        np.testing.assert_array_equal(dec, ref["vals_sorted"].astype(np.int64))
    else:
        np.testing.assert_array_equal(vals, ref["vals_sorted"])
    timed = rast.forward(scene, spec, bg, render_variant=variant, measure_render=True)
    np.testing.assert_array_equal(timed[0].cpu().numpy(), got)
