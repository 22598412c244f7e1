"""Batched frames (ps_forward_batch) against single frames (ps_forward) and
the oracle, through the C ABI.

A batch runs one preprocess, one sort, one range build and one persistent
render over several frames (different scenes, cameras and resolutions).
Bars: each frame's sorted keys (with the frame's tile offset removed) and
values are bit-exact against the oracle, and each image is byte-identical
to the same frame rendered alone with the persistent renderer.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from polysplat_amd import scenes
from polysplat_amd.metrics import diff_metrics
from tests import capi
from tests.gpu_helpers import PREC, Gpu, oracle_frame, say

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="the batched path faulted the GPU in one run and its cause has not "
                            "been found; ps_forward_batch is switched off until it is"),
]


def _frames():
    # Different scenes, resolutions (one not a multiple of 16) and poses.
    a = scenes.make_config_scene("c1", seed=21, P=30_000)
    b = scenes.make_scene(20_000, 8.0, [(0.85, -4.0, 0.5), (0.12, -2.2, 0.4), (0.03, -0.7, 0.3)],
                          seed=22)
    c = scenes.make_config_scene("c1", seed=23, P=12_000)
    return [
        (a, scenes.config_camera("c1"), (0.0, 0.0, 0.0)),
        (b, scenes.make_camera(1000, 700, 8.0), (0.1, 0.2, 0.3)),
        (c, scenes.make_camera(333, 250, 10.0, distance=9.0), (1.0, 1.0, 1.0)),
    ]


class Batch:
    """Shared key/sort/range buffers for a batch of Gpu frames."""

    def __init__(self, gpus, capacity):
        dev = gpus[0].dev
        self.gpus = gpus
        self.cap = capacity
        self.ku = torch.zeros(capacity, device=dev, dtype=torch.int64)
        self.vu = torch.zeros(capacity, device=dev, dtype=torch.int32)
        self.ks = torch.full((capacity,), -1, device=dev, dtype=torch.int64)
        self.vs = torch.full((capacity,), -1, device=dev, dtype=torch.int32)
        self.curr = torch.zeros(1, device=dev, dtype=torch.int32)
        L = capi.lib()
        self.sort_bytes = L.ps_get_sort_buffer_size(capacity)
        self.sort_space = torch.zeros(self.sort_bytes, device=dev, dtype=torch.uint8)
        self.tiles = [g.tiles for g in gpus]
        self.ranges = torch.zeros(sum(self.tiles), 2, device=dev, dtype=torch.int32)
        P = (C.c_int * len(gpus))(*[g.P for g in gpus])
        self.ws_bytes = L.ps_forward_batch_workspace_size(len(gpus), P)
        self.ws = torch.zeros(self.ws_bytes, device=dev, dtype=torch.uint8)

    def run(self, bgs, sync=True, ws=None):
        gpus = self.gpus
        frames = (capi.PsFrame * len(gpus))()
        for f, (g, bg) in enumerate(zip(gpus, bgs)):
            q = frames[f]
            q.P = g.P
            q.positions, q.shs = g.pos.data_ptr(), g.shs.data_ptr()
            q.opacities, q.cov3d = g.opac.data_ptr(), g.cov.data_ptr()
            q.cam = g.cam
            for i in range(3):
                q.bg[i] = bg[i]
            q.points_xy, q.rgb_depth, q.conic_opacity = (g.xy.data_ptr(), g.rgbd.data_ptr(),
                                                         g.conic.data_ptr())
            q.records = g.records.data_ptr()
            g.out.fill_(77)
            q.out_color = g.out.data_ptr()
        fb = capi.PsFrameBuffers()
        fb.curr_offset = self.curr.data_ptr()
        fb.sort_space, fb.sort_space_bytes = self.sort_space.data_ptr(), self.sort_bytes
        fb.keys_unsorted, fb.values_unsorted = self.ku.data_ptr(), self.vu.data_ptr()
        fb.keys_sorted, fb.values_sorted = self.ks.data_ptr(), self.vs.data_ptr()
        fb.capacity = self.cap
        fb.tile_ranges = self.ranges.data_ptr()
        w = self.ws if ws is None else ws
        fb.workspace, fb.workspace_bytes = w.data_ptr(), w.numel()
        counts = (C.c_int * len(gpus))()
        g0 = gpus[0]
        st = capi.lib().ps_forward_batch(len(gpus), frames, PREC[g0.sh_type], PREC[g0.feat_prec],
                                         C.byref(fb), 1 if sync else 0, counts, capi.stream())
        torch.cuda.synchronize()
        return st, list(counts), [g.out.cpu().numpy().copy() for g in gpus]


@pytest.fixture(scope="module")
def batch_case(gpu):
    items = _frames()
    refs = [oracle_frame(s, sp, sh_type="fp16", feat_prec="fp16", bg=bg, check_every=8,
                         fp64=False) for s, sp, bg in items]
    say("batch oracle", [r["pre"]["M"] for r in refs])
    return items, refs


def _gpus(items, refs, prec="fp16", extra=4096):
    return [Gpu(s, sp, sh_type=prec, feat_prec=prec, capacity=r["pre"]["M"] + extra)
            for (s, sp, _), r in zip(items, refs)]


def test_batch_matches_single_frames_and_oracle(batch_case):
    items, refs = batch_case
    gpus = _gpus(items, refs)
    bgs = [bg for _, _, bg in items]
    singles = []
    for g, bg in zip(gpus, bgs):
        st, n, img = g.forward(bg, capi.V_LITE)
        capi.check(st)
        singles.append((n, img))
    total = sum(r["pre"]["M"] for r in refs)
    b = Batch(gpus, total + 1000)
    say("batch run", total)
    st, counts, imgs = b.run(bgs)
    capi.check(st)
    assert counts == [r["pre"]["M"] for r in refs]
    assert int(b.curr.item()) == total
    ks = b.ks[:total].cpu().numpy().view(np.uint64)
    vs = b.vs[:total].cpu().numpy().view(np.uint32)
    off, tile0 = 0, 0
    for f, (r, (n, img)) in enumerate(zip(refs, singles)):
        M = r["pre"]["M"]
        k = ks[off:off + M]
        # frame f's keys carry tile ids offset by the tiles of the frames before it
        np.testing.assert_array_equal((k >> np.uint64(32)) - np.uint64(tile0),
                                      r["keys_sorted"] >> np.uint64(32), err_msg=f"frame {f}")
        np.testing.assert_array_equal(k & np.uint64(0xffffffff),
                                      r["keys_sorted"] & np.uint64(0xffffffff))
        np.testing.assert_array_equal(vs[off:off + M], r["vals_sorted"])
        np.testing.assert_array_equal(imgs[f], img, err_msg=f"frame {f} batch != single")
        m = diff_metrics(imgs[f], r["image"])
        say("frame", f, m)
        assert m["max_abs_diff"] <= 1, m
        off += M
        tile0 += gpus[f].tiles
    rng = b.ranges.cpu().numpy()
    t0 = 0
    for f, r in enumerate(refs):
        T = gpus[f].tiles
        ends = rng[t0:t0 + T, 1] - rng[t0:t0 + T, 0]
        np.testing.assert_array_equal(ends, r["ranges"][:, 1] - r["ranges"][:, 0])
        t0 += T
    # repeated batch: identical (deterministic, queue words reset)
    st, counts2, imgs2 = b.run(bgs)
    capi.check(st)
    for a, c in zip(imgs, imgs2):
        np.testing.assert_array_equal(a, c)


def test_batch_order_and_subsets_do_not_change_results(batch_case):
    items, refs = batch_case
    gpus = _gpus(items, refs)
    bgs = [bg for _, _, bg in items]
    total = sum(r["pre"]["M"] for r in refs)
    full = Batch(gpus, total + 1000)
    st, _, imgs = full.run(bgs)
    capi.check(st)
    # reversed order, and each frame alone as a batch of one
    rev = Batch(gpus[::-1], total + 1000)
    st, counts, imgs_rev = rev.run(bgs[::-1])
    capi.check(st)
    assert counts == [r["pre"]["M"] for r in refs][::-1]
    for a, c in zip(imgs, imgs_rev[::-1]):
        np.testing.assert_array_equal(a, c)
    for f in range(len(gpus)):
        one = Batch([gpus[f]], refs[f]["pre"]["M"] + 10)
        st, counts, im = one.run([bgs[f]])
        capi.check(st)
        assert counts == [refs[f]["pre"]["M"]]
        np.testing.assert_array_equal(im[0], imgs[f])


def test_batch_with_empty_scene_and_bf16(batch_case):
    items, refs = batch_case
    s0, sp0, bg0 = items[0]
    empty = scenes.SceneArrays(s0.position[:0], s0.shs[:0], s0.opacity[:0], s0.cov3d[:0])
    g_empty = Gpu(empty, scenes.make_camera(100, 60, 2.0), sh_type="bf16", feat_prec="bf16")
    g0 = Gpu(s0, sp0, sh_type="bf16", feat_prec="bf16", capacity=refs[0]["pre"]["M"] + 4096)
    st, n, single = g0.forward(bg0, capi.V_LITE)
    capi.check(st)
    b = Batch([g_empty, g0, g_empty], n + 64)
    st, counts, imgs = b.run([(0.5, 0.25, 1.0), bg0, (0.5, 0.25, 1.0)])
    capi.check(st)
    assert counts == [0, n, 0]
    np.testing.assert_array_equal(imgs[1], single)
    for f in (0, 2):
        assert (imgs[f][..., 0] == 127).all() and (imgs[f][..., 1] == 63).all()
        assert (imgs[f][..., 2] == 255).all()


def test_batch_overflow_is_reported(batch_case):
    items, refs = batch_case
    gpus = _gpus(items, refs)
    total = sum(r["pre"]["M"] for r in refs)
    b = Batch(gpus, total - 1)
    b.ku.fill_(-7)
    st, counts, imgs = b.run([(0.0, 0.0, 0.0)] * 3)
    assert st == capi.PS_ERR_OVERFLOW
    assert int(b.curr.item()) == total
    assert (b.ku == -7).all().item()
    for im in imgs:  # nothing sorted: every tile shows the background
        assert (im == 0).all()
    # exactly enough room works, with the same workspace
    b2 = Batch(gpus, total)
    st, counts, _ = b2.run([(0.0, 0.0, 0.0)] * 3, ws=b.ws)
    capi.check(st)
    assert counts == [r["pre"]["M"] for r in refs]
