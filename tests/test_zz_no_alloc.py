"""ps_render_16x16_persistent_noalloc: the persistent renderer with its tile
queue in the library's static device memory, against
ps_render_16x16_persistent with a caller-owned queue, through the C ABI.

Two calls in a row, the second on a smaller frame, each byte-identical to
the allocating entry point and within one 8-bit level of the oracle; both
range modes (precomputed tile ranges and the in-kernel search) and both
feature layouts (FP32 arrays and FP16 records)."""
import numpy as np
import pytest
import torch

from polysplat_amd import scenes
from polysplat_amd.metrics import diff_metrics
from tests import capi
from tests.gpu_helpers import PREC, Gpu, oracle_frame, say

pytestmark = pytest.mark.gpu

BG = (0.25, 0.5, 0.75)


def _frame(P, W, H, seed, prec):
    arrays = scenes.make_scene(P, 8.0, [(0.85, -4.0, 0.5), (0.12, -2.2, 0.4), (0.03, -0.7, 0.3)],
                               seed=seed)
    spec = scenes.make_camera(W, H, 8.0)
    ref = oracle_frame(arrays, spec, sh_type=prec, feat_prec=prec, bg=BG, check_every=8,
                       fp64=False)
    g = Gpu(arrays, spec, sh_type=prec, feat_prec=prec, capacity=ref["pre"]["M"] + 1024)
    g.preprocess()
    g.sort()
    g.ranges_bsearch()
    return g, ref


def _noalloc(g, use_ranges):
    g.out.fill_(7)
    capi.check(capi.lib().ps_render_16x16_persistent_noalloc(
        g.M, None, g.spec.width, g.spec.height, PREC[g.feat_prec], capi.ptr(g.xy),
        capi.ptr(g.rgbd), capi.ptr(g.conic), capi.ptr(g.records), 0, capi.ptr(g.ks),
        capi.ptr(g.vs), capi.ptr(g.ranges) if use_ranges else None, 0, capi.floats3(BG),
        capi.ptr(g.out), capi.stream()))
    torch.cuda.synchronize()
    return g.out.cpu().numpy().copy()


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_noalloc_render_matches_allocating_entry_point(gpu, prec):
    big, ref_big = _frame(60_000, 1000, 700, 41, prec)
    small, ref_small = _frame(8_000, 320, 240, 42, prec)
    for use_ranges in (True, False):
        for g, ref in ((big, ref_big), (small, ref_small)):
            say("noalloc", prec, g.spec.width, g.spec.height, "ranges" if use_ranges else "search")
            want = g.render_persistent(BG, use_ranges=use_ranges)
            got = _noalloc(g, use_ranges)
            np.testing.assert_array_equal(got, want)
            if use_ranges:
                m = diff_metrics(got, ref["image"])
                assert m["max_abs_diff"] <= 1, m
