"""Top-k cache renderers (render.cu:396-513, :600-623, :6698-6800;
example.py:168-197) through the host emulation of the kernels.

Sorted values naming the most-referenced Gaussians are packed as
0x80000000 | slot of compact FP32 feature arrays holding copies of their
features; the plain and the persistent top-k renderer must then give the
image of the plain renderer over the unpacked values, byte for byte (same
features, same saturation cadence of 2).  A packed value beyond the cache
contributes nothing, and no kernel reads outside its arrays (the emulated
"device" memory is host memory).
"""
import ctypes as C
import subprocess
import sys

import numpy as np
import pytest

from polysplat_amd import scenes
from tests import capi
from tests.emu import frames as HF


@pytest.fixture(scope="module")
def emu():
    subprocess.run([sys.executable, str(HF.HERE / "build.py")], check=True,
                   stdout=subprocess.DEVNULL)
    return HF.load(asan=False)


def pack_topk(vals, P, k):
    """Top-k Gaussians by reference count (ties: smaller id first) and the
    values packed against them (example.py:168-190)."""
    counts = np.bincount(vals.astype(np.int64), minlength=P)
    ids = np.argsort(-counts, kind="stable")[:k]
    lookup = np.full(P, -1, np.int64)
    lookup[ids] = np.arange(len(ids))
    slots = lookup[vals.astype(np.int64)]
    packed = vals.astype(np.uint32).copy()
    packed[slots >= 0] = (slots[slots >= 0].astype(np.uint32) | np.uint32(0x80000000))
    return ids, packed


def test_topk_renders_equal_plain_render(emu):
    L = emu
    arr = scenes.make_config_scene("c1", seed=9, P=4000)
    spec = scenes.make_camera(200, 150, 10.0, distance=9.0)
    bg = (0.3, 0.1, 0.2)
    f = HF.HostFrame(L, arr, spec, sh_type="fp32", feat_prec="fp32")
    st, n, plain = f.forward(bg, capi.V_DEFAULT)
    capi.check_with(L, st)
    vals = f.vs[:n].copy()
    for k in (0, 1, 100, 4000):
        ids, packed = pack_topk(vals, f.P, k)
        assert (packed >= 0x80000000).sum() == np.isin(vals, ids).sum()
        txy = np.ascontiguousarray(f.xy[ids]) if k else np.zeros((1, 2), np.float32)
        trgbd = np.ascontiguousarray(f.rgbd[ids]) if k else np.zeros((1, 4), np.float32)
        tcon = np.ascontiguousarray(f.conic[ids]) if k else np.zeros((1, 4), np.float32)
        vs = np.ascontiguousarray(packed)
        q = np.zeros(512, np.int32)
        for persistent in (0, 1):
            out = np.full_like(plain, 9)
            capi.check_with(L, L.ps_render_16x16_topk(
                n, None, spec.width, spec.height, HF.ptr(f.xy), HF.ptr(f.rgbd), HF.ptr(f.conic),
                HF.ptr(f.ks), HF.ptr(vs), HF.ptr(txy), HF.ptr(trgbd), HF.ptr(tcon), len(ids),
                persistent, capi.floats3(bg), HF.ptr(out), HF.ptr(q), None))
            np.testing.assert_array_equal(out, plain, err_msg=f"k={k} persistent={persistent}")
        assert not q.any()


def test_topk_slot_beyond_the_cache_contributes_nothing(emu):
    L = emu
    arr = scenes.make_config_scene("c1", seed=10, P=3000)
    spec = scenes.make_camera(160, 112, 10.0, distance=9.0)
    f = HF.HostFrame(L, arr, spec, sh_type="fp32", feat_prec="fp32")
    st, n, plain = f.forward((0, 0, 0), capi.V_DEFAULT)
    capi.check_with(L, st)
    ids, packed = pack_topk(f.vs[:n].copy(), f.P, 10)
    txy, trgbd, tcon = (np.ascontiguousarray(a[ids]) for a in (f.xy, f.rgbd, f.conic))
    # slots 0..9 exist; a value naming slot 10 has no cache entry
    bad = packed.copy()
    hit = np.nonzero(bad >= 0x80000000)[0]
    assert len(hit) > 0
    bad[hit[0]] = np.uint32(0x80000000 | 10)
    out = np.zeros_like(plain)
    capi.check_with(L, L.ps_render_16x16_topk(
        n, None, spec.width, spec.height, HF.ptr(f.xy), HF.ptr(f.rgbd), HF.ptr(f.conic),
        HF.ptr(f.ks), HF.ptr(np.ascontiguousarray(bad)), HF.ptr(txy), HF.ptr(trgbd),
        HF.ptr(tcon), len(ids), 0, capi.floats3((0, 0, 0)), HF.ptr(out), None, None))
    # the same as removing that pair from its tile: the other tiles unchanged
    key = int(f.ks[hit[0]]) >> 32
    gx = (spec.width + 15) // 16
    ty, tx = divmod(key, gx)
    mask = np.ones(plain.shape[:2], bool)
    mask[16 * ty:16 * ty + 16, 16 * tx:16 * tx + 16] = False
    np.testing.assert_array_equal(out[mask], plain[mask])
    # bad arguments
    assert L.ps_render_16x16_topk(n, None, spec.width, spec.height, HF.ptr(f.xy), HF.ptr(f.rgbd),
                                  HF.ptr(f.conic), HF.ptr(f.ks), HF.ptr(f.vs), None, None, None,
                                  5, 0, capi.floats3((0, 0, 0)), HF.ptr(out), None,
                                  None) == capi.PS_ERR_INVALID
    assert L.ps_render_16x16_topk(n, None, spec.width, spec.height, HF.ptr(f.xy), HF.ptr(f.rgbd),
                                  HF.ptr(f.conic), HF.ptr(f.ks), HF.ptr(f.vs), None, None, None,
                                  0, 1, capi.floats3((0, 0, 0)), HF.ptr(out), None,
                                  None) == capi.PS_ERR_INVALID
