"""Tile reordering on the GPU (render.cu:4411-7919, example.py:128-136,
:304-389, :609-718): the drop-in ops with the reference's positional
signatures (pybind.cpp:461-596) and the Rasterizer's reordered variants.

Bars: tile orders, sorted keys and packed descriptors bit-exact against the
oracle (oracle/oracle.c orc_tile_order); images of the reordered renderers
byte-identical to the same renderer without a tile order, and within one
8-bit level of the oracle run with the same precisions and saturation
cadence.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
import polysplat_amd as polysplat
from polysplat_amd import rasterizer as R
from polysplat_amd import scenes
from polysplat_amd.metrics import diff_metrics
from tests import capi
from tests.gpu_helpers import PREC, Gpu, oracle_frame, say
from tests.test_gpu_graph import _fb

pytestmark = pytest.mark.gpu

ops = polysplat.ops
BG = (0.1, 0.2, 0.3)
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def case(gpu):
    arrays = scenes.make_scene(50_000, 8.0, [(0.85, -4.0, 0.5), (0.12, -2.2, 0.4),
                                             (0.03, -0.7, 0.3)], seed=41)
    spec = scenes.make_camera(1000, 700, 8.0)
    scene = polysplat.Scene.from_arrays(arrays, DEV)
    return arrays, spec, scene


_refs = {}


def oracle(arrays, spec, sh="fp32", feat="fp32", cad=8):
    k = (sh, feat, cad)
    if k not in _refs:
        _refs[k] = oracle_frame(arrays, spec, sh_type=sh, feat_prec=feat, bg=BG,
                                check_every=cad, fp64=False)
        say("oracle", k, _refs[k]["pre"]["M"])
    return _refs[k]


def _z(*shape, dt=torch.int32):
    return torch.zeros(*shape, device=DEV, dtype=dt)


class Frame:
    """A sorted frame through the ops, with the reference's buffers."""

    def __init__(self, case, half=False):
        arrays, spec, scene = case
        ref = oracle(arrays, spec)
        M = ref["pre"]["M"] + 4096
        P = scene.num_vertex
        self.spec = spec
        self.gx = (spec.width + 15) // 16
        self.T = self.gx * ((spec.height + 15) // 16)
        self.ku, self.ks = _z(M, dt=torch.int64), _z(M, dt=torch.int64)
        self.vu, self.vs = _z(M), _z(M)
        self.space = _z(ops.get_sort_buffer_size(M), dt=torch.int8)
        self.curr = _z(1)
        self.xy, self.rgbd, self.con = (_z(P, 2, dt=torch.float32), _z(P, 4, dt=torch.float32),
                                        _z(P, 4, dt=torch.float32))
        self.ranges = _z(self.T, 2)
        self.cam = (spec.width, spec.height, 16, 16, spec.position, spec.rotation,
                    spec.focal_x, spec.focal_y, spec.zFar, spec.zNear)
        fn = ops.preprocess_half_sh if half else ops.preprocess
        fn(scene.position, scene.shs_half if half else scene.shs, scene.opacity, scene.cov3d,
           *self.cam, self.xy, self.rgbd, self.con, self.ku, self.vu, self.curr)
        self.M = int(self.curr.item())
        ops.sort_gaussian(self.M, spec.width, spec.height, 16, 16, self.space, self.ku, self.vu,
                          self.ks, self.vs)
        ops.precompute_tile_ranges(self.M, spec.width, spec.height, 16, 16, self.ks, self.ranges)
        torch.cuda.synchronize()
        self.order, self.counts, self.ids = _z(self.T), _z(self.T), _z(self.T)
        self.desc = _z(self.T, 4)
        self.temp = _z(ops.get_tile_order_sort_temp_size(self.T), dt=torch.int8)
        self.out = _z(spec.height, spec.width, 3, dt=torch.uint8)

    def image(self):
        torch.cuda.synchronize()
        return self.out.cpu().numpy().copy()


def _bg():
    return torch.tensor(BG, dtype=torch.float32)


def _u32(t):
    return t.cpu().numpy().view(np.uint32)


def test_compute_tile_order_ops_bit_exact(case):
    f = Frame(case)
    rg = f.ranges.cpu().numpy()
    # compute_tile_order (pybind.cpp:461-477)
    f.ids.fill_(-1)
    ops.compute_tile_order(f.ranges, f.T, f.order, f.counts, f.ids, f.temp)
    o_order, o_keys, _ = O.tile_order(rg, f.gx)
    np.testing.assert_array_equal(_u32(f.order), o_order)
    np.testing.assert_array_equal(_u32(f.counts), o_keys)
    np.testing.assert_array_equal(_u32(f.ids), np.arange(f.T))
    # the reference's own statement of the order: a stable descending sort
    cnt = (rg[:, 1] - rg[:, 0]).astype(np.int64)
    np.testing.assert_array_equal(o_order, np.argsort(-cnt, kind="stable"))
    # compute_tile_order_packed (pybind.cpp:506-528) and _morton (:530-554)
    for zig, bucket in ((0, 0), (264, 0), (7, 0), (0, 1280), (792, 1280), (5, 16)):
        f.desc.fill_(-7)
        if bucket:
            ops.compute_tile_order_packed_morton(f.ranges, f.T, f.gx, f.desc, f.order, f.counts,
                                                 f.ids, f.temp, zig, bucket)
        else:
            ops.compute_tile_order_packed(f.ranges, f.T, f.gx, f.desc, f.order, f.counts, f.ids,
                                          f.temp, zig)
        o_order, o_keys, o_desc = O.tile_order(rg, f.gx, zig, bucket)
        np.testing.assert_array_equal(_u32(f.order), o_order, err_msg=f"{zig} {bucket}")
        np.testing.assert_array_equal(_u32(f.counts), o_keys, err_msg=f"{zig} {bucket}")
        np.testing.assert_array_equal(f.desc.cpu().numpy(), o_desc, err_msg=f"{zig} {bucket}")


def test_tile_order_of_counts_differing_in_every_byte(gpu):
    # a synthetic range table (the op reads nothing else): counts up to 2^30
    # with ties, so every digit pass of the sort runs
    rng = np.random.default_rng(17)
    T = 70_000
    c = rng.integers(0, 1 << 30, T)
    c[::5] = c[11]
    rg = np.stack([np.zeros_like(c), c], 1).astype(np.int32)
    ranges = torch.from_numpy(rg).to(DEV)
    order, counts, ids = _z(T), _z(T), _z(T)
    temp = _z(ops.get_tile_order_sort_temp_size(T), dt=torch.int8)
    ops.compute_tile_order(ranges, T, order, counts, ids, temp)
    o_order, o_keys, _ = O.tile_order(rg, 1)
    np.testing.assert_array_equal(_u32(order), o_order)
    np.testing.assert_array_equal(_u32(counts), o_keys)


def test_reordered_render_ops(case):
    f = Frame(case)
    spec = f.spec
    W, H = spec.width, spec.height
    arrays = case[0]
    # the plain renderer without an order is the reference image here
    f.out.fill_(3)
    ops.render_16x16(f.M, W, H, f.xy, f.rgbd, f.con, f.ks, f.vs, _bg(), f.out)
    plain = f.image()
    m = diff_metrics(plain, oracle(arrays, spec, cad=2)["image"])
    assert m["max_abs_diff"] <= 1, m
    ops.compute_tile_order(f.ranges, f.T, f.order, f.counts, f.ids, f.temp)
    # render_16x16_reordered (pybind.cpp:484-504)
    f.out.fill_(9)
    ops.render_16x16_reordered(f.M, W, H, f.xy, f.rgbd, f.con, f.ks, f.vs, f.ranges, f.order,
                               _bg(), f.out)
    np.testing.assert_array_equal(f.image(), plain)
    # render_16x16_reordered_v2 (pybind.cpp:556-575) over zig-zagged descriptors
    ops.compute_tile_order_packed(f.ranges, f.T, f.gx, f.desc, f.order, f.counts, f.ids, f.temp,
                                  132)
    f.out.fill_(9)
    ops.render_16x16_reordered_v2(f.M, W, H, f.xy, f.rgbd, f.con, f.ks, f.vs, f.desc, _bg(),
                                  f.out)
    np.testing.assert_array_equal(f.image(), plain)
    # persistent: the unordered one, then over each kind of descriptor
    f.out.fill_(5)
    ops.render_16x16_preranges_smem_persistent_lite(f.M, W, H, f.xy, f.rgbd, f.con, f.vs,
                                                    f.ranges, _bg(), f.out)
    lite = f.image()
    m = diff_metrics(lite, oracle(arrays, spec, cad=8)["image"])
    assert m["max_abs_diff"] <= 1, m
    for zig, bucket in ((0, 0), (264, 0), (792, 1280)):
        if bucket:
            ops.compute_tile_order_packed_morton(f.ranges, f.T, f.gx, f.desc, f.order, f.counts,
                                                 f.ids, f.temp, zig, bucket)
        else:
            ops.compute_tile_order_packed(f.ranges, f.T, f.gx, f.desc, f.order, f.counts, f.ids,
                                          f.temp, zig)
        f.out.fill_(9)
        # render_16x16_reordered_persistent (pybind.cpp:577-596)
        ops.render_16x16_reordered_persistent(f.M, W, H, f.xy, f.rgbd, f.con, f.ks, f.vs, f.desc,
                                              _bg(), f.out)
        np.testing.assert_array_equal(f.image(), lite, err_msg=f"{zig} {bucket}")


def test_reordered_records_render(case):
    arrays, spec, scene = case
    r = oracle(arrays, spec, sh="fp16", feat="fp16", cad=8)
    M = r["pre"]["M"] + 4096
    P = scene.num_vertex
    T = ((spec.width + 15) // 16) * ((spec.height + 15) // 16)
    gx = (spec.width + 15) // 16
    cam = (spec.width, spec.height, 16, 16, spec.position, spec.rotation, spec.focal_x,
           spec.focal_y, spec.zFar, spec.zNear)
    ku, ks, vu, vs = _z(M, dt=torch.int64), _z(M, dt=torch.int64), _z(M), _z(M)
    curr, rec = _z(1), _z(24 * P, dt=torch.uint8)
    ops.preprocess_records(scene.position, scene.shs_half, scene.opacity, scene.cov3d, *cam, rec,
                           1, ku, vu, curr)
    n = int(curr.item())
    ops.sort_gaussian(n, spec.width, spec.height, 16, 16,
                      _z(ops.get_sort_buffer_size(M), dt=torch.int8), ku, vu, ks, vs)
    ranges = _z(T, 2)
    ops.precompute_tile_ranges(n, spec.width, spec.height, 16, 16, ks, ranges)
    out = _z(spec.height, spec.width, 3, dt=torch.uint8)
    ops.render_records(n, spec.width, spec.height, rec, 1, False, vs, ranges, _bg(), out)
    base = out.cpu().numpy().copy()
    assert diff_metrics(base, r["image"])["max_abs_diff"] <= 1
    desc, order, counts, ids = _z(T, 4), _z(T), _z(T), _z(T)
    temp = _z(ops.get_tile_order_sort_temp_size(T), dt=torch.int8)
    ops.compute_tile_order_packed(ranges, T, gx, desc, order, counts, ids, temp, 264)
    out.fill_(9)
    ops.render_records_reordered(n, spec.width, spec.height, rec, 1, vs, desc, _bg(), out)
    np.testing.assert_array_equal(out.cpu().numpy(), base)


ORDER_VARIANTS = ["reordered", "reordered_v2", "reordered_zigzag", "reordered_persistent",
                  "reordered_persistent_zigzag264", "reordered_persistent_morton1280",
                  "reordered_persistent_morton1280_zigzag792", "reordered_persistent_adaptive",
                  "preranges_smem_persistent", "preranges_smem_persistent_morton_zigzag"]


@pytest.mark.parametrize("feature_dtype", ["fp32", "fp16"])
@pytest.mark.parametrize("variant", ORDER_VARIANTS)
def test_rasterizer_reordered_variants(case, variant, feature_dtype):
    arrays, spec, scene = case
    T = ((spec.width + 15) // 16) * ((spec.height + 15) // 16)
    kind = R.tile_order_plan(variant, T)[0]
    persistent = kind == "persistent"
    feat = feature_dtype if persistent else "fp32"
    if variant.startswith("preranges_smem_persistent"):
        sh = feat if feat != "fp32" else "fp16"
    else:
        sh = "fp32"
    r = oracle(arrays, spec, sh=sh, feat=feat, cad=8 if persistent else 2)
    M = r["pre"]["M"]
    rast = polysplat.Rasterizer(scene, MAX_NUM_RENDERED=M + 4096, MAX_NUM_TILES=T,
                                feature_dtype=feature_dtype)
    img, stats = rast.forward(scene, spec, _bg(), render_variant=variant, return_stats=True)
    plain = img.cpu().numpy().copy()
    m = diff_metrics(plain, r["image"])
    say(variant, feature_dtype, m)
    assert m["max_abs_diff"] <= 1, (variant, m)
    assert stats["num_rendered"] == M and stats["requested_render_variant"] == variant
    res = rast.forward(scene, spec, _bg(), render_variant=variant, measure_preprocess=True,
                       measure_sort=True, measure_render=True)
    assert len(res) == 4 and all(isinstance(x, float) and x > 0 for x in res[1:])
    np.testing.assert_array_equal(res[0].cpu().numpy(), plain)


@pytest.mark.parametrize("variant,prec", [(capi.V_LITE, "fp16"), (capi.V_LITE, "fp32"),
                                          (capi.V_DT, "fp16"),
                                          (capi.V_LITE | capi.DEPTH_FIRST, "bf16")])
def test_forward_with_tile_order_flag(case, variant, prec):
    """ps_forward with PS_FORWARD_TILE_ORDER: the same image, count and
    sorted pairs as without it, eagerly and replayed from a graph."""
    arrays, spec, scene = case
    cap = 2_000_000
    a = Gpu(arrays, spec, sh_type=prec, feat_prec=prec, capacity=cap)
    st, n, ref = a.forward(BG, variant)
    capi.check(st)
    b = Gpu(arrays, spec, sh_type=prec, feat_prec=prec, capacity=cap)
    st, m, img = b.forward(BG, variant | capi.TILE_ORDER)
    capi.check(st)
    assert m == n
    np.testing.assert_array_equal(img, ref)
    np.testing.assert_array_equal(b.vs[:n].cpu().numpy(), a.vs[:n].cpu().numpy())
    L = capi.lib()
    graph = C.c_void_p()
    capi.check(L.ps_frame_graph_create(b.P, capi.ptr(b.pos), capi.ptr(b.shs), PREC[prec],
                                       capi.ptr(b.opac), capi.ptr(b.cov), C.byref(b.cam),
                                       variant | capi.TILE_ORDER, PREC[prec],
                                       C.byref(_fb(b)), C.byref(graph)))
    try:
        for _ in range(3):
            out = torch.full_like(b.out, 5)
            k = C.c_int(-1)
            capi.check(L.ps_frame_graph_launch(graph, C.byref(b.cam), capi.floats3(BG),
                                               capi.ptr(out), 1, C.byref(k), capi.stream()))
            assert k.value == n
            np.testing.assert_array_equal(out.cpu().numpy(), ref)
    finally:
        L.ps_frame_graph_destroy(graph)
    # the flag needs tile ranges, and sort space for the order
    assert b.forward(BG, capi.V_FUSED | capi.TILE_ORDER)[0] == capi.PS_ERR_INVALID
    small = Gpu(arrays, spec, sh_type=prec, feat_prec=prec, capacity=64)
    assert small.sort_bytes < L.ps_forward_tile_order_space(small.tiles)
    assert small.forward(BG, variant | capi.TILE_ORDER)[0] == capi.PS_ERR_INVALID
