"""The drop-in Python surface on the GPU.

Every hot-path operator of ``polysplat_amd.ops`` is called with the
reference's positional signatures (pybind.cpp:122-181, :333-364, :441-459,
:616-632, :691-745, :932-1149) on buffers laid out as the reference's
Rasterizer allocates them (example.py:56-139), and ``Rasterizer.forward``
runs every implemented variant with each feature precision and each return
form (example.py:293-824).  Bars as in the parity suite: keys, values, sort
order, ranges and gathers bit-exact; images within one 8-bit level of the
oracle run with the same SH precision, feature precision and saturation
cadence.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import polysplat_amd as polysplat
from polysplat_amd import scenes
from polysplat_amd.metrics import diff_metrics
from tests import capi
from tests.gpu_helpers import oracle_frame, say

pytestmark = pytest.mark.gpu

ops = polysplat.ops
BG = (0.1, 0.2, 0.3)
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def case(gpu):
    # Mixed scale regimes (single, fragmented and massive emission routes),
    # an image size that is not a multiple of 16, a non-black background.
    arrays = scenes.make_scene(50_000, 8.0, [(0.85, -4.0, 0.5), (0.12, -2.2, 0.4),
                                             (0.03, -0.7, 0.3)], seed=31)
    spec = scenes.make_camera(1000, 700, 8.0)
    ref = oracle_frame(arrays, spec, bg=BG, check_every=8, fp64=False)
    scene = polysplat.Scene.from_arrays(arrays, DEV)
    say("dropin case", ref["pre"]["M"])
    return arrays, spec, scene, ref


_refs = {}


def oracle(arrays, spec, sh="fp32", feat="fp32", cad=8, dt=False):
    k = (id(arrays), sh, feat, cad, dt)
    if k not in _refs:
        _refs[k] = oracle_frame(arrays, spec, sh_type=sh, feat_prec=feat, bg=BG,
                                check_every=cad, fp64=False, dt=dt)
    return _refs[k]


class Bufs:
    """The reference Rasterizer's per-frame buffers (example.py:56-139)."""

    def __init__(self, P, M, tiles, W, H, records=False):
        z = lambda *s, dt=torch.float32: torch.zeros(*s, device=DEV, dtype=dt)  # noqa: E731
        self.ku = z(M, dt=torch.int64)
        self.vu = z(M, dt=torch.int32)
        self.ks = z(M, dt=torch.int64)
        self.vs = z(M, dt=torch.int32)
        self.space = z(ops.get_sort_buffer_size(M), dt=torch.int8)
        self.curr = z(1, dt=torch.int32)
        self.xy = z(P, 2)
        self.rgbd = z(P, 4)
        self.con = z(P, 4)
        self.rec = z(24 * P, dt=torch.uint8) if records else None
        self.ranges = z(tiles, 2, dt=torch.int32)
        self.out = z(H, W, 3, dt=torch.uint8)


def _cam(spec):
    return (spec.width, spec.height, 16, 16, spec.position, spec.rotation, spec.focal_x,
            spec.focal_y, spec.zFar, spec.zNear)


def _bufs(case, extra=4096, records=False):
    arrays, spec, scene, ref = case
    T = ((spec.width + 15) // 16) * ((spec.height + 15) // 16)
    return Bufs(scene.num_vertex, ref["pre"]["M"] + extra, T, spec.width, spec.height, records)


def _bg():
    return torch.tensor(BG, dtype=torch.float32)


def _img(b):
    torch.cuda.synchronize()
    return b.out.cpu().numpy().copy()


def _assert_close(img, ref, what):
    m = diff_metrics(img, ref["image"])
    say(what, m)
    assert m["max_abs_diff"] <= 1, (what, m)


def _preprocess(case, b, half=False):
    arrays, spec, scene, ref = case
    fn = ops.preprocess_half_sh if half else ops.preprocess
    b.curr.zero_()
    fn(scene.position, scene.shs_half if half else scene.shs, scene.opacity, scene.cov3d,
       *_cam(spec), b.xy, b.rgbd, b.con, b.ku, b.vu, b.curr)
    torch.cuda.synchronize()
    return int(b.curr.item())


def _sorted(case, b, M):
    spec = case[1]
    ops.sort_gaussian(M, spec.width, spec.height, 16, 16, b.space, b.ku, b.vu, b.ks, b.vs)
    ops.precompute_tile_ranges(M, spec.width, spec.height, 16, 16, b.ks, b.ranges)


@pytest.mark.parametrize("half", [False, True])
def test_preprocess_ops_bit_exact(case, half):
    arrays, spec, scene, ref = case
    b = _bufs(case, extra=ref["pre"]["M"] + 4096)
    M = _preprocess(case, b, half)
    r = oracle(arrays, spec, sh="fp16" if half else "fp32")
    assert M == r["pre"]["M"]
    np.testing.assert_array_equal(b.ku[:M].cpu().numpy().view(np.uint64), r["pre"]["keys"])
    np.testing.assert_array_equal(b.vu[:M].cpu().numpy().view(np.uint32), r["pre"]["vals"])
    ids = np.unique(r["pre"]["vals"])
    for t, k in ((b.xy, "xy"), (b.rgbd, "rgbd"), (b.con, "conic")):
        np.testing.assert_array_equal(t.cpu().numpy()[ids].view(np.uint32),
                                      r["pre"][k][ids].view(np.uint32), err_msg=k)
    # a second call appends at curr_offset (the reference clears it first)
    fn = ops.preprocess_half_sh if half else ops.preprocess
    fn(scene.position, scene.shs_half if half else scene.shs, scene.opacity, scene.cov3d,
       *_cam(spec), b.xy, b.rgbd, b.con, b.ku, b.vu, b.curr)
    torch.cuda.synchronize()
    assert int(b.curr.item()) == 2 * M
    np.testing.assert_array_equal(b.ku[M:2 * M].cpu().numpy().view(np.uint64), r["pre"]["keys"])


def test_sort_ranges_ops_bit_exact(case):
    arrays, spec, scene, ref = case
    b = _bufs(case)
    M = _preprocess(case, b)
    _sorted(case, b, M)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(b.ks[:M].cpu().numpy().view(np.uint64), ref["keys_sorted"])
    np.testing.assert_array_equal(b.vs[:M].cpu().numpy().view(np.uint32), ref["vals_sorted"])
    np.testing.assert_array_equal(b.ranges.cpu().numpy(), ref["ranges"])
    ops.precompute_tile_ranges_scan(M, spec.width, spec.height, 16, 16, b.ks, b.ranges)
    torch.cuda.synchronize()
    import oracle as O
    np.testing.assert_array_equal(b.ranges.cpu().numpy(),
                                  O.tile_ranges(ref["keys_sorted"], b.ranges.shape[0], 1))
    # the sort preserves its inputs
    np.testing.assert_array_equal(b.ku[:M].cpu().numpy().view(np.uint64), ref["pre"]["keys"])


TILE_WORD = np.uint64(0xFFFFFFFF00000000)


@pytest.mark.parametrize("feat", ["fp32", "fp16"])
def test_depth_first_ops_bit_exact(case, feat):
    """ops.preprocess_es (count pass, depth sort, pairs in depth order) and
    ops.es_tile_sort with the reference's es_tile_sort signature
    (pybind.cpp:306-327): sorted values and ranges as the oracle's, keys
    holding the tile word only (sort.cu:205-215), features as preprocess
    writes them."""
    arrays, spec, scene, ref = case
    b = _bufs(case, records=True)
    M0 = ref["pre"]["M"]
    cap = b.ku.numel()
    tk = b.ku.view(torch.int32)
    prec = 0 if feat == "fp32" else 1
    b.curr.fill_(12345)  # overwritten, not appended to
    ops.preprocess_es(scene.position, scene.shs if prec == 0 else scene.shs_half, scene.opacity,
                      scene.cov3d, *_cam(spec), b.xy, b.rgbd, b.con, b.rec, prec, tk[:cap], b.vu,
                      b.curr, b.space)
    torch.cuda.synchronize()
    M = int(b.curr.item())
    assert M == M0
    ops.es_tile_sort(M, spec.width, spec.height, 16, 16, tk[:cap], b.vu, tk[cap:], b.vs, b.ks,
                     b.space, tile_ranges=b.ranges)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(b.ks[:M].cpu().numpy().view(np.uint64),
                                  ref["keys_sorted"] & TILE_WORD)
    np.testing.assert_array_equal(b.vs[:M].cpu().numpy().view(np.uint32), ref["vals_sorted"])
    import oracle as O
    np.testing.assert_array_equal(b.ranges.cpu().numpy(),
                                  O.tile_ranges(ref["keys_sorted"], b.ranges.shape[0], 1))
    if prec == 0:
        ids = np.unique(ref["pre"]["vals"])
        for t, k in ((b.xy, "xy"), (b.rgbd, "rgbd"), (b.con, "conic")):
            np.testing.assert_array_equal(t.cpu().numpy()[ids].view(np.uint32),
                                          ref["pre"][k][ids].view(np.uint32), err_msg=k)
    # the reference signature without ranges, and its bounds
    b.ks.zero_()
    ops.es_tile_sort(M, spec.width, spec.height, 16, 16, tk[:cap], b.vu, tk[cap:], b.vs, b.ks,
                     b.space)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(b.ks[:M].cpu().numpy().view(np.uint64),
                                  ref["keys_sorted"] & TILE_WORD)
    with pytest.raises(Exception):
        ops.es_tile_sort(cap + 1, spec.width, spec.height, 16, 16, tk[:cap], b.vu, tk[cap:],
                         b.vs, b.ks, b.space)


def test_gather_ops_and_gathered_render(case):
    arrays, spec, scene, ref = case
    b = _bufs(case, records=True)
    M = _preprocess(case, b)
    _sorted(case, b, M)
    dxy = torch.full((M, 2), -1.0, device=DEV)
    drgbd = torch.full((M, 4), -1.0, device=DEV)
    dcon = torch.full((M, 4), -1.0, device=DEV)
    ops.gather_features(M, b.vs, b.xy, b.rgbd, b.con, dxy, drgbd, dcon)
    idx = b.vs[:M].long()
    torch.cuda.synchronize()
    assert torch.equal(dxy, b.xy[idx]) and torch.equal(drgbd, b.rgbd[idx])
    assert torch.equal(dcon, b.con[idx])
    # FP32 features: persistent render over the gathered arrays (features by
    # pair position) equals the render through the id indirection.
    L = capi.lib()

    def persistent(prec, xy, rgbd, con, rec, gathered):
        b.out.zero_()
        capi.check(L.ps_render_16x16_persistent(
            M, None, spec.width, spec.height, prec, capi.ptr(xy), capi.ptr(rgbd), capi.ptr(con),
            capi.ptr(rec), 1 if gathered else 0, capi.ptr(b.ks), capi.ptr(b.vs),
            capi.ptr(b.ranges), 0, capi.floats3(BG), capi.ptr(b.out), capi.ptr(b.q),
            capi.stream()))
        return _img(b)

    b.q = torch.zeros(512, device=DEV, dtype=torch.int32)
    direct = persistent(capi.F32, b.xy, b.rgbd, b.con, None, False)
    gathered = persistent(capi.F32, dxy, drgbd, dcon, None, True)
    np.testing.assert_array_equal(gathered, direct)
    _assert_close(direct, ref, "f32 persistent")
    # 24-byte records: gather_records then the gathered record render.
    b.curr.zero_()
    ops.preprocess_records(scene.position, scene.shs_half, scene.opacity, scene.cov3d,
                           *_cam(spec), b.rec, capi.F16, b.ku, b.vu, b.curr)
    torch.cuda.synchronize()
    assert int(b.curr.item()) == M
    drec = torch.zeros(24 * M, device=DEV, dtype=torch.uint8)
    ops.gather_records(M, b.vs, b.rec, drec)
    torch.cuda.synchronize()
    assert torch.equal(drec.view(M, 24), b.rec.view(-1, 24)[idx])
    direct = persistent(capi.F16, None, None, None, b.rec, False)
    gathered = persistent(capi.F16, None, None, None, drec, True)
    np.testing.assert_array_equal(gathered, direct)
    _assert_close(direct, oracle(arrays, spec, sh="fp16", feat="fp16"), "f16 persistent")


def test_render_ops(case):
    arrays, spec, scene, ref = case
    b = _bufs(case, records=True)
    M = _preprocess(case, b)
    _sorted(case, b, M)
    W, H = spec.width, spec.height
    b.out.zero_()
    ops.render_16x16(M, W, H, b.xy, b.rgbd, b.con, b.ks, b.vs, _bg(), b.out)
    _assert_close(_img(b), oracle(arrays, spec, cad=2), "render_16x16")
    b.out.zero_()
    ops.render_16x16_preranges_smem_persistent_lite(M, W, H, b.xy, b.rgbd, b.con, b.vs,
                                                   b.ranges, _bg(), b.out)
    _assert_close(_img(b), ref, "lite")
    b.out.zero_()
    ops.render_16x16_preranges_smem_persistent_lite_dt(M, W, H, b.xy, b.rgbd, b.con, b.vs,
                                                      b.ranges, _bg(), b.out)
    _assert_close(_img(b), oracle(arrays, spec, dt=True), "lite_dt")
    b.out.zero_()
    ops.render_16x16_preranges_smem_persistent_lite_fused(M, W, H, b.xy, b.rgbd, b.con, b.ks,
                                                         b.vs, _bg(), b.out)
    _assert_close(_img(b), oracle(arrays, spec, cad=16), "lite_fused")
    # reduced-precision records, with and without the alpha < 1/255 skip
    for prec, name in ((capi.F16, "fp16"), (capi.BF16, "bf16")):
        b.curr.zero_()
        shs = scene.shs_half if prec == capi.F16 else scene.shs_bf16
        ops.preprocess_records(scene.position, shs, scene.opacity, scene.cov3d, *_cam(spec),
                               b.rec, prec, b.ku, b.vu, b.curr)
        for dt in (False, True):
            b.out.zero_()
            ops.render_records(M, W, H, b.rec, prec, False, b.vs, b.ranges, _bg(), b.out, dt)
            _assert_close(_img(b), oracle(arrays, spec, sh=name, feat=name, dt=dt),
                          f"render_records {name} dt={dt}")


def test_forward_ops(case):
    arrays, spec, scene, ref = case
    b = _bufs(case)
    M = ref["pre"]["M"]
    args = (*_cam(spec),)
    n = ops.forward_fused(scene.position, scene.shs, scene.opacity, scene.cov3d, *args,
                          "default", b.curr, b.space, b.ku, b.vu, b.ks, b.vs, b.xy, b.rgbd, b.con,
                          _bg(), b.out)
    assert n == M
    _assert_close(_img(b), oracle(arrays, spec, cad=2), "forward_fused")
    np.testing.assert_array_equal(b.ks[:M].cpu().numpy().view(np.uint64), ref["keys_sorted"])
    n = ops.forward_fused_e2e(scene.position, scene.shs_half, scene.opacity, scene.cov3d, *args,
                              b.curr, b.space, b.ku, b.vu, b.ks, b.vs, b.xy, b.rgbd, b.con,
                              _bg(), b.out)
    assert n == M
    _assert_close(_img(b), oracle(arrays, spec, sh="fp16", cad=16), "forward_fused_e2e")
    n = ops.forward_fused_e2e_preranges(scene.position, scene.shs_half, scene.opacity,
                                        scene.cov3d, *args, b.curr, b.space, b.ku, b.vu, b.ks,
                                        b.vs, b.xy, b.rgbd, b.con, b.ranges, _bg(), b.out)
    assert n == M
    _assert_close(_img(b), oracle(arrays, spec, sh="fp16"), "forward_fused_e2e_preranges")
    ends = b.ranges.cpu().numpy()
    np.testing.assert_array_equal(ends[:, 1] - ends[:, 0], ref["ranges"][:, 1] - ref["ranges"][:, 0])
    with pytest.raises(ValueError):
        ops.forward_fused(scene.position, scene.shs, scene.opacity, scene.cov3d, *args, "plain",
                          b.curr, b.space, b.ku, b.vu, b.ks, b.vs, b.xy, b.rgbd, b.con, _bg(),
                          b.out)
    # feature buffers that cannot hold every Gaussian are refused before any launch
    small = torch.zeros(10, 4, device=DEV)
    with pytest.raises(ValueError):
        ops.forward_fused(scene.position, scene.shs, scene.opacity, scene.cov3d, *args,
                          "default", b.curr, b.space, b.ku, b.vu, b.ks, b.vs, b.xy, small, b.con,
                          _bg(), b.out)
    # too few key slots: RuntimeError with the reference's message
    t = Bufs(scene.num_vertex, M // 2, b.ranges.shape[0], spec.width, spec.height)
    with pytest.raises(RuntimeError, match="Too many k-v pairs"):
        ops.forward_fused(scene.position, scene.shs, scene.opacity, scene.cov3d, *args,
                          "default", t.curr, t.space, t.ku, t.vu, t.ks, t.vs, t.xy, t.rgbd, t.con,
                          _bg(), t.out)
    torch.cuda.synchronize()
    assert int(t.curr.item()) == M


def test_fetch_num_rendered(case):
    arrays, spec, scene, ref = case
    b = _bufs(case)
    _preprocess(case, b)
    n = C.c_int(-1)
    capi.check(capi.lib().ps_fetch_num_rendered(capi.ptr(b.curr), C.byref(n), capi.stream()))
    assert n.value == ref["pre"]["M"]


VARIANTS = {
    # variant: (SH precision follows feature_dtype?, cadence, dt)
    "default": (False, 2, False),
    "preranges_smem_persistent_lite_e2e": (True, 16, False),
    "preranges_smem_persistent_lite_e2e_preranges": (True, 8, False),
    "preranges_smem_persistent_lite": (True, 8, False),
    "preranges_smem_persistent_lite_dt": (True, 8, True),
}


@pytest.mark.parametrize("feature_dtype", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("variant", list(VARIANTS))
def test_rasterizer_forward(case, variant, feature_dtype):
    arrays, spec, scene, ref = case
    M = ref["pre"]["M"]
    T = ((spec.width + 15) // 16) * ((spec.height + 15) // 16)
    rast = polysplat.Rasterizer(scene, MAX_NUM_RENDERED=M + 4096, MAX_NUM_TILES=T,
                                feature_dtype=feature_dtype)
    half, cad, dt = VARIANTS[variant]
    sh = feature_dtype if half and feature_dtype != "fp32" else ("fp16" if half else "fp32")
    feat = feature_dtype if variant != "default" else "fp32"
    r = oracle(arrays, spec, sh=sh, feat=feat, cad=cad, dt=dt)
    bg = _bg()
    img = rast.forward(scene, spec, bg, render_variant=variant)
    plain = img.cpu().numpy().copy()
    _assert_close(plain, r, f"{variant} {feature_dtype}")
    # the output buffer is reused at the same resolution (example.py:141-144)
    out2, stats = rast.forward(scene, spec, bg, render_variant=variant, return_stats=True)
    assert out2.data_ptr() == img.data_ptr()
    np.testing.assert_array_equal(out2.cpu().numpy(), plain)
    assert stats["num_rendered"] == M and stats["requested_render_variant"] == variant
    for k in ("render_variant", "resolved_block_x", "resolved_block_y", "preprocess_ms",
              "sort_ms", "render_ms"):
        assert k in stats
    # per-stage timing: the tuple form (example.py:810-824)
    res = rast.forward(scene, spec, bg, render_variant=variant, measure_preprocess=True,
                       measure_sort=True, measure_render=True)
    assert len(res) == 4 and all(isinstance(x, float) and x > 0 for x in res[1:])
    timed = res[0].cpu().numpy()
    if variant == "preranges_smem_persistent_lite_e2e":
        # timing runs the staged lite path (example.py:362-368): cadence 8
        _assert_close(timed, oracle(arrays, spec, sh=sh, feat=feat, cad=8), "e2e timed")
    else:
        np.testing.assert_array_equal(timed, plain, err_msg=f"{variant}: timing changed the image")
    _, s2 = rast.forward(scene, spec, bg, render_variant=variant, measure_render=True,
                         return_stats=True)
    assert s2["render_ms"] > 0 and s2["preprocess_ms"] is None


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_rasterizer_depth_first(case, variant):
    """use_es=True (example.py:246-291, 494-576) gives the same images as the
    default pipeline in every variant and return form; "auto" follows the
    previous frame's pair count; enable_es=False never takes it."""
    arrays, spec, scene, ref = case
    M = ref["pre"]["M"]
    T = ((spec.width + 15) // 16) * ((spec.height + 15) // 16)
    rast = polysplat.Rasterizer(scene, MAX_NUM_RENDERED=M + 4096, MAX_NUM_TILES=T,
                                feature_dtype="fp16")
    bg = _bg()
    base = rast.forward(scene, spec, bg, render_variant=variant, use_es=False).cpu().numpy().copy()
    img, stats = rast.forward(scene, spec, bg, render_variant=variant, use_es=True,
                              return_stats=True)
    assert stats["used_es"] is True and stats["num_rendered"] == M
    np.testing.assert_array_equal(img.cpu().numpy(), base)
    np.testing.assert_array_equal(rast.gaussian_values_sorted[:M].cpu().numpy().view(np.uint32),
                                  ref["vals_sorted"])
    res = rast.forward(scene, spec, bg, render_variant=variant, use_es=True,
                       measure_preprocess=True, measure_sort=True, measure_render=True)
    assert len(res) == 4 and all(isinstance(x, float) and x > 0 for x in res[1:])
    timed_base = rast.forward(scene, spec, bg, render_variant=variant, use_es=False,
                              measure_render=True)[0].cpu().numpy().copy()
    np.testing.assert_array_equal(res[0].cpu().numpy(), timed_base)
    # "auto": never on a first frame, then by the last pair count
    rast2 = polysplat.Rasterizer(scene, MAX_NUM_RENDERED=M + 4096, MAX_NUM_TILES=T,
                                 feature_dtype="fp16")
    rast2._es_threshold_M, rast2._es_ratio_thresh = 0, 0.0
    _, s1 = rast2.forward(scene, spec, bg, render_variant=variant, return_stats=True)
    _, s2 = rast2.forward(scene, spec, bg, render_variant=variant, return_stats=True)
    assert s1["used_es"] is False and s2["used_es"] is True
    rast2._es_ratio_thresh = 1e9
    _, s3 = rast2.forward(scene, spec, bg, render_variant=variant, return_stats=True)
    assert s3["used_es"] is False
    rast3 = polysplat.Rasterizer(scene, MAX_NUM_RENDERED=M + 4096, MAX_NUM_TILES=T,
                                 enable_es=False, feature_dtype="fp16")
    _, s4 = rast3.forward(scene, spec, bg, render_variant=variant, use_es=True,
                          return_stats=True)
    assert s4["used_es"] is False


def test_rasterizer_auto_and_unsupported(case):
    arrays, spec, scene, ref = case
    T = ((spec.width + 15) // 16) * ((spec.height + 15) // 16)
    rast = polysplat.Rasterizer(scene, MAX_NUM_RENDERED=ref["pre"]["M"] + 4096, MAX_NUM_TILES=T)
    img, stats = rast.forward(scene, spec, _bg(), render_variant="auto", return_stats=True)
    assert stats["requested_render_variant"] == "auto"
    assert stats["num_rendered"] == ref["pre"]["M"]
    for bad in ("topk_smem_v3", "reordered_persistent_hilbert", "split"):
        with pytest.raises(ValueError):
            rast.forward(scene, spec, _bg(), render_variant=bad)
    with pytest.raises(ValueError):
        rast.forward(scene, spec, _bg(), block_x=32)
