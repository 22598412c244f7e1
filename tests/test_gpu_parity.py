"""GPU parity: every stage of the HIP path against the CPU oracle, through
the C ABI.

Bars: keys, values, sort order and tile ranges bit-exact; per-Gaussian
features bit-exact for every Gaussian that emits a key; images within one
8-bit level (1/255) of the FP32 oracle run with the same early-exit cadence,
and PSNR >= 50 dB against the FP64 evaluation of the same blend.
"""
import numpy as np
import pytest
import torch

import oracle as O
from polysplat_amd import scenes
from polysplat_amd.metrics import diff_metrics
from tests import capi
from tests.gpu_helpers import Gpu, oracle_frame, say

pytestmark = pytest.mark.gpu

BG = (0.0, 0.0, 0.0)
TILE_WORD = np.uint64(0xFFFFFFFF00000000)  # sorted keys of the depth-first pipeline


def _mixed_scene(P=60_000, seed=3):
    # Mixture with large splats close to the camera: exercises the single,
    # fragmented and massive (>= 64 candidate tiles) emission routes.
    return scenes.make_scene(P, 8.0, [(0.85, -4.0, 0.5), (0.12, -2.2, 0.4), (0.03, -0.7, 0.3)],
                             seed=seed)


CASES = {
    "c1": (lambda: scenes.make_config_scene("c1", seed=1), lambda: scenes.config_camera("c1")),
    "mixed_odd": (_mixed_scene, lambda: scenes.make_camera(1000, 700, 8.0)),
    "mixed_near": (lambda: _mixed_scene(40_000, 9), lambda: scenes.make_camera(640, 480, 8.0, distance=7.0)),
}


@pytest.fixture(scope="module", params=list(CASES))
def case(request, gpu):
    name = request.param
    scene = CASES[name][0]()
    spec = CASES[name][1]()
    ref = oracle_frame(scene, spec, bg=BG, check_every=8)
    say("case", name, "P", scene.num_vertex, "M", ref["pre"]["M"], spec.width, spec.height)
    return name, scene, spec, ref


def _cap(ref):
    return ref["pre"]["M"] + 4096


def _referenced(ref):
    return np.unique(ref["pre"]["vals"])


def test_preprocess_keys_values_bit_exact(case):
    name, scene, spec, ref = case
    g = Gpu(scene, spec, capacity=_cap(ref))
    say("preprocess", name)
    M = g.preprocess()
    assert M == ref["pre"]["M"]
    np.testing.assert_array_equal(g.ku[:M].cpu().numpy().view(np.uint64), ref["pre"]["keys"])
    np.testing.assert_array_equal(g.vu[:M].cpu().numpy().view(np.uint32), ref["pre"]["vals"])
    ids = _referenced(ref)
    for t, r in ((g.xy, "xy"), (g.rgbd, "rgbd"), (g.conic, "conic")):
        got = t.cpu().numpy()[ids]
        np.testing.assert_array_equal(got.view(np.uint32), ref["pre"][r][ids].view(np.uint32),
                                      err_msg=r)
    assert g.status_word() == 0


def test_routes_are_exercised(case):
    name, scene, spec, ref = case
    # candidate-rect areas decide the route; check the mixed scenes hit all three
    counts = ref["pre"]["counts"]
    if name.startswith("mixed"):
        assert (counts == 1).sum() > 0 and ((counts > 1) & (counts < 64)).sum() > 0
        assert (counts >= 64).sum() > 0


def test_sort_and_ranges_bit_exact(case):
    name, scene, spec, ref = case
    g = Gpu(scene, spec, capacity=_cap(ref))
    g.preprocess()
    say("sort", name, g.M)
    g.sort()
    M = g.M
    np.testing.assert_array_equal(g.ks[:M].cpu().numpy().view(np.uint64), ref["keys_sorted"])
    np.testing.assert_array_equal(g.vs[:M].cpu().numpy().view(np.uint32), ref["vals_sorted"])
    say("ranges", name)
    np.testing.assert_array_equal(g.ranges_bsearch(), ref["ranges"])
    T = g.tiles
    np.testing.assert_array_equal(g.ranges_scan(), O.tile_ranges(ref["keys_sorted"], T, 1))


def test_ranges_scan_any_alignment_and_count(case):
    # The boundary scan reads two keys per 16-byte load when the keys are
    # 16-byte aligned and one key at a time otherwise: both forms, with the
    # pair count from the host and from device memory, even and odd.
    name, scene, spec, ref = case
    g = Gpu(scene, spec, capacity=_cap(ref))
    g.preprocess()
    g.sort()
    M, T = g.M, g.tiles
    say("ranges alignment", name, M)
    buf = torch.zeros(M + 2, dtype=torch.int64, device=g.ks.device)
    buf[1:M + 1] = g.ks[:M]
    views = {"aligned": g.ks, "unaligned": buf[1:]}
    assert views["aligned"].data_ptr() % 16 == 0 and views["unaligned"].data_ptr() % 16 == 8
    for n in sorted({M, max(M - 1, 0)}):
        want = O.tile_ranges(ref["keys_sorted"][:n], T, 1)
        d_n = torch.tensor([n], dtype=torch.int32, device=g.ks.device)
        for kind, keys in views.items():
            for dev_count in (False, True):
                out = torch.full((T, 2), -5, dtype=torch.int32, device=g.ks.device)
                capi.check(capi.lib().ps_precompute_tile_ranges_scan(
                    0 if dev_count else n, capi.ptr(d_n) if dev_count else None, M,
                    spec.width, spec.height, 16, 16, capi.ptr(keys), capi.ptr(out),
                    capi.stream()))
                torch.cuda.synchronize()
                np.testing.assert_array_equal(out.cpu().numpy(), want,
                                              err_msg=f"{kind} n={n} device count={dev_count}")


@pytest.mark.parametrize("mode", ["plain", "persistent", "persistent_tiles", "fused", "dt"])
def test_render_fp32_parity(case, mode, monkeypatch):
    # These frames have fewer tiles than the persistent renderer's waves, so
    # it takes (tile, quadrant) items; "persistent_tiles" forces whole tiles.
    if mode == "persistent_tiles":
        monkeypatch.setenv("POLYSPLAT_RENDER_QUADS", "0")
    name, scene, spec, ref = case
    g = Gpu(scene, spec, capacity=_cap(ref))
    g.preprocess()
    g.sort()
    g.ranges_bsearch()
    say("render", name, mode)
    if mode == "plain":
        img = g.render_plain(BG)
        cad, dt = 2, False
    elif mode in ("persistent", "persistent_tiles"):
        img = g.render_persistent(BG, use_ranges=True)
        cad, dt = 8, False
    elif mode == "fused":
        img = g.render_persistent(BG, use_ranges=False)
        cad, dt = 16, False
    else:
        img = g.render_persistent(BG, use_ranges=True, dt=True)
        cad, dt = 8, True
    r = ref if (cad == 8 and not dt) else oracle_frame(scene, spec, bg=BG, check_every=cad,
                                                       fp64=False, dt=dt)
    m32 = diff_metrics(img, r["image"])
    m64 = diff_metrics(img, ref["image64"])
    say(mode, "vs fp32", m32, "vs fp64", m64)
    assert m32["max_abs_diff"] <= 1, m32
    if not dt:
        assert m64["psnr_db"] >= 50.0, m64
        assert m64["max_abs_diff"] <= 1, m64


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_render_reduced_precision_features(case, prec):
    name, scene, spec, ref = case
    g = Gpu(scene, spec, sh_type=prec, feat_prec=prec, capacity=_cap(ref))
    g.preprocess()
    g.sort()
    g.ranges_bsearch()
    say("render records", name, prec)
    img = g.render_persistent(BG, use_ranges=True)
    r = oracle_frame(scene, spec, sh_type=prec, feat_prec=prec, bg=BG, check_every=8, fp64=True)
    m32 = diff_metrics(img, r["image"])
    m64 = diff_metrics(img, ref["image64"])  # FP64 of the FP32 scene (SH unrounded)
    say(prec, "vs fp32-quantised", m32, "vs fp64", m64)
    assert m32["max_abs_diff"] <= 1, m32
    assert m64["psnr_db"] >= 50.0, m64


def test_forward_pipeline_matches_stages(case):
    name, scene, spec, ref = case
    g = Gpu(scene, spec, capacity=_cap(ref))
    for variant, cad in ((capi.V_DEFAULT, 2), (capi.V_LITE, 8), (capi.V_FUSED, 16)):
        say("forward", name, variant)
        st, n, img = g.forward(BG, variant)
        capi.check(st)
        assert n == ref["pre"]["M"]
        np.testing.assert_array_equal(g.ks[:n].cpu().numpy().view(np.uint64), ref["keys_sorted"])
        r = ref if cad == 8 else oracle_frame(scene, spec, bg=BG, check_every=cad, fp64=False)
        m = diff_metrics(img, r["image"])
        assert m["max_abs_diff"] <= 1, (variant, m)
    # repeated frames give identical images (deterministic pipeline)
    st, n, img1 = g.forward(BG, capi.V_LITE)
    st, n, img2 = g.forward(BG, capi.V_LITE)
    np.testing.assert_array_equal(img1, img2)


@pytest.mark.parametrize("variant", [capi.V_DEFAULT, capi.V_LITE, capi.V_FUSED])
def test_depth_first_pipeline(case, variant):
    """Depth-first key pipeline (Gaussians sorted by depth, pairs emitted in
    depth order as 32-bit tile ids, stable tile sort, rebuilt keys): the
    oracle's sorted values and ranges, its sorted keys' tile words (the
    rebuilt keys carry the tile id only, sort.cu:205-215), and the same image
    bytes as the default pipeline of the variant."""
    name, scene, spec, ref = case
    g = Gpu(scene, spec, capacity=_cap(ref))
    say("depth first", name, variant)
    st, n, img0 = g.forward(BG, variant)
    capi.check(st)
    st, n, img = g.forward(BG, variant | capi.DEPTH_FIRST)
    capi.check(st)
    M = ref["pre"]["M"]
    assert n == M
    np.testing.assert_array_equal(g.ks[:M].cpu().numpy().view(np.uint64),
                                  ref["keys_sorted"] & TILE_WORD)
    np.testing.assert_array_equal(g.vs[:M].cpu().numpy().view(np.uint32), ref["vals_sorted"])
    if variant == capi.V_LITE:
        np.testing.assert_array_equal(g.ranges.cpu().numpy(),
                                      O.tile_ranges(ref["keys_sorted"], g.tiles, 1))
    np.testing.assert_array_equal(img, img0)


def test_depth_first_overflow(gpu):
    scene = scenes.make_config_scene("c1", seed=2, P=20_000)
    spec = scenes.config_camera("c1")
    ref = oracle_frame(scene, spec, fp64=False)
    M = ref["pre"]["M"]
    g = Gpu(scene, spec, capacity=M - 1)
    say("depth first overflow", M)
    st, n, _ = g.forward(BG, capi.V_LITE | capi.DEPTH_FIRST)
    assert st == capi.PS_ERR_OVERFLOW and n == M
    g = Gpu(scene, spec, capacity=M)
    st, n, img = g.forward(BG, capi.V_LITE | capi.DEPTH_FIRST)
    capi.check(st)
    assert n == M
    np.testing.assert_array_equal(g.ks[:M].cpu().numpy().view(np.uint64),
                                  ref["keys_sorted"] & TILE_WORD)
    np.testing.assert_array_equal(g.vs[:M].cpu().numpy().view(np.uint32), ref["vals_sorted"])


@pytest.mark.parametrize("block", [(8, 8), (32, 32), (12, 20)])
def test_preprocess_tile_sizes(gpu, block):
    """ps_preprocess with tile sizes other than 16x16 against the oracle run
    with the same tile size (keys and values bit-exact)."""
    scene = _mixed_scene(20_000, 5)
    spec = scenes.make_camera(1000, 700, 8.0)
    bx, by = block
    cam_o = O.camera(spec.width, spec.height, spec.position.numpy(), spec.rotation.numpy(),
                     spec.focal_x, spec.focal_y, spec.zFar, spec.zNear, bx, by)
    ref = O.preprocess(scene.position.numpy(), scene.shs.numpy(), scene.opacity.numpy(),
                       scene.cov3d.numpy(), cam_o)
    g = Gpu(scene, spec, capacity=ref["M"] + 64)
    g.cam.block_x, g.cam.block_y = bx, by
    say("tile size", block, ref["M"])
    assert g.preprocess() == ref["M"]
    np.testing.assert_array_equal(g.ku[:ref["M"]].cpu().numpy().view(np.uint64), ref["keys"])
    np.testing.assert_array_equal(g.vu[:ref["M"]].cpu().numpy().view(np.uint32), ref["vals"])


def test_overflow_is_reported_and_nothing_is_written_past_capacity(gpu):
    scene = scenes.make_config_scene("c1", seed=2, P=20_000)
    spec = scenes.config_camera("c1")
    ref = oracle_frame(scene, spec, fp64=False)
    M = ref["pre"]["M"]
    g = Gpu(scene, spec, capacity=M + 4096)
    say("overflow", M)
    for cap in (M - 1, M // 3):
        g.cap = cap  # the key arrays stay M + 4096 long: the tail is a guard
        g.ku.fill_(-7)
        g.vu.fill_(-7)
        assert g.preprocess() == M  # the count still reports what is needed
        assert g.status_word() & 1
        assert (g.ku[cap:] == -7).all().item() and (g.vu[cap:] == -7).all().item()
        # below capacity: only whole workgroups' canonical pairs
        ku = g.ku[:cap].cpu().numpy()
        done = ku != -7
        np.testing.assert_array_equal(ku[done].view(np.uint64), ref["pre"]["keys"][:cap][done])
        st, n, _ = g.forward(BG, capi.V_LITE)
        assert st == capi.PS_ERR_OVERFLOW and n == M
    # exactly enough room works
    g2 = Gpu(scene, spec, capacity=M)
    assert g2.preprocess() == M and g2.status_word() == 0
    np.testing.assert_array_equal(g2.ku.cpu().numpy().view(np.uint64), ref["pre"]["keys"])


def test_preprocess_appends_at_curr_offset(gpu):
    scene = scenes.make_config_scene("c1", seed=4, P=5_000)
    spec = scenes.config_camera("c1")
    g = Gpu(scene, spec)
    M = g.preprocess()
    first = g.ku[:M].clone()
    # second call without clearing: pairs land after the first batch
    L = capi.lib()
    import ctypes as C
    capi.check(L.ps_preprocess(g.P, capi.ptr(g.pos), capi.ptr(g.shs), 0, capi.ptr(g.opac),
                               capi.ptr(g.cov), C.byref(g.cam), capi.ptr(g.xy), capi.ptr(g.rgbd),
                               capi.ptr(g.conic), None, 1, capi.ptr(g.ku), capi.ptr(g.vu), g.cap,
                               capi.ptr(g.curr), capi.ptr(g.ws), g.ws_bytes, capi.stream()))
    torch.cuda.synchronize()
    assert int(g.curr.item()) == 2 * M
    assert torch.equal(g.ku[M:2 * M], first) and torch.equal(g.ku[:M], first)


@pytest.mark.parametrize("depth_first", [False, True])
@pytest.mark.parametrize("P", [0, 1, 2])
def test_tiny_and_empty_scenes(gpu, P, depth_first):
    scene = scenes.make_scene(max(P, 1), 2.0, [(1.0, -2.0, 0.3)], seed=11)
    if P == 0:
        scene = scenes.SceneArrays(scene.position[:0], scene.shs[:0], scene.opacity[:0],
                                   scene.cov3d[:0])
    spec = scenes.make_camera(50, 30, 2.0)
    g = Gpu(scene, spec)
    say("tiny", P, depth_first)
    st, n, img = g.forward((0.2, 0.4, 0.6), capi.V_LITE | (capi.DEPTH_FIRST if depth_first else 0))
    capi.check(st)
    ref = oracle_frame(scene, spec, bg=(0.2, 0.4, 0.6), fp64=False)
    assert n == ref["pre"]["M"]
    assert diff_metrics(img, ref["image"])["max_abs_diff"] <= 1


def test_all_gaussians_behind_camera(gpu):
    scene = scenes.make_scene(3000, 2.0, [(1.0, -3.0, 0.3)], seed=12)
    scene.position[:, 2] -= 50.0  # everything behind the camera at z = -3
    spec = scenes.make_camera(64, 64, 2.0)
    g = Gpu(scene, spec)
    st, n, img = g.forward((1.0, 0.5, 0.0), capi.V_FUSED)
    capi.check(st)
    assert n == 0
    assert (img[..., 0] == 255).all() and (img[..., 1] == 127).all() and (img[..., 2] == 0).all()


def test_c2_full_size_bit_exact(gpu):
    """Config 2 at full size: 3M Gaussians, 1920x1080, FP32."""
    scene = scenes.make_config_scene("c2", seed=1)
    spec = scenes.config_camera("c2")
    say("c2 oracle")
    ref = oracle_frame(scene, spec, bg=BG, check_every=8, fp64=True)
    M = ref["pre"]["M"]
    g = Gpu(scene, spec, capacity=M + 4096)
    say("c2 gpu", M)
    assert g.preprocess() == M
    np.testing.assert_array_equal(g.ku[:M].cpu().numpy().view(np.uint64), ref["pre"]["keys"])
    np.testing.assert_array_equal(g.vu[:M].cpu().numpy().view(np.uint32), ref["pre"]["vals"])
    g.sort()
    np.testing.assert_array_equal(g.ks[:M].cpu().numpy().view(np.uint64), ref["keys_sorted"])
    np.testing.assert_array_equal(g.vs[:M].cpu().numpy().view(np.uint32), ref["vals_sorted"])
    np.testing.assert_array_equal(g.ranges_bsearch(), ref["ranges"])
    img = g.render_persistent(BG, use_ranges=True)
    m32 = diff_metrics(img, ref["image"])
    m64 = diff_metrics(img, ref["image64"])
    say("c2 vs fp32", m32, "vs fp64", m64)
    assert m32["max_abs_diff"] <= 1, m32
    assert m64["max_abs_diff"] <= 1 and m64["psnr_db"] >= 50.0, m64


def test_forward_workspace_reused_across_scene_sizes(gpu):
    """One workspace for a large then a small scene: the persistent queue
    words must not pick up the larger scene's preprocess data."""
    big_s = scenes.make_config_scene("c1", seed=21, P=30_000)
    small_s = scenes.make_config_scene("c1", seed=23, P=12_000)
    big = Gpu(big_s, scenes.config_camera("c1"), sh_type="fp16", feat_prec="fp16",
              capacity=400_000)
    small = Gpu(small_s, scenes.make_camera(333, 250, 10.0, distance=9.0), sh_type="fp16",
                feat_prec="fp16", capacity=400_000)
    st, _, fresh = small.forward(BG, capi.V_LITE)
    capi.check(st)
    shared = torch.zeros(max(big.ws_bytes, small.ws_bytes), device=big.dev, dtype=torch.uint8)
    for _ in range(2):
        say("ws reuse big")
        st, _, _ = big.forward(BG, capi.V_LITE, ws=shared)
        capi.check(st)
        say("ws reuse small")
        st, _, img = small.forward(BG, capi.V_LITE, ws=shared)
        capi.check(st)
        np.testing.assert_array_equal(img, fresh)


@pytest.mark.parametrize("name", ["c1_small", "mixed_small"])
def test_golden_fixtures(gpu, name):
    """The committed fixtures (tests/golden/): keys, order and both range
    conventions bit-exact, every image within one 8-bit level."""
    from tests.test_golden import arrays_of, load, spec_of

    d = load(name)
    M = int(d["num_rendered"])
    arrays, spec, bg = arrays_of(d), spec_of(d), tuple(d["bg"])
    g = Gpu(arrays, spec, capacity=M + 64)
    assert g.preprocess() == M
    np.testing.assert_array_equal(g.ku[:M].cpu().numpy().view(np.uint64), d["keys"])
    np.testing.assert_array_equal(g.vu[:M].cpu().numpy().view(np.uint32), d["values"])
    g.sort()
    np.testing.assert_array_equal(g.ks[:M].cpu().numpy().view(np.uint64), d["keys_sorted"])
    np.testing.assert_array_equal(g.vs[:M].cpu().numpy().view(np.uint32), d["values_sorted"])
    np.testing.assert_array_equal(g.ranges_scan(), d["ranges_scan"])
    np.testing.assert_array_equal(g.ranges_bsearch(), d["ranges"])
    for img, ref in ((g.render_plain(bg), "image_check2"),
                     (g.render_persistent(bg, use_ranges=True), "image_check8")):
        m = diff_metrics(img, d[ref])
        assert m["max_abs_diff"] <= 1, (ref, m)
        assert diff_metrics(img, d["image_fp64"])["max_abs_diff"] <= 1
    h = Gpu(arrays, spec, sh_type="fp16", feat_prec="fp16", capacity=M + 64)
    st, n, img = h.forward(bg, capi.V_LITE)
    capi.check(st)
    assert n == M
    assert diff_metrics(img, d["image_fp16_check8"])["max_abs_diff"] <= 1


def test_posted_frames_and_their_counts(gpu):
    """ps_forward(sync_count=2): frames enqueued back to back, counts waited
    for afterwards; same images and counts as synchronous frames, overflow
    at the wait, at most 64 outstanding tickets per thread."""
    import ctypes as C

    scene = scenes.make_config_scene("c1", seed=6, P=20_000)
    spec = scenes.config_camera("c1")
    ref = oracle_frame(scene, spec, fp64=False)
    M = ref["pre"]["M"]
    g = Gpu(scene, spec, sh_type="fp16", feat_prec="fp16", capacity=M + 64)
    st, n, want = g.forward(BG, capi.V_LITE)
    capi.check(st)
    L = capi.lib()
    tickets = []
    for _ in range(64):
        st, t, img = g.forward(BG, capi.V_LITE, sync=2)
        capi.check(st)
        tickets.append(t)
        np.testing.assert_array_equal(img, want)
    assert sorted(tickets) == list(range(64))
    # a 65th outstanding ticket is refused before anything runs (the helper
    # cleared the image first, and it stays cleared)
    st, _, img = g.forward(BG, capi.V_LITE, sync=2)
    assert st == capi.PS_ERR_INVALID and not img.any()
    torch.cuda.synchronize()
    out = C.c_int(-1)
    for t in tickets:
        assert L.ps_forward_query(t) == 1
        capi.check(L.ps_forward_wait(t, C.byref(out)))
        assert out.value == M
    g.cap = M - 1
    st, t, _ = g.forward(BG, capi.V_LITE, sync=2)
    capi.check(st)
    assert L.ps_forward_wait(t, C.byref(out)) == capi.PS_ERR_OVERFLOW and out.value == M
    assert L.ps_forward_wait(t, C.byref(out)) == capi.PS_ERR_INVALID
