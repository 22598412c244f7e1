"""GPU parity at the benchmark workloads, at full size.

* 6M Gaussians at 3840x2160, log-scales from the 90/9/1 mixture
  N(-5, 0.4) / N(-3, 0.4) / N(-1, 0.3): keys, values, sort order and both
  range conventions bit-exact; the plain FP32 path and the persistent path
  with FP16 and with BF16 SH and features within one 8-bit level of the
  oracle run at the same precision, and PSNR >= 50 dB against the FP64
  evaluation of the FP32 scene.
* The 56M-Gaussian city (64x64 blocks on a 2 km plane, FP16 SH) at
  9216x6912 through the persistent renderer: keys, order and ranges
  bit-exact, the image within one level of the oracle, PSNR >= 50 dB
  against FP64.
* The serving stream: four resident scenes (1M, 3M, 10M, 30M Gaussians,
  FP16 SH and features) and requests at all four resolutions
  (1280x720 .. 7680x4320) through the serving loop; every image
  byte-identical to the same request rendered alone, and one request per
  scene against the oracle.

Scenes are generated on the device; the oracle reads host copies.
"""
import math

import numpy as np
import pytest
import torch

import oracle as O
from polysplat_amd import scenes
from polysplat_amd import serving as SV
from polysplat_amd.metrics import diff_metrics
from tests import capi
from tests.gpu_helpers import Gpu, oracle_frame, oracle_sh, say

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
BG = (0.0, 0.0, 0.0)
DEPTH_FIRST = capi.DEPTH_FIRST
TILE_WORD = np.uint64(0xFFFFFFFF00000000)  # sorted keys of the depth-first pipeline


def _host_view(a):
    """Positions, opacities and covariances on the host for the oracle; SH
    stay on the device (the oracle helper copies them at the precision it
    reads)."""
    return scenes.SceneArrays(a.position.cpu(), a.shs, a.opacity.cpu(), a.cov3d.cpu())


def _keys(t, n):
    return t[:n].cpu().numpy().view(np.uint64)


def _vals(t, n):
    return t[:n].cpu().numpy().view(np.uint32)


def _quality(img, ref, ref64, what):
    m = diff_metrics(img, ref["image"])
    m64 = diff_metrics(img, ref64)
    say(what, "vs same-precision oracle", m, "vs fp64", m64)
    assert m["max_abs_diff"] <= 1, (what, m)
    assert m64["psnr_db"] >= 50.0, (what, m64)
    return m64


# ------------------------------------------------------------------ C3
@pytest.fixture(scope="module")
def c3(gpu):
    arrays = scenes.make_config_scene("c3", seed=1, device=DEV)
    spec = scenes.config_camera("c3")
    host = _host_view(arrays)
    say("c3 oracle")
    ref = oracle_frame(host, spec, bg=BG, check_every=8, fp64=True)
    say("c3", "M", ref["pre"]["M"])
    return arrays, host, spec, ref


def test_c3_keys_order_ranges_features_bit_exact(c3):
    arrays, host, spec, ref = c3
    M = ref["pre"]["M"]
    g = Gpu(arrays, spec, capacity=M + 4096)
    assert g.preprocess() == M
    np.testing.assert_array_equal(_keys(g.ku, M), ref["pre"]["keys"])
    np.testing.assert_array_equal(_vals(g.vu, M), ref["pre"]["vals"])
    ids = np.unique(ref["pre"]["vals"])
    for t, k in ((g.xy, "xy"), (g.rgbd, "rgbd"), (g.conic, "conic")):
        np.testing.assert_array_equal(t.cpu().numpy()[ids].view(np.uint32),
                                      ref["pre"][k][ids].view(np.uint32), err_msg=k)
    # every emission route is exercised at this size
    counts = ref["pre"]["counts"]
    assert (counts == 1).sum() > 0 and ((counts > 1) & (counts < 64)).sum() > 0
    assert (counts >= 64).sum() > 0
    g.sort()
    np.testing.assert_array_equal(_keys(g.ks, M), ref["keys_sorted"])
    np.testing.assert_array_equal(_vals(g.vs, M), ref["vals_sorted"])
    np.testing.assert_array_equal(g.ranges_bsearch(), ref["ranges"])
    np.testing.assert_array_equal(g.ranges_scan(), O.tile_ranges(ref["keys_sorted"], g.tiles, 1))


@pytest.mark.parametrize("depth_first", [False, True])
def test_c3_fp32_plain_path(c3, depth_first):
    arrays, host, spec, ref = c3
    M = ref["pre"]["M"]
    g = Gpu(arrays, spec, capacity=M + 4096)
    st, n, img = g.forward(BG, capi.V_DEFAULT | (DEPTH_FIRST if depth_first else 0))
    capi.check(st)
    assert n == M
    np.testing.assert_array_equal(_keys(g.ks, M),
                                  ref["keys_sorted"] & (TILE_WORD if depth_first else ~np.uint64(0)))
    np.testing.assert_array_equal(_vals(g.vs, M), ref["vals_sorted"])
    r2 = oracle_frame(host, spec, bg=BG, check_every=2, fp64=False)
    m64 = _quality(img, r2, ref["image64"], "c3 fp32 plain")
    assert m64["max_abs_diff"] <= 1


@pytest.mark.parametrize("prec,depth_first", [("fp16", False), ("bf16", False),
                                               ("fp16", True)])
def test_c3_reduced_precision_persistent_path(c3, prec, depth_first):
    arrays, host, spec, ref = c3
    M = ref["pre"]["M"]
    g = Gpu(arrays, spec, sh_type=prec, feat_prec=prec, capacity=M + 4096)
    st, n, img = g.forward(BG, capi.V_LITE | (DEPTH_FIRST if depth_first else 0))
    capi.check(st)
    assert n == M
    # the SH and feature precision change colours only: keys and order are
    # those of the FP32 scene
    np.testing.assert_array_equal(_keys(g.ks, M),
                                  ref["keys_sorted"] & (TILE_WORD if depth_first else ~np.uint64(0)))
    np.testing.assert_array_equal(_vals(g.vs, M), ref["vals_sorted"])
    np.testing.assert_array_equal(g.ranges.cpu().numpy(),
                                  O.tile_ranges(ref["keys_sorted"], g.tiles, 1))
    rq = oracle_frame(host, spec, sh_type=prec, feat_prec=prec, bg=BG, check_every=8, fp64=False)
    m64 = _quality(img, rq, ref["image64"], f"c3 {prec} persistent")
    print(f"c3 {prec} features: PSNR {m64['psnr_db']:.2f} dB vs FP64, max-abs "
          f"{m64['max_abs_diff']}", flush=True)


def _check_reordered(g, img, M, prec, what):
    """Tile orders of the frame bit-exact against the oracle, and the
    persistent render over each (packed or plain order) identical to the
    unordered render."""
    L = capi.lib()
    T, gx = g.tiles, (g.spec.width + 15) // 16
    rg = g.ranges.cpu().numpy()
    tb = L.ps_get_tile_order_sort_temp_size(T)
    temp = torch.zeros(tb, dtype=torch.uint8, device=DEV)
    order, counts, ids = (torch.zeros(T, dtype=torch.int32, device=DEV) for _ in range(3))
    desc = torch.zeros(T, 4, dtype=torch.int32, device=DEV)
    for zig, bucket in ((0, 0), (264, 0), (792, 1280)):
        capi.check(L.ps_compute_tile_order_packed(
            capi.ptr(g.ranges), T, gx, capi.ptr(desc), capi.ptr(order), capi.ptr(counts),
            capi.ptr(ids), capi.ptr(temp), tb, zig, bucket, capi.stream()))
        torch.cuda.synchronize()
        o_order, o_keys, o_desc = O.tile_order(rg, gx, zig, bucket)
        np.testing.assert_array_equal(order.cpu().numpy().view(np.uint32), o_order)
        np.testing.assert_array_equal(counts.cpu().numpy().view(np.uint32), o_keys)
        np.testing.assert_array_equal(desc.cpu().numpy(), o_desc)
        for use_desc in (True, False):
            g.out.fill_(9)
            capi.check(L.ps_render_16x16_reordered_persistent(
                M, None, g.spec.width, g.spec.height, capi.PREC_IDS[prec], capi.ptr(g.xy),
                capi.ptr(g.rgbd), capi.ptr(g.conic), capi.ptr(g.records), capi.ptr(g.vs),
                capi.ptr(g.ranges), capi.ptr(order), capi.ptr(desc) if use_desc else None, 0,
                capi.floats3(BG), capi.ptr(g.out), capi.ptr(g.queue), capi.stream()))
            torch.cuda.synchronize()
            np.testing.assert_array_equal(g.out.cpu().numpy(), img,
                                          err_msg=f"{what} reordered {zig} {bucket} {use_desc}")
    say(what, "tile orders and reordered renders ok")


def test_c3_tile_orders_and_reordered_render(c3):
    arrays, host, spec, ref = c3
    M = ref["pre"]["M"]
    g = Gpu(arrays, spec, sh_type="fp16", feat_prec="fp16", capacity=M + 4096)
    st, n, img = g.forward(BG, capi.V_LITE)
    capi.check(st)
    assert n == M
    _check_reordered(g, img, M, "fp16", "c3")


# ------------------------------------------------------------------ C4
@pytest.fixture(scope="module")
def c4(gpu):
    arrays = scenes.make_config_scene("c4", seed=1, device=DEV)
    arrays.shs = arrays.shs.half()  # FP16 SH storage; halves the device copy
    spec = scenes.config_camera("c4")
    host = _host_view(arrays)
    say("c4 oracle")
    ref = oracle_frame(host, spec, sh_type="fp16", feat_prec="fp16", bg=BG, check_every=8,
                       fp64=True)
    say("c4", "M", ref["pre"]["M"])
    return arrays, host, spec, ref


@pytest.mark.parametrize("depth_first", [False, True])
def test_c4_city_bit_exact_and_image(c4, depth_first):
    arrays, host, spec, ref = c4
    M = ref["pre"]["M"]
    g = Gpu(arrays, spec, sh_type="fp16", feat_prec="fp16", capacity=M + 4096)
    say("c4 forward", "depth first" if depth_first else "default")
    st, n, img = g.forward(BG, capi.V_LITE | (DEPTH_FIRST if depth_first else 0))
    capi.check(st)
    assert n == M
    if not depth_first:  # the depth-first pipeline keeps 32-bit tile ids there
        np.testing.assert_array_equal(_keys(g.ku, M), ref["pre"]["keys"])
    np.testing.assert_array_equal(_keys(g.ks, M),
                                  ref["keys_sorted"] & (TILE_WORD if depth_first else ~np.uint64(0)))
    np.testing.assert_array_equal(_vals(g.vs, M), ref["vals_sorted"])
    rng = g.ranges.cpu().numpy()  # the sort's boundary scan: empty tiles (0, 0)
    np.testing.assert_array_equal(rng, O.tile_ranges(ref["keys_sorted"], g.tiles, 1))
    m64 = _quality(img, ref, ref["image64"], "c4 fp16 persistent")
    print(f"c4 city: {M} pairs, PSNR {m64['psnr_db']:.2f} dB vs FP64", flush=True)
    if not depth_first:
        _check_reordered(g, img, M, "fp16", "c4")


# ------------------------------------------------------------------ C5
def _c5_requests(seed=5, per_combo=3, rate=240.0):
    """Every (scene, resolution) pair `per_combo` times at random orbit
    azimuths, in random order, arriving as a Poisson process."""
    rng = np.random.default_rng(seed)
    combos = [(s, w, h) for s in range(len(SV.SCENE_SIZES)) for (w, h) in SV.RESOLUTIONS]
    items = [c for c in combos for _ in range(per_combo)]
    order = rng.permutation(len(items))
    t = np.cumsum(rng.exponential(1.0 / rate, len(items)))
    reqs = []
    for i, k in enumerate(order):
        s, w, h = items[k]
        r = SV.Request(i, s, w, h, float(rng.uniform(0.0, 2.0 * math.pi)), float(t[i]))
        r.bg = (float(rng.uniform()), float(rng.uniform()), float(rng.uniform()))
        reqs.append(r)
    return reqs


def _alone(sc, r):
    cap = r.pairs + 16
    z = lambda n, dt: torch.zeros(n, dtype=dt, device=DEV)  # noqa: E731
    ku, vu, ks, vs = z(cap, torch.int64), z(cap, torch.int32), z(cap, torch.int64), z(cap, torch.int32)
    space = z(SV._ops().get_sort_buffer_size(cap), torch.int8)
    curr = z(1, torch.int32)
    P = sc["P"]
    e = torch.empty(0, dtype=torch.float32, device=DEV)
    rec = z(24 * P, torch.uint8)
    ranges = torch.zeros((r.tiles, 2), dtype=torch.int32, device=DEV)
    out = torch.zeros((r.height, r.width, 3), dtype=torch.uint8, device=DEV)
    ws = z(SV._ops().forward_workspace_size(P), torch.uint8)
    cam = SV.orbit_camera(r.width, r.height, r.azimuth)
    n = SV._ops().forward_frame(sc["pos"], sc["shs"], sc["opac"], sc["cov"], r.width, r.height,
                                cam.position, cam.rotation, cam.focal_x, cam.focal_y, cam.zFar,
                                cam.zNear, 1, 1, curr, space, ku, vu, ks, vs, e, e, e, rec,
                                ranges, torch.tensor(r.bg, dtype=torch.float32), out, ws, True)
    return n, out.cpu().numpy()


def test_c5_serving_stream_same_alone_and_in_the_mix(gpu):
    sc = SV.make_resident_scenes(SV.SCENE_SIZES, device="cuda:0")
    ex = SV.make_executor(sc)
    est = SV.Estimates()
    say("c5 calibrate", type(ex).__name__)
    SV.calibrate(ex, est, len(sc), repeats=1)
    sched = SV.Scheduler(SV.SCENE_SIZES, 48 << 30, est, max_frames=8, max_batch_ms=6.0)
    reqs = _c5_requests()
    loop = SV.ServingLoop(sched, ex, keep_outputs=True)
    say("c5 run", len(reqs))
    done = loop.run(reqs)
    torch.cuda.synchronize()
    assert len(done) == len(reqs) and not sched.rejected
    summary = SV.summarize(done, sched.rejected, wall_s=ex.now_s(), batches=loop.batches)
    say("c5 summary", summary)
    seen = {(r.scene, r.width, r.height) for r in done}
    assert len(seen) == len(SV.SCENE_SIZES) * len(SV.RESOLUTIONS)
    for r in done:
        n, img = _alone(sc[r.scene], r)
        assert n == r.pairs, (r.rid, n, r.pairs)
        mix = loop.outputs[r.rid].cpu().numpy()
        np.testing.assert_array_equal(mix, img, err_msg=f"request {r.rid} differs in the mix")
    # one request per scene, at a different resolution each, against the oracle
    picks = {}
    for r in done:
        if r.scene not in picks and (r.width, r.height) == SV.RESOLUTIONS[-1 - r.scene]:
            picks[r.scene] = r
    assert len(picks) == len(SV.SCENE_SIZES)
    for s, r in sorted(picks.items()):
        d = sc[s]
        host = scenes.SceneArrays(d["pos"].cpu(), d["shs"], d["opac"].cpu(), d["cov"].cpu())
        cam = SV.orbit_camera(r.width, r.height, r.azimuth)
        ref = O.frame(host.position.numpy(), oracle_sh(d["shs"], "fp16"), host.opacity.numpy(),
                      host.cov3d.numpy(), O.camera_from_spec(cam), r.bg, "fp16", "fp16", 8,
                      False)
        assert ref["pre"]["M"] == r.pairs
        m = diff_metrics(loop.outputs[r.rid].cpu().numpy(), ref["image"])
        say("c5 oracle", SV.SCENE_SIZES[s], r.width, r.height, m)
        assert m["max_abs_diff"] <= 1, (s, r.width, r.height, m)
