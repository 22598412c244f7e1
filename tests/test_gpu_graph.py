"""Frame graphs on the GPU: a frame's pipeline captured once
(ps_frame_graph_create) and replayed with other poses, backgrounds and
output images (ps_frame_graph_launch), through the C ABI and through
``ops.FrameGraph``.

Every replay must equal ps_forward run eagerly for the same pose on buffers
of its own, byte for byte (image, pair count, sorted keys and values), for
each variant, feature precision and the depth-first pipeline; one pose is
also checked against the oracle.  Replays alternate with eager work and
host synchronisations, and a graph is replayed several times, so a capture
that kept state from one replay to the next would show.
"""
import ctypes as C
import math

import numpy as np
import pytest
import torch

import polysplat_amd as polysplat
from polysplat_amd import scenes
from polysplat_amd import serving as SV
from polysplat_amd.metrics import diff_metrics
from tests import capi
from tests.gpu_helpers import PREC, Gpu, oracle_frame, say

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
W, H = 1000, 700
POSES = [(0.0, (0.1, 0.2, 0.3)), (0.5, (1.0, 0.0, 0.5)), (-0.7, (0.0, 0.0, 0.0)),
         (0.0, (0.1, 0.2, 0.3)), (2.4, (0.3, 0.3, 0.9))]
CAP = 4_000_000  # key buffers: room for every pose's pairs


@pytest.fixture(scope="module")
def arrays(gpu):
    return scenes.make_scene(50_000, 8.0, [(0.85, -4.0, 0.5), (0.12, -2.2, 0.4),
                                           (0.03, -0.7, 0.3)], seed=51)


def _spec(az):
    return SV.orbit_camera(W, H, az, radius=12.0)


def _fb(g):
    fb = capi.PsFrameBuffers()
    fb.curr_offset = g.curr.data_ptr()
    fb.sort_space, fb.sort_space_bytes = g.sort_space.data_ptr(), g.sort_bytes
    fb.keys_unsorted, fb.values_unsorted = g.ku.data_ptr(), g.vu.data_ptr()
    fb.keys_sorted, fb.values_sorted = g.ks.data_ptr(), g.vs.data_ptr()
    fb.capacity = g.cap
    fb.points_xy, fb.rgb_depth = g.xy.data_ptr(), g.rgbd.data_ptr()
    fb.conic_opacity, fb.records = g.conic.data_ptr(), g.records.data_ptr()
    fb.tile_ranges = g.ranges.data_ptr()
    fb.workspace, fb.workspace_bytes = g.ws.data_ptr(), g.ws.numel()
    return fb


def _eager(arrays, az, bg, variant, prec, cap):
    e = Gpu(arrays, _spec(az), sh_type=prec, feat_prec=prec, capacity=cap)
    st, n, img = e.forward(bg, variant)
    capi.check(st)
    return e, n, img


@pytest.mark.parametrize("variant,prec", [
    (capi.V_DEFAULT, "fp32"), (capi.V_LITE, "fp16"), (capi.V_FUSED, "fp16"),
    (capi.V_DT, "fp16"), (capi.V_LITE, "bf16"), (capi.V_LITE | capi.DEPTH_FIRST, "fp16"),
    (capi.V_DEFAULT | capi.DEPTH_FIRST, "fp32")])
def test_graph_replays_equal_eager_frames(arrays, variant, prec):
    L = capi.lib()
    g = Gpu(arrays, _spec(POSES[0][0]), sh_type=prec, feat_prec=prec, capacity=CAP)
    graph = C.c_void_p()
    capi.check(L.ps_frame_graph_create(g.P, capi.ptr(g.pos), capi.ptr(g.shs), PREC[prec],
                                       capi.ptr(g.opac), capi.ptr(g.cov), C.byref(g.cam), variant,
                                       PREC[prec], C.byref(_fb(g)), C.byref(graph)))
    try:
        for k, (az, bg) in enumerate(POSES):
            spec = _spec(az)
            cam = capi.camera(spec)
            out = torch.full((H, W, 3), 77, dtype=torch.uint8, device=DEV)
            n = C.c_int(-1)
            say("graph replay", variant, prec, k)
            capi.check(L.ps_frame_graph_launch(graph, C.byref(cam), capi.floats3(bg),
                                               capi.ptr(out), 1, C.byref(n), capi.stream()))
            img = out.cpu().numpy()
            # eager work with its own buffers between the replays
            e, m, ref = _eager(arrays, az, bg, variant, prec, g.cap)
            torch.cuda.synchronize()
            assert n.value == m, (k, n.value, m)
            np.testing.assert_array_equal(img, ref, err_msg=f"replay {k}")
            np.testing.assert_array_equal(g.ks[:m].cpu().numpy(), e.ks[:m].cpu().numpy())
            np.testing.assert_array_equal(g.vs[:m].cpu().numpy(), e.vs[:m].cpu().numpy())
            if k == 1:
                cad = {capi.V_DEFAULT: 2, capi.V_FUSED: 16}.get(variant & 0xff, 8)
                r = oracle_frame(arrays, spec, sh_type=prec, feat_prec=prec, bg=bg,
                                 check_every=cad, fp64=False, dt=(variant & 0xff) == capi.V_DT)
                m_ = diff_metrics(img, r["image"])
                say("graph vs oracle", m_)
                assert m_["max_abs_diff"] <= 1, m_
                assert m == r["pre"]["M"]
        assert not g.ws[:1028].any()  # the queue words are left zero
    finally:
        L.ps_frame_graph_destroy(graph)


def test_graph_posted_counts_and_overflow(arrays):
    L = capi.lib()
    g = Gpu(arrays, _spec(0.0), sh_type="fp16", feat_prec="fp16", capacity=CAP)
    graph = C.c_void_p()
    capi.check(L.ps_frame_graph_create(g.P, capi.ptr(g.pos), capi.ptr(g.shs), capi.F16,
                                       capi.ptr(g.opac), capi.ptr(g.cov), C.byref(g.cam),
                                       capi.V_LITE, capi.F16, C.byref(_fb(g)), C.byref(graph)))
    try:
        outs, tickets = [], []
        for az, bg in POSES[:3]:
            out = torch.zeros((H, W, 3), dtype=torch.uint8, device=DEV)
            t = C.c_int(-1)
            capi.check(L.ps_frame_graph_launch(graph, C.byref(capi.camera(_spec(az))),
                                               capi.floats3(bg), capi.ptr(out), 2, C.byref(t),
                                               capi.stream()))
            outs.append(out)
            tickets.append(t.value)
        for (az, bg), out, t in zip(POSES[:3], outs, tickets):
            n = C.c_int(-1)
            capi.check(L.ps_forward_wait(t, C.byref(n)))
            _, m, ref = _eager(arrays, az, bg, capi.V_LITE, "fp16", g.cap)
            assert n.value == m
            np.testing.assert_array_equal(out.cpu().numpy(), ref)
    finally:
        L.ps_frame_graph_destroy(graph)
    # key buffers too small: the replay reports the overflow
    small = Gpu(arrays, _spec(0.0), sh_type="fp16", feat_prec="fp16", capacity=1000)
    graph = C.c_void_p()
    capi.check(L.ps_frame_graph_create(small.P, capi.ptr(small.pos), capi.ptr(small.shs),
                                       capi.F16, capi.ptr(small.opac), capi.ptr(small.cov),
                                       C.byref(small.cam), capi.V_LITE, capi.F16,
                                       C.byref(_fb(small)), C.byref(graph)))
    try:
        out = torch.zeros((H, W, 3), dtype=torch.uint8, device=DEV)
        n = C.c_int(-1)
        st = L.ps_frame_graph_launch(graph, C.byref(small.cam), capi.floats3((0, 0, 0)),
                                     capi.ptr(out), 1, C.byref(n), capi.stream())
        assert st == capi.PS_ERR_OVERFLOW and n.value > 1000
        bad = capi.camera(SV.orbit_camera(W + 16, H, 0.0))
        assert L.ps_frame_graph_launch(graph, C.byref(bad), capi.floats3((0, 0, 0)),
                                       capi.ptr(out), 1, None, capi.stream()) == capi.PS_ERR_INVALID
    finally:
        L.ps_frame_graph_destroy(graph)


def test_ops_frame_graph_on_a_side_stream(arrays):
    """ops.FrameGraph (the pybind surface), replayed on a torch stream other
    than the default one, against ops.forward_frame."""
    ops = polysplat.ops
    scene = polysplat.Scene.from_arrays(arrays, DEV)
    P = scene.num_vertex
    cap = CAP
    z = lambda n, dt: torch.zeros(n, dtype=dt, device=DEV)  # noqa: E731

    def bufs():
        return dict(curr=z(1, torch.int32), space=z(ops.get_sort_buffer_size(cap), torch.int8),
                    ku=z(cap, torch.int64), vu=z(cap, torch.int32), ks=z(cap, torch.int64),
                    vs=z(cap, torch.int32), rec=z(24 * P, torch.uint8),
                    ranges=torch.zeros((((W + 15) // 16) * ((H + 15) // 16), 2),
                                       dtype=torch.int32, device=DEV),
                    ws=z(ops.forward_workspace_size(P), torch.uint8))

    e = torch.empty(0, dtype=torch.float32, device=DEV)
    b = bufs()
    s0 = _spec(0.0)
    fg = ops.FrameGraph(scene.position, scene.shs_half, scene.opacity, scene.cov3d, W, H,
                        s0.position, s0.rotation, s0.focal_x, s0.focal_y, s0.zFar, s0.zNear, 1, 1,
                        b["curr"], b["space"], b["ku"], b["vu"], b["ks"], b["vs"], e, e, e,
                        b["rec"], b["ranges"], b["ws"])
    side = torch.cuda.Stream(device=DEV)
    ref_b = bufs()
    for az, bg in POSES:
        sp = _spec(az)
        bgt = torch.tensor(bg, dtype=torch.float32)
        out = torch.zeros((H, W, 3), dtype=torch.uint8, device=DEV)
        with torch.cuda.stream(side):
            n = fg.launch(sp.position, sp.rotation, sp.focal_x, sp.focal_y, sp.zFar, sp.zNear, bgt,
                          out, 1)
        side.synchronize()
        ref = torch.zeros_like(out)
        m = ops.forward_frame(scene.position, scene.shs_half, scene.opacity, scene.cov3d, W, H,
                              sp.position, sp.rotation, sp.focal_x, sp.focal_y, sp.zFar, sp.zNear,
                              1, 1, ref_b["curr"], ref_b["space"], ref_b["ku"], ref_b["vu"],
                              ref_b["ks"], ref_b["vs"], e, e, e, ref_b["rec"], ref_b["ranges"],
                              bgt, ref, ref_b["ws"], True)
        torch.cuda.synchronize()
        assert n == m
        np.testing.assert_array_equal(out.cpu().numpy(), ref.cpu().numpy())
    # a graph needs the caller's workspace
    with pytest.raises(ValueError):
        ops.FrameGraph(scene.position, scene.shs_half, scene.opacity, scene.cov3d, W, H,
                       s0.position, s0.rotation, s0.focal_x, s0.focal_y, s0.zFar, s0.zNear, 1, 1,
                       b["curr"], b["space"], b["ku"], b["vu"], b["ks"], b["vs"], e, e, e,
                       b["rec"], b["ranges"], torch.empty(0, dtype=torch.uint8, device=DEV))
    del fg
