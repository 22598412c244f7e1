"""Frame graphs (ps_frame_graph_*) through the host emulation of the kernels.

A frame's pipeline is captured once and replayed with other camera poses,
backgrounds and output images; every replay must equal ps_forward run
eagerly for the same pose on buffers of its own (image, pair count, sorted
keys and values), for each variant and the depth-first pipeline.  The
emulated runtime records the launches of a capturing stream and refuses
copies and event records on it, so a captured path that would put a copy
into the graph fails here.
"""
import ctypes as C
import math
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


def _orbit(W, H, az, dist=9.0):
    s, c = math.sin(az), math.cos(az)
    spec = scenes.make_camera(W, H, 10.0, distance=dist)
    spec.position[:] = scenes.torch.tensor([dist * s, 0.3, -dist * c])
    spec.rotation[:] = scenes.torch.tensor([[c, 0.0, -s], [0.0, 1.0, 0.0], [s, 0.0, c]])
    return spec


def _fb(f):
    fb = capi.PsFrameBuffers()
    fb.curr_offset = f.curr.ctypes.data
    fb.sort_space, fb.sort_space_bytes = f.sort_space.ctypes.data, f.sort_bytes
    fb.keys_unsorted, fb.values_unsorted = f.ku.ctypes.data, f.vu.ctypes.data
    fb.keys_sorted, fb.values_sorted = f.ks.ctypes.data, f.vs.ctypes.data
    fb.capacity = f.cap
    fb.points_xy, fb.rgb_depth = f.xy.ctypes.data, f.rgbd.ctypes.data
    fb.conic_opacity, fb.records = f.conic.ctypes.data, f.records.ctypes.data
    fb.tile_ranges = f.ranges.ctypes.data
    fb.workspace, fb.workspace_bytes = f.ws.ctypes.data, f.ws_bytes
    return fb


# Two variants here (the plain and the persistent renderer; the
# depth-first keys cost minutes in the emulation); the GPU test runs all of
# them, the depth-first pipeline included.
@pytest.mark.parametrize("variant,prec", [(capi.V_DEFAULT, "fp32"), (capi.V_LITE, "fp16")])
def test_graph_replays_equal_eager_frames(emu, variant, prec):
    L = emu
    arr = scenes.make_config_scene("c1", seed=7, P=3000)
    W, H = 160, 112
    poses = [(0.0, (0.1, 0.2, 0.3)), (0.6, (1.0, 0.0, 0.5)), (0.0, (0.1, 0.2, 0.3))]
    g_frame = HF.HostFrame(L, arr, _orbit(W, H, poses[0][0]), sh_type=prec, feat_prec=prec,
                           garbage=True, seed=3)
    graph = C.c_void_p()
    st = L.ps_frame_graph_create(g_frame.P, HF.ptr(g_frame.pos), HF.ptr(g_frame.shs),
                                 HF.PREC[prec], HF.ptr(g_frame.opac), HF.ptr(g_frame.cov),
                                 C.byref(g_frame.cam), variant, HF.PREC[prec],
                                 C.byref(_fb(g_frame)), C.byref(graph))
    capi.check_with(L, st)
    try:
        for k, (az, bg) in enumerate(poses):
            spec = _orbit(W, H, az)
            cam = capi.camera(spec)
            out = np.full((H, W, 3), 77, np.uint8)
            n = C.c_int(-1)
            capi.check_with(L, L.ps_frame_graph_launch(graph, C.byref(cam), capi.floats3(bg),
                                                       HF.ptr(out), 1, C.byref(n), None))
            eager = HF.HostFrame(L, arr, spec, sh_type=prec, feat_prec=prec)
            st, m, img = eager.forward(bg, variant)
            capi.check_with(L, st)
            assert n.value == m, (k, n.value, m)
            np.testing.assert_array_equal(out, img, err_msg=f"replay {k}")
            np.testing.assert_array_equal(g_frame.ks[:m], eager.ks[:m])
            np.testing.assert_array_equal(g_frame.vs[:m], eager.vs[:m])
            assert not g_frame.ws[:1028].any()  # queue words left zero
        # a posted count through the same graph
        cam = capi.camera(_orbit(W, H, 0.2))
        out = np.zeros((H, W, 3), np.uint8)
        ticket = C.c_int(-1)
        capi.check_with(L, L.ps_frame_graph_launch(graph, C.byref(cam), capi.floats3((0, 0, 0)),
                                                   HF.ptr(out), 2, C.byref(ticket), None))
        n = C.c_int(-1)
        capi.check_with(L, L.ps_forward_wait(ticket.value, C.byref(n)))
        assert n.value > 0
        # wrong resolution / missing output: refused before anything runs
        bad = capi.camera(_orbit(W + 16, H, 0.0))
        assert L.ps_frame_graph_launch(graph, C.byref(bad), capi.floats3((0, 0, 0)), HF.ptr(out),
                                       1, None, None) == capi.PS_ERR_INVALID
        assert L.ps_frame_graph_launch(graph, C.byref(cam), capi.floats3((0, 0, 0)), None, 1,
                                       None, None) == capi.PS_ERR_INVALID
    finally:
        L.ps_frame_graph_destroy(graph)


def test_graph_overflow_and_bad_arguments(emu):
    L = emu
    arr = scenes.make_config_scene("c1", seed=8, P=2000)
    spec = _orbit(160, 120, 0.0)
    f = HF.HostFrame(L, arr, spec, sh_type="fp16", feat_prec="fp16", capacity=64)
    graph = C.c_void_p()
    capi.check_with(L, L.ps_frame_graph_create(f.P, HF.ptr(f.pos), HF.ptr(f.shs), 1,
                                               HF.ptr(f.opac), HF.ptr(f.cov), C.byref(f.cam),
                                               capi.V_LITE, 1, C.byref(_fb(f)), C.byref(graph)))
    try:
        out = np.zeros((120, 160, 3), np.uint8)
        n = C.c_int(-1)
        st = L.ps_frame_graph_launch(graph, C.byref(f.cam), capi.floats3((0, 0, 0)), HF.ptr(out),
                                     1, C.byref(n), None)
        assert st == capi.PS_ERR_OVERFLOW and n.value > 64
    finally:
        L.ps_frame_graph_destroy(graph)
    # workspace too small: refused, no graph
    fb = _fb(f)
    fb.workspace_bytes = 16
    g2 = C.c_void_p(1)
    assert L.ps_frame_graph_create(f.P, HF.ptr(f.pos), HF.ptr(f.shs), 1, HF.ptr(f.opac),
                                   HF.ptr(f.cov), C.byref(f.cam), capi.V_LITE, 1, C.byref(fb),
                                   C.byref(g2)) == capi.PS_ERR_INVALID
    assert not g2.value
    # a bad variant fails inside the capture: reported, no graph
    assert L.ps_frame_graph_create(f.P, HF.ptr(f.pos), HF.ptr(f.shs), 1, HF.ptr(f.opac),
                                   HF.ptr(f.cov), C.byref(f.cam), 9, 1, C.byref(_fb(f)),
                                   C.byref(g2)) == capi.PS_ERR_INVALID
    assert not g2.value
