"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, and its host-only entry points agree with the oracle.  No GPU call
is made here."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from polysplat_amd import scenes
from tests import capi


def test_library_exports_every_header_symbol():
    L = capi.lib()
    names = capi.header_functions()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_strings_and_version():
    L = capi.lib()
    assert L.ps_status_string(capi.PS_OK) == b"ok"
    assert b"pairs" in L.ps_status_string(capi.PS_ERR_OVERFLOW)
    assert L.ps_version().startswith(b"polysplat_amd")


@pytest.mark.parametrize("cfg", ["c1", "c2", "c3"])
def test_camera_matrices_bit_exact_with_oracle(cfg):
    spec = scenes.config_camera(cfg)
    view, proj, tf = (np.zeros(16, np.float32), np.zeros(16, np.float32), np.zeros(2, np.float32))
    cam = capi.camera(spec)
    capi.check(capi.lib().ps_camera_matrices(C.byref(cam), view.ctypes.data, proj.ctypes.data,
                                             tf.ctypes.data))
    v2, p2, t2 = O.camera_matrices(O.camera_from_spec(spec))
    np.testing.assert_array_equal(view, v2)
    np.testing.assert_array_equal(proj, p2)
    np.testing.assert_array_equal(tf, t2)


def test_camera_matrices_rotated_pose_bit_exact():
    ang = 0.7
    rot = np.array([[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]],
                   np.float32)
    spec = scenes.make_camera(1280, 720, 10.0)
    spec.rotation = __import__("torch").from_numpy(rot)
    spec.position = __import__("torch").tensor([1.5, -0.25, -9.0])
    view, proj, tf = (np.zeros(16, np.float32), np.zeros(16, np.float32), np.zeros(2, np.float32))
    cam = capi.camera(spec)
    capi.check(capi.lib().ps_camera_matrices(C.byref(cam), view.ctypes.data, proj.ctypes.data,
                                             tf.ctypes.data))
    v2, p2, t2 = O.camera_matrices(O.camera_from_spec(spec))
    np.testing.assert_array_equal(view, v2)
    np.testing.assert_array_equal(proj, p2)


def test_workspace_sizes_are_monotone():
    L = capi.lib()
    a = [L.ps_get_sort_buffer_size(n) for n in (0, 1, 1000, 10**6, 10**7)]
    assert a == sorted(a) and a[-1] >= 12 * 10**7
    b = [L.ps_preprocess_workspace_size(n) for n in (0, 1, 10**6)]
    assert b == sorted(b) and b[-1] >= 4 * 10**6
    assert L.ps_forward_workspace_size(1000) > L.ps_preprocess_workspace_size(1000)


def test_invalid_arguments_are_rejected_without_gpu():
    L = capi.lib()
    # NULL camera / NULL buffers fail argument validation before any HIP call.
    st = L.ps_preprocess(10, None, None, 0, None, None, None, None, None, None, None, 0, None,
                         None, 0, None, None, 0, None)
    assert st == capi.PS_ERR_INVALID
    assert b"camera" in L.ps_last_error() or b"NULL" in L.ps_last_error()
    st = L.ps_sort_gaussian(5, None, 4, 64, 64, 16, 16, None, 0, None, None, None, None, None)
    assert st == capi.PS_ERR_INVALID


def test_batch_workspace_and_argument_checks_without_gpu():
    import ctypes as C

    L = capi.lib()
    P = (C.c_int * 3)(1000, 0, 50_000)
    one = (C.c_int * 1)(1000)
    w3 = L.ps_forward_batch_workspace_size(3, P)
    assert w3 > L.ps_forward_batch_workspace_size(1, one) > 0
    assert L.ps_forward_batch_workspace_size(0, P) == 0
    frames = (capi.PsFrame * 2)()
    fb = capi.PsFrameBuffers()
    # too many frames, no frames, NULL arguments: rejected before any HIP call
    assert L.ps_forward_batch(capi.MAX_BATCH_FRAMES + 1, frames, 0, 1, C.byref(fb), 0, None,
                              None) == capi.PS_ERR_INVALID
    assert L.ps_forward_batch(0, frames, 0, 1, C.byref(fb), 0, None, None) == capi.PS_ERR_INVALID
    assert L.ps_forward_batch(1, None, 0, 1, C.byref(fb), 0, None, None) == capi.PS_ERR_INVALID
    # frame without 16x16 tiles
    frames[0].cam.width, frames[0].cam.height = 64, 64
    frames[0].cam.block_x, frames[0].cam.block_y = 8, 8
    fb.curr_offset = 1
    fb.tile_ranges = 1
    assert L.ps_forward_batch(1, frames, 0, 1, C.byref(fb), 0, None, None) == capi.PS_ERR_INVALID
    assert b"16x16" in L.ps_last_error()
