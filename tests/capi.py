"""ctypes binding of include/polysplat_amd.h for the tests.

This is the binding a ctypes user of the library would write: plain
pointers, sizes and a stream handle.  GPU tests pass torch tensors' device
pointers; CPU tests only load the library and call host-only entry points.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "polysplat_amd" / "libpolysplat_amd.so"
HEADER = ROOT / "include" / "polysplat_amd.h"

PS_OK, PS_ERR_INVALID, PS_ERR_HIP, PS_ERR_OVERFLOW, PS_ERR_UNSUPPORTED = range(5)
F32, F16, BF16 = 0, 1, 2
PREC_IDS = {"fp32": F32, "fp16": F16, "bf16": BF16}
V_DEFAULT, V_LITE, V_FUSED, V_DT = 0, 1, 2, 3
DEPTH_FIRST = 0x100  # PS_FORWARD_DEPTH_FIRST, OR-ed into a variant
TILE_ORDER = 0x200  # PS_FORWARD_TILE_ORDER
MAX_BATCH_FRAMES = 64


class PsCamera(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("block_x", C.c_int),
                ("block_y", C.c_int), ("position", C.c_float * 3), ("rotation", C.c_float * 9),
                ("focal_x", C.c_float), ("focal_y", C.c_float), ("z_far", C.c_float),
                ("z_near", C.c_float)]


class PsFrameBuffers(C.Structure):
    _fields_ = [("curr_offset", C.c_void_p), ("sort_space", C.c_void_p),
                ("sort_space_bytes", C.c_size_t), ("keys_unsorted", C.c_void_p),
                ("values_unsorted", C.c_void_p), ("keys_sorted", C.c_void_p),
                ("values_sorted", C.c_void_p), ("capacity", C.c_int), ("points_xy", C.c_void_p),
                ("rgb_depth", C.c_void_p), ("conic_opacity", C.c_void_p),
                ("records", C.c_void_p), ("tile_ranges", C.c_void_p),
                ("workspace", C.c_void_p), ("workspace_bytes", C.c_size_t)]


class PsFrame(C.Structure):
    _fields_ = [("P", C.c_int), ("positions", C.c_void_p), ("shs", C.c_void_p),
                ("opacities", C.c_void_p), ("cov3d", C.c_void_p), ("cam", PsCamera),
                ("bg", C.c_float * 3), ("points_xy", C.c_void_p), ("rgb_depth", C.c_void_p),
                ("conic_opacity", C.c_void_p), ("records", C.c_void_p),
                ("out_color", C.c_void_p)]


def header_functions() -> list[str]:
    """Names of the functions declared in the public header."""
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ps_[a-z0-9_]+)\s*\(", text)))


_lib = None


def lib():
    """The GPU library with every exported function's signature applied."""
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is not built; run python polysplat_amd/_build.py")
        _lib = bind(C.CDLL(str(LIB)))
    return _lib


def bind(L):
    """Applies the C ABI signatures to a loaded library (GPU or host emulation)."""
    P, I, S = C.c_void_p, C.c_int, C.c_size_t
    sig = {
        "ps_status_string": ([I], C.c_char_p),
        "ps_last_error": ([], C.c_char_p),
        "ps_version": ([], C.c_char_p),
        "ps_camera_matrices": ([P, P, P, P], I),
        "ps_preprocess_workspace_size": ([I], S),
        "ps_preprocess": ([I, P, P, I, P, P, P, P, P, P, P, I, P, P, I, P, P, S, P], I),
        "ps_preprocess_count_ptr": ([P], P),
        "ps_preprocess_status_ptr": ([P], P),
        "ps_preprocess_depth_first": ([I, P, P, I, P, P, P, P, P, P, P, I, P, P, I, P, P, S, P,
                                       S, P], I),
        "ps_es_tile_sort": ([I, P, I, I, I, I, I, P, S, P, P, P, P, P, P, P], I),
        "ps_get_sort_buffer_size": ([I], S),
        "ps_sort_gaussian": ([I, P, I, I, I, I, I, P, S, P, P, P, P, P], I),
        "ps_fetch_num_rendered": ([P, P, P], I),
        "ps_precompute_tile_ranges": ([I, P, I, I, I, I, P, P, P], I),
        "ps_precompute_tile_ranges_scan": ([I, P, I, I, I, I, I, P, P, P], I),
        "ps_gather_features": ([I, P, I, P, P, P, P, P, P, P, P], I),
        "ps_gather_records": ([I, P, I, P, P, P, P], I),
        "ps_render_16x16": ([I, P, I, I, P, P, P, P, P, P, P, P], I),
        "ps_render_16x16_persistent": ([I, P, I, I, I, P, P, P, P, I, P, P, P, I, P, P, P, P],
                                       I),
        "ps_render_16x16_persistent_noalloc": ([I, P, I, I, I, P, P, P, P, I, P, P, P, I, P, P,
                                                P], I),
        "ps_get_tile_order_sort_temp_size": ([I], S),
        "ps_compute_tile_order": ([P, I, P, P, P, P, S, P], I),
        "ps_compute_tile_order_packed": ([P, I, I, P, P, P, P, P, S, I, I, P], I),
        "ps_render_16x16_reordered": ([I, P, I, I, P, P, P, P, P, P, P, P, P, P], I),
        "ps_render_16x16_reordered_desc": ([I, P, I, I, P, P, P, P, P, P, P, P], I),
        "ps_render_16x16_topk": ([I, P, I, I, P, P, P, P, P, P, P, P, I, I, P, P, P, P], I),
        "ps_render_16x16_reordered_persistent": ([I, P, I, I, I, P, P, P, P, P, P, P, P, I, P,
                                                  P, P, P], I),
        "ps_frame_graph_create": ([I, P, P, I, P, P, P, I, I, P, P], I),
        "ps_frame_graph_launch": ([P, P, P, P, I, P, P], I),
        "ps_frame_graph_destroy": ([P], I),
        "ps_forward_workspace_size": ([I], S),
        "ps_forward_tile_order_space": ([I], S),
        "ps_forward": ([I, P, P, I, P, P, P, I, I, P, P, P, I, P, P], I),
        "ps_forward_wait": ([I, P], I),
        "ps_forward_query": ([I], I),
        "ps_forward_batch_workspace_size": ([I, P], S),
        "ps_forward_batch": ([I, P, I, I, P, I, P, P], I),
        "ps_forward_batch_counts_ptr": ([P, I, P], P),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    return L


def check(status: int) -> None:
    check_with(lib() if status != PS_OK else None, status)


def check_with(L, status: int) -> None:
    """Raises with the library's own message when status is not PS_OK."""
    if status != PS_OK:
        raise RuntimeError(f"{L.ps_status_string(status).decode()}: {L.ps_last_error().decode()}")


def camera(spec, block=16) -> PsCamera:
    c = PsCamera()
    c.width, c.height, c.block_x, c.block_y = spec.width, spec.height, block, block
    pos = spec.position.reshape(3).tolist()
    rot = spec.rotation.reshape(9).tolist()
    for i in range(3):
        c.position[i] = pos[i]
    for i in range(9):
        c.rotation[i] = rot[i]
    c.focal_x, c.focal_y, c.z_far, c.z_near = spec.focal_x, spec.focal_y, spec.zFar, spec.zNear
    return c


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def floats3(v):
    return (C.c_float * 3)(*[float(x) for x in v])
