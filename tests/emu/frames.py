"""Frames through the host emulation of the kernels (test infrastructure).

Same C ABI calls as the GPU tests, with numpy arrays as "device" memory:
the emulated kernels dereference them directly, so an out-of-bounds access
is a host memory error (AddressSanitizer reports it when the ASan build is
loaded with libasan preloaded).  Buffers can start as random bytes instead
of zeros, so a kernel that reads memory it never wrote is visible too.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from tests import capi

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libpolysplat_emu.so"
LIB_ASAN = HERE / "_build" / "libpolysplat_emu_asan.so"

PREC = {"fp32": 0, "fp16": 1, "bf16": 2}


def load(asan: bool | None = None):
    if asan is None:
        asan = "libasan" in os.environ.get("LD_PRELOAD", "")
    path = LIB_ASAN if asan else LIB
    if not path.exists():
        raise RuntimeError(f"{path} is not built; run python tests/emu/build.py")
    return capi.bind(C.CDLL(str(path)))


def _sh(shs, sh_type):
    import torch

    if sh_type == "fp32":
        return np.ascontiguousarray(shs.float().numpy())
    dt = torch.float16 if sh_type == "fp16" else torch.bfloat16
    return np.ascontiguousarray(shs.to(dt).view(torch.int16).numpy().view(np.uint16))


class Buffers:
    """Allocates arrays either zeroed or filled with random bytes."""

    def __init__(self, garbage: bool, seed: int = 0):
        self.garbage = garbage
        self.rng = np.random.default_rng(seed)

    def __call__(self, shape, dtype):
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        if self.garbage:
            raw = self.rng.integers(0, 256, max(n, 1), dtype=np.uint8)
        else:
            raw = np.zeros(max(n, 1), np.uint8)
        return raw[:n].view(dtype).reshape(shape)


def ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class HostFrame:
    """One scene + camera with its own buffers, like the GPU tests' Gpu."""

    def __init__(self, L, scene, spec, sh_type="fp16", feat_prec="fp16", capacity=None,
                 garbage=False, seed=0):
        self.L = L
        alloc = Buffers(garbage, seed)
        self.spec = spec
        self.P = scene.num_vertex
        self.sh_type, self.feat_prec = sh_type, feat_prec
        self.pos = np.ascontiguousarray(scene.position.numpy())
        self.shs = _sh(scene.shs, sh_type)
        self.opac = np.ascontiguousarray(scene.opacity.numpy())
        self.cov = np.ascontiguousarray(scene.cov3d.numpy())
        P = max(self.P, 1)
        self.xy = alloc((P, 2), np.float32)
        self.rgbd = alloc((P, 4), np.float32)
        self.conic = alloc((P, 4), np.float32)
        self.records = alloc((24 * P,), np.uint8)
        self.cap = int(capacity if capacity is not None else 8 * P + 1024)
        self.ku = alloc((self.cap,), np.uint64)
        self.vu = alloc((self.cap,), np.uint32)
        self.ks = alloc((self.cap,), np.uint64)
        self.vs = alloc((self.cap,), np.uint32)
        self.curr = alloc((1,), np.int32)
        self.sort_bytes = L.ps_get_sort_buffer_size(self.cap)
        self.sort_space = alloc((self.sort_bytes,), np.uint8)
        self.ws_bytes = L.ps_forward_workspace_size(self.P)
        self.ws = alloc((self.ws_bytes,), np.uint8)
        self.ws[:2048] = 0  # persistent-queue words start at zero
        self.tiles = ((spec.width + 15) // 16) * ((spec.height + 15) // 16)
        self.ranges = alloc((self.tiles, 2), np.int32)
        self.out = alloc((spec.height, spec.width, 3), np.uint8)
        self.cam = capi.camera(spec)

    def forward(self, bg, variant=capi.V_LITE, sync=True):
        fb = capi.PsFrameBuffers()
        fb.curr_offset = self.curr.ctypes.data
        fb.sort_space, fb.sort_space_bytes = self.sort_space.ctypes.data, self.sort_bytes
        fb.keys_unsorted, fb.values_unsorted = self.ku.ctypes.data, self.vu.ctypes.data
        fb.keys_sorted, fb.values_sorted = self.ks.ctypes.data, self.vs.ctypes.data
        fb.capacity = self.cap
        fb.points_xy, fb.rgb_depth = self.xy.ctypes.data, self.rgbd.ctypes.data
        fb.conic_opacity, fb.records = self.conic.ctypes.data, self.records.ctypes.data
        fb.tile_ranges = self.ranges.ctypes.data
        fb.workspace, fb.workspace_bytes = self.ws.ctypes.data, self.ws_bytes
        n = C.c_int(-1)
        st = self.L.ps_forward(self.P, ptr(self.pos), ptr(self.shs), PREC[self.sh_type],
                               ptr(self.opac), ptr(self.cov), C.byref(self.cam), variant,
                               PREC[self.feat_prec], C.byref(fb), capi.floats3(bg), ptr(self.out),
                               int(sync), C.byref(n), None)
        return st, n.value, self.out.copy()


class HostBatch:
    """Shared buffers for ps_forward_batch over HostFrames."""

    def __init__(self, L, frames, capacity, garbage=False, seed=1):
        alloc = Buffers(garbage, seed)
        self.L = L
        self.frames = frames
        self.cap = capacity
        self.ku = alloc((capacity,), np.uint64)
        self.vu = alloc((capacity,), np.uint32)
        self.ks = alloc((capacity,), np.uint64)
        self.vs = alloc((capacity,), np.uint32)
        self.curr = alloc((1,), np.int32)
        self.sort_bytes = L.ps_get_sort_buffer_size(capacity)
        self.sort_space = alloc((self.sort_bytes,), np.uint8)
        self.ranges = alloc((sum(f.tiles for f in frames), 2), np.int32)
        P = (C.c_int * len(frames))(*[f.P for f in frames])
        self.ws_bytes = L.ps_forward_batch_workspace_size(len(frames), P)
        self.ws = alloc((self.ws_bytes,), np.uint8)
        self.ws[:2048] = 0

    def run(self, bgs, sync=True):
        fr = (capi.PsFrame * len(self.frames))()
        for q, f, bg in zip(fr, self.frames, bgs):
            q.P = f.P
            q.positions, q.shs = f.pos.ctypes.data, f.shs.ctypes.data
            q.opacities, q.cov3d = f.opac.ctypes.data, f.cov.ctypes.data
            q.cam = f.cam
            for i in range(3):
                q.bg[i] = bg[i]
            q.points_xy, q.rgb_depth = f.xy.ctypes.data, f.rgbd.ctypes.data
            q.conic_opacity, q.records = f.conic.ctypes.data, f.records.ctypes.data
            q.out_color = f.out.ctypes.data
        fb = capi.PsFrameBuffers()
        fb.curr_offset = self.curr.ctypes.data
        fb.sort_space, fb.sort_space_bytes = self.sort_space.ctypes.data, self.sort_bytes
        fb.keys_unsorted, fb.values_unsorted = self.ku.ctypes.data, self.vu.ctypes.data
        fb.keys_sorted, fb.values_sorted = self.ks.ctypes.data, self.vs.ctypes.data
        fb.capacity = self.cap
        fb.tile_ranges = self.ranges.ctypes.data
        fb.workspace, fb.workspace_bytes = self.ws.ctypes.data, self.ws_bytes
        counts = (C.c_int * len(self.frames))()
        f0 = self.frames[0]
        st = self.L.ps_forward_batch(len(self.frames), fr, PREC[f0.sh_type], PREC[f0.feat_prec],
                                     C.byref(fb), 1 if sync else 0, counts, None)
        return st, list(counts), [f.out.copy() for f in self.frames]
