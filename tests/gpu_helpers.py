"""Stage-by-stage GPU runs through the C ABI, for the parity tests."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import torch

import oracle as O
from tests import capi

DT = {"fp32": torch.float32, "fp16": torch.float16, "bf16": torch.bfloat16}
PREC = {"fp32": capi.F32, "fp16": capi.F16, "bf16": capi.BF16}


def say(*a):
    print("[gpu]", *a, file=sys.stderr, flush=True)


def oracle_sh(shs: torch.Tensor, sh_type: str) -> np.ndarray:
    """SH array as the oracle reads it for a given storage type."""
    if sh_type == "fp32":
        return shs.float().cpu().numpy()
    return shs.to(DT[sh_type]).cpu().view(torch.int16).numpy().view(np.uint16)


class Gpu:
    def __init__(self, scene, spec, sh_type="fp32", feat_prec="fp32", capacity=None,
                 device="cuda:0"):
        self.dev = torch.device(device)
        self.spec = spec
        self.P = scene.num_vertex
        self.sh_type = sh_type
        self.feat_prec = feat_prec
        self.pos = scene.position.to(self.dev).contiguous()
        self.shs = scene.shs.to(self.dev, DT[sh_type]).contiguous()
        self.opac = scene.opacity.to(self.dev).contiguous()
        self.cov = scene.cov3d.to(self.dev).contiguous()
        P = max(self.P, 1)
        self.xy = torch.zeros(P, 2, device=self.dev)
        self.rgbd = torch.zeros(P, 4, device=self.dev)
        self.conic = torch.zeros(P, 4, device=self.dev)
        self.records = torch.zeros(24 * P, device=self.dev, dtype=torch.uint8)
        self.cap = int(capacity if capacity is not None else 8 * P + 1024)
        self.ku = torch.zeros(self.cap, device=self.dev, dtype=torch.int64)
        self.vu = torch.zeros(self.cap, device=self.dev, dtype=torch.int32)
        self.ks = torch.zeros(self.cap, device=self.dev, dtype=torch.int64)
        self.vs = torch.zeros(self.cap, device=self.dev, dtype=torch.int32)
        self.curr = torch.zeros(1, device=self.dev, dtype=torch.int32)
        L = capi.lib()
        self.sort_bytes = L.ps_get_sort_buffer_size(self.cap)
        self.sort_space = torch.zeros(self.sort_bytes, device=self.dev, dtype=torch.uint8)
        self.ws_bytes = L.ps_forward_workspace_size(self.P)
        self.ws = torch.zeros(self.ws_bytes, device=self.dev, dtype=torch.uint8)
        self.queue = torch.zeros(512, device=self.dev, dtype=torch.int32)
        gx, gy = (spec.width + 15) // 16, (spec.height + 15) // 16
        self.tiles = gx * gy
        self.ranges = torch.zeros(self.tiles, 2, device=self.dev, dtype=torch.int32)
        self.out = torch.zeros(spec.height, spec.width, 3, device=self.dev, dtype=torch.uint8)
        self.cam = capi.camera(spec)
        self.M = None

    def preprocess(self, records=None):
        L = capi.lib()
        rec = self.feat_prec != "fp32" if records is None else records
        self.curr.zero_()
        st = L.ps_preprocess(self.P, capi.ptr(self.pos), capi.ptr(self.shs),
                             PREC[self.sh_type], capi.ptr(self.opac), capi.ptr(self.cov),
                             C.byref(self.cam), capi.ptr(self.xy), capi.ptr(self.rgbd),
                             capi.ptr(self.conic), capi.ptr(self.records) if rec else None,
                             PREC[self.feat_prec] if rec else capi.F16, capi.ptr(self.ku),
                             capi.ptr(self.vu), self.cap, capi.ptr(self.curr),
                             capi.ptr(self.ws), self.ws_bytes, capi.stream())
        capi.check(st)
        torch.cuda.synchronize()
        self.M = int(self.curr.item())
        return self.M

    def status_word(self):
        return int(self.ws[8:12].view(torch.int32).item())

    def sort(self, n=None):
        L = capi.lib()
        n = self.M if n is None else n
        capi.check(L.ps_sort_gaussian(n, None, self.cap, self.spec.width, self.spec.height, 16, 16,
                                      capi.ptr(self.sort_space), self.sort_bytes,
                                      capi.ptr(self.ku), capi.ptr(self.vu), capi.ptr(self.ks),
                                      capi.ptr(self.vs), capi.stream()))
        torch.cuda.synchronize()

    def ranges_bsearch(self):
        capi.check(capi.lib().ps_precompute_tile_ranges(
            self.M, None, self.spec.width, self.spec.height, 16, 16, capi.ptr(self.ks),
            capi.ptr(self.ranges), capi.stream()))
        torch.cuda.synchronize()
        return self.ranges.cpu().numpy().copy()

    def ranges_scan(self):
        capi.check(capi.lib().ps_precompute_tile_ranges_scan(
            self.M, None, self.M, self.spec.width, self.spec.height, 16, 16, capi.ptr(self.ks),
            capi.ptr(self.ranges), capi.stream()))
        torch.cuda.synchronize()
        return self.ranges.cpu().numpy().copy()

    def render_plain(self, bg):
        self.out.zero_()
        capi.check(capi.lib().ps_render_16x16(
            self.M, None, self.spec.width, self.spec.height, capi.ptr(self.xy),
            capi.ptr(self.rgbd), capi.ptr(self.conic), capi.ptr(self.ks), capi.ptr(self.vs),
            capi.floats3(bg), capi.ptr(self.out), capi.stream()))
        torch.cuda.synchronize()
        return self.out.cpu().numpy().copy()

    def render_persistent(self, bg, use_ranges=True, dt=False, prec=None, gathered=False,
                          feats=None):
        prec = self.feat_prec if prec is None else prec
        xy, rgbd, con, rec = (self.xy, self.rgbd, self.conic, self.records) if feats is None else feats
        self.out.zero_()
        capi.check(capi.lib().ps_render_16x16_persistent(
            self.M, None, self.spec.width, self.spec.height, PREC[prec], capi.ptr(xy),
            capi.ptr(rgbd), capi.ptr(con), capi.ptr(rec), 1 if gathered else 0,
            capi.ptr(self.ks), capi.ptr(self.vs), capi.ptr(self.ranges) if use_ranges else None,
            1 if dt else 0, capi.floats3(bg), capi.ptr(self.out), capi.ptr(self.queue),
            capi.stream()))
        torch.cuda.synchronize()
        return self.out.cpu().numpy().copy()

    def forward(self, bg, variant, sync=True, ws=None):
        fb = capi.PsFrameBuffers()
        fb.curr_offset = self.curr.data_ptr()
        fb.sort_space = self.sort_space.data_ptr()
        fb.sort_space_bytes = self.sort_bytes
        fb.keys_unsorted, fb.values_unsorted = self.ku.data_ptr(), self.vu.data_ptr()
        fb.keys_sorted, fb.values_sorted = self.ks.data_ptr(), self.vs.data_ptr()
        fb.capacity = self.cap
        fb.points_xy, fb.rgb_depth, fb.conic_opacity = (self.xy.data_ptr(), self.rgbd.data_ptr(),
                                                        self.conic.data_ptr())
        fb.records = self.records.data_ptr()
        fb.tile_ranges = self.ranges.data_ptr()
        w = self.ws if ws is None else ws
        fb.workspace, fb.workspace_bytes = w.data_ptr(), w.numel()
        n = C.c_int(-1)
        self.out.zero_()
        st = capi.lib().ps_forward(self.P, capi.ptr(self.pos), capi.ptr(self.shs),
                                   PREC[self.sh_type], capi.ptr(self.opac), capi.ptr(self.cov),
                                   C.byref(self.cam), variant, PREC[self.feat_prec], C.byref(fb),
                                   capi.floats3(bg), capi.ptr(self.out), int(sync),
                                   C.byref(n), capi.stream())
        torch.cuda.synchronize()
        return st, n.value, self.out.cpu().numpy().copy()


def oracle_frame(scene, spec, sh_type="fp32", feat_prec="fp32", bg=(0, 0, 0), check_every=8,
                 fp64=True, dt=False):
    return O.frame(scene.position.numpy(), oracle_sh(scene.shs, sh_type), scene.opacity.numpy(),
                   scene.cov3d.numpy(), O.camera_from_spec(spec), bg, sh_type, feat_prec,
                   check_every, fp64, dt)
