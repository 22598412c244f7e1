"""CPU oracle for the forward rasterizer -- TEST INFRASTRUCTURE ONLY.

ctypes/numpy front end of ``oracle/liboracle_ps.so`` (plain C, built by
``oracle/Makefile``).  Only tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg use it, as the checker; the product never
imports it.

``frame()`` runs the whole FP32 restatement of the reference's forward pass
(preprocess -> stable sort -> tile ranges -> blend) and, on request, the FP64
evaluation of the same blend with the FP32 discrete decisions.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_ps.so"

SH_TYPES = {"fp32": 0, "fp16": 1, "bf16": 2}
PREC = {"fp32": 0, "fp16": 1, "bf16": 2}


def build(force: bool = False) -> Path:
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "-B" if force else "liboracle_ps.so"],
                       check=True)
    return LIB_PATH


class OrcCamera(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("block_x", C.c_int),
                ("block_y", C.c_int), ("position", C.c_float * 3), ("rotation", C.c_float * 9),
                ("focal_x", C.c_float), ("focal_y", C.c_float), ("z_far", C.c_float),
                ("z_near", C.c_float)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        L.orc_camera_matrices.argtypes = [P, P, P, P]
        L.orc_log2.argtypes = [C.c_float]
        L.orc_log2.restype = C.c_float
        L.orc_cov2d.argtypes = [P, P, P, P]
        L.orc_sh_color.argtypes = [P, P, P, P]
        L.orc_tile_hits.argtypes = [C.c_int, C.c_int] + [C.c_float] * 6 + [C.c_int, C.c_int]
        L.orc_tile_hits.restype = C.c_int
        L.orc_preprocess.argtypes = [C.c_int, P, P, C.c_int, P, P, P, P, P, P, P, P, P,
                                     C.c_longlong]
        L.orc_preprocess.restype = C.c_longlong
        L.orc_sort_pairs.argtypes = [C.c_longlong, P, P]
        L.orc_morton.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_morton.restype = C.c_uint32
        L.orc_tile_order.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P]
        L.orc_tile_ranges.argtypes = [C.c_longlong, P, C.c_int, P, C.c_int]
        L.orc_render_f32.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P, C.c_int, C.c_int, P, P,
                                     P, P]
        L.orc_features_f64.argtypes = [C.c_int, P, P, P, C.c_int, P, P, P, P, P, P]
        L.orc_render_f64.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P, P, P, P, P]
        L.orc_f32_to_f16.argtypes = [C.c_float]
        L.orc_f32_to_f16.restype = C.c_uint16
        L.orc_f32_to_bf16.argtypes = [C.c_float]
        L.orc_f32_to_bf16.restype = C.c_uint16
        L.orc_quantize_features.argtypes = [C.c_int, C.c_int, P, P, P, P]
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the oracle must be contiguous"
    return a.ctypes.data_as(C.c_void_p)


def camera(width, height, position, rotation, focal_x, focal_y, z_far=100.0, z_near=0.01,
           block_x=16, block_y=16) -> OrcCamera:
    c = OrcCamera()
    c.width, c.height, c.block_x, c.block_y = int(width), int(height), int(block_x), int(block_y)
    pos = np.asarray(position, dtype=np.float32).reshape(3)
    rot = np.asarray(rotation, dtype=np.float32).reshape(9)
    for i in range(3):
        c.position[i] = float(pos[i])
    for i in range(9):
        c.rotation[i] = float(rot[i])
    c.focal_x, c.focal_y, c.z_far, c.z_near = float(focal_x), float(focal_y), float(z_far), float(z_near)
    return c


def camera_from_spec(spec) -> OrcCamera:
    return camera(spec.width, spec.height, spec.position.numpy(), spec.rotation.numpy(),
                  spec.focal_x, spec.focal_y, spec.zFar, spec.zNear)


def camera_matrices(cam: OrcCamera):
    view = np.zeros(16, np.float32)
    proj = np.zeros(16, np.float32)
    tf = np.zeros(2, np.float32)
    lib().orc_camera_matrices(C.byref(cam), _p(view), _p(proj), _p(tf))
    return view, proj, tf


def log2(x: float) -> float:
    return float(lib().orc_log2(C.c_float(x)))


def cov2d(p, cov3, cam):
    out = np.zeros(3, np.float32)
    lib().orc_cov2d(_p(np.ascontiguousarray(p, np.float32)),
                    _p(np.ascontiguousarray(cov3, np.float32)), C.byref(cam), _p(out))
    return out


def sh_color(p, campos, sh):
    out = np.zeros(3, np.float32)
    lib().orc_sh_color(_p(np.ascontiguousarray(p, np.float32)),
                       _p(np.ascontiguousarray(campos, np.float32)),
                       _p(np.ascontiguousarray(sh, np.float32).reshape(48)), _p(out))
    return out


def tile_hits(tx, ty, cx, cy, ca, cb, cc, power, bx=16, by=16) -> bool:
    return bool(lib().orc_tile_hits(tx, ty, cx, cy, ca, cb, cc, power, bx, by))


def _sh_buffer(shs: np.ndarray, sh_type: str) -> np.ndarray:
    if sh_type == "fp32":
        return np.ascontiguousarray(shs, np.float32).reshape(-1, 48)
    # f16/bf16 pass their raw 16-bit patterns
    return np.ascontiguousarray(shs).view(np.uint16).reshape(-1, 48)


def preprocess(pos, shs, opac, cov3, cam: OrcCamera, sh_type="fp32", capacity=None):
    P = int(opac.shape[0])
    pos = np.ascontiguousarray(pos, np.float32)
    opac = np.ascontiguousarray(opac, np.float32)
    cov3 = np.ascontiguousarray(cov3, np.float32)
    shb = _sh_buffer(shs, sh_type)
    xy = np.zeros((P, 2), np.float32)
    rgbd = np.zeros((P, 4), np.float32)
    conic = np.zeros((P, 4), np.float32)
    counts = np.zeros(P, np.int32)
    L = lib()
    total = L.orc_preprocess(P, _p(pos), _p(shb), SH_TYPES[sh_type], _p(opac), _p(cov3),
                             C.byref(cam), _p(xy), _p(rgbd), _p(conic), _p(counts), None, None, 0)
    cap = total if capacity is None else capacity
    keys = np.zeros(max(int(cap), 1), np.uint64)
    vals = np.zeros(max(int(cap), 1), np.uint32)
    if total <= cap:
        L.orc_preprocess(P, _p(pos), _p(shb), SH_TYPES[sh_type], _p(opac), _p(cov3),
                         C.byref(cam), None, None, None, None, _p(keys), _p(vals), int(cap))
    return dict(M=int(total), keys=keys[:total], vals=vals[:total], xy=xy, rgbd=rgbd,
                conic=conic, counts=counts)


def sort_pairs(keys: np.ndarray, vals: np.ndarray):
    k = np.array(keys, np.uint64, copy=True)
    v = np.array(vals, np.uint32, copy=True)
    lib().orc_sort_pairs(len(k), _p(k), _p(v))
    return k, v


def tile_ranges(keys: np.ndarray, tiles: int, mode: int = 0) -> np.ndarray:
    r = np.zeros((tiles, 2), np.int32)
    k = np.ascontiguousarray(keys, np.uint64)
    lib().orc_tile_ranges(len(k), _p(k), tiles, _p(r), mode)
    return r


def render_f32(W, H, ranges, vals, xy, rgbd, conic, bg, check_every=8, dt=False):
    gx, gy = (W + 15) // 16, (H + 15) // 16
    out = np.zeros((H, W, 3), np.uint8)
    out_f = np.zeros((H, W, 3), np.float32)
    stop = np.zeros(gx * gy, np.int32)
    need = np.zeros(gx * gy, np.int32)
    bgv = np.asarray(bg, np.float32).reshape(3)
    lib().orc_render_f32(W, H, _p(np.ascontiguousarray(ranges, np.int32)),
                         _p(np.ascontiguousarray(vals, np.uint32)),
                         _p(np.ascontiguousarray(xy, np.float32)),
                         _p(np.ascontiguousarray(rgbd, np.float32)),
                         _p(np.ascontiguousarray(conic, np.float32)), _p(bgv), int(check_every),
                         1 if dt else 0, _p(out), _p(out_f), _p(stop), _p(need))
    return out, out_f, stop, need


def render_f64(W, H, ranges, vals, stop, pos, shs, opac, cov3, cam, bg, sh_type="fp32"):
    """FP64 blend with the FP32 run's tile lists and stop points."""
    vals = np.ascontiguousarray(vals, np.uint32)
    ids = np.unique(vals).astype(np.uint32)
    P = int(opac.shape[0])
    slot = np.full(P, -1, np.int32)
    slot[ids] = np.arange(len(ids), dtype=np.int32)
    n = len(ids)
    xy = np.zeros((max(n, 1), 2))
    con = np.zeros((max(n, 1), 4))
    rgb = np.zeros((max(n, 1), 3))
    shb = _sh_buffer(shs, sh_type)
    L = lib()
    L.orc_features_f64(n, _p(ids), _p(np.ascontiguousarray(pos, np.float32)), _p(shb),
                       SH_TYPES[sh_type], _p(np.ascontiguousarray(opac, np.float32)),
                       _p(np.ascontiguousarray(cov3, np.float32)), C.byref(cam), _p(xy), _p(con),
                       _p(rgb))
    out = np.zeros((H, W, 3), np.uint8)
    out_f = np.zeros((H, W, 3), np.float64)
    bgv = np.asarray(bg, np.float32).reshape(3)
    L.orc_render_f64(W, H, _p(np.ascontiguousarray(ranges, np.int32)), _p(vals),
                     _p(np.ascontiguousarray(stop, np.int32)), _p(slot), _p(xy), _p(con),
                     _p(rgb), _p(bgv), _p(out), _p(out_f))
    return out, out_f


def quantize_features(rgbd: np.ndarray, conic: np.ndarray, prec: str):
    P = rgbd.shape[0]
    rq = np.zeros_like(rgbd, dtype=np.float32)
    cq = np.zeros_like(conic, dtype=np.float32)
    lib().orc_quantize_features(P, PREC[prec], _p(np.ascontiguousarray(rgbd, np.float32)),
                                _p(np.ascontiguousarray(conic, np.float32)), _p(rq), _p(cq))
    return rq, cq


def morton(x: int, y: int) -> int:
    return int(lib().orc_morton(x, y))


def tile_order(ranges: np.ndarray, gx: int, zigzag: int = 0, bucket: int = 0):
    """(order, sorted keys, packed descriptors [T, 4]) of the reordered
    renderers for tile ranges [T, 2]."""
    rg = np.ascontiguousarray(ranges, dtype=np.int32).reshape(-1, 2)
    T = rg.shape[0]
    order = np.zeros(max(T, 1), np.uint32)
    keys = np.zeros(max(T, 1), np.uint32)
    desc = np.zeros((max(T, 1), 4), np.int32)
    lib().orc_tile_order(_p(rg), T, gx, zigzag, bucket, _p(order), _p(keys), _p(desc))
    return order[:T], keys[:T], desc[:T]


def tiles_of(W, H, bx=16, by=16):
    return ((W + bx - 1) // bx) * ((H + by - 1) // by)


def frame(pos, shs, opac, cov3, cam: OrcCamera, bg, sh_type="fp32", feat_prec="fp32",
          check_every=8, fp64=False, dt=False):
    """Whole FP32 forward pass; optionally the FP64 evaluation as well."""
    pre = preprocess(pos, shs, opac, cov3, cam, sh_type)
    ks, vs = sort_pairs(pre["keys"], pre["vals"])
    T = tiles_of(cam.width, cam.height, cam.block_x, cam.block_y)
    ranges = tile_ranges(ks, T, 0)
    rgbd, conic = pre["rgbd"], pre["conic"]
    if feat_prec != "fp32":
        rgbd, conic = quantize_features(rgbd, conic, feat_prec)
    img, img_f, stop, need = render_f32(cam.width, cam.height, ranges, vs, pre["xy"], rgbd,
                                        conic, bg, check_every, dt)
    res = dict(pre=pre, keys_sorted=ks, vals_sorted=vs, ranges=ranges, image=img, image_f=img_f,
               stop=stop, need=need)
    if fp64:
        res["image64"], res["image64_f"] = render_f64(cam.width, cam.height, ranges, vs, stop,
                                                      pos, shs, opac, cov3, cam, bg, sh_type)
    return res
