"""Scene / Camera / Rasterizer facade with the reference's Python surface.

Mirrors example.py:12-824 for the forward rasterizer variants this build
implements:

=============================================  ==========================================
render_variant                                 pipeline
=============================================  ==========================================
"default"                                      FP32 SH -> sort -> plain per-tile render
                                               (forward_fused, pybind.cpp:932)
"preranges_smem_persistent_lite_e2e"           FP16 SH -> sort -> persistent render with
                                               in-kernel range search (pybind.cpp:1085)
"preranges_smem_persistent_lite_e2e_preranges" FP16 SH -> sort -> ranges -> persistent
                                               render (pybind.cpp:1016)
"preranges_smem_persistent_lite[_dt]"          stage-by-stage from Python: FP16 SH
                                               preprocess, host count, sort, ranges,
                                               persistent render (example.py:480-781)
"reordered"                                    FP32 SH -> sort -> ranges -> tile order
                                               (heaviest first) -> plain render in that
                                               order (example.py:617-625, :698-704)
"reordered_v2", "reordered_zigzag<G>"          same with packed tile descriptors, groups
                                               of G positions zig-zagged (G = 132)
"reordered_persistent[_zigzag<G>]",            FP32 SH -> ... -> packed descriptors ->
"reordered_persistent_morton<B>[_zigzag<G>]",  persistent render over them (Morton-bucket
"reordered_persistent_adaptive"                keys of bucket B; example.py:626-645,
                                               :712-718, :234-243)
"preranges_smem_persistent[_morton..|_zigzag..  FP16 SH -> ... -> the same ordered
 |_adaptive]"                                  persistent render (example.py:326-359,
                                               :646-669)
"topk_smem", "topk_smem_persistent"           FP32 SH -> sort -> top-k cache of the most
                                               referenced Gaussians' features, values
                                               packed (bit 31 | slot) -> plain / persistent
                                               render (example.py:168-197, :676-697)
"topk_smem_persistent_v2"                      the persistent top-k renderer with no cache
"auto"                                         resolves to the fastest of the above
=============================================  ==========================================

``feature_dtype`` ("fp32", "fp16", "bf16") selects how the persistent
variants store per-Gaussian render features: FP32 in the reference's three
arrays, or 24-byte records with FP16/BF16 conic and colour (projection,
opacity and blending stay FP32).  The "default" variant always uses FP32.

``use_es`` selects the depth-first ("ES", early sorting) key pipeline of
example.py:246-291 for every variant here, fused or staged: the Gaussians
sorted by depth, their pairs emitted in that order as 32-bit tile ids and
stably sorted by tile id.  Images and sorted values are identical to the
default pipeline's; the sorted keys carry the tile id only (low word zero),
as the reference rebuilds them.  "auto" takes it when the previous frame's
pair count M is large in absolute terms and per Gaussian (M > T_M and
M > rho * P, thresholds measured on MI355X; environment variables
POLYSPLAT_ES_THRESHOLD_M and POLYSPLAT_ES_RATIO_THRESH override them,
POLYSPLAT_ES_FORCE_TWO_PASS=1 forces it), never on a first frame.

The reordered variants change only the order in which tiles are rendered;
their images are identical to the unordered renderers' (the persistent ones
render 24-byte records when ``feature_dtype`` is fp16/bf16, the plain ones
always FP32 features).  The top-k variants leave gaussian_values_sorted
packed, as the reference does.  Other variants of the reference (split
warps, other tile shapes) are not part of this build and raise ValueError.
"""
from __future__ import annotations

import os

import torch

from . import _C

ops = _C.ops

_PREC = {"fp32": 0, "fp16": 1, "bf16": 2}
VARIANT_IDS = {
    "default": 0,
    "preranges_smem_persistent_lite_e2e_preranges": 1,
    "preranges_smem_persistent_lite_e2e": 2,
}
STAGED = ("preranges_smem_persistent_lite", "preranges_smem_persistent_lite_dt")
TOPK = ("topk_smem", "topk_smem_persistent", "topk_smem_persistent_v2")
SMEM_TOPK_GAUSSIANS = 1365  # cached Gaussians of the top-k variants (example.py:9)
AUTO_CHOICE = "preranges_smem_persistent_lite_e2e_preranges"
DEPTH_FIRST = 0x100  # PS_FORWARD_DEPTH_FIRST: OR-ed into a ps_forward variant

# Depth-first pipeline selection ("auto").  On MI355X its extra per-Gaussian
# work (depth sort, counts and records gathered in depth order) costs more
# than the per-tile depth sort it removes unless frames hold many pairs per
# Gaussian: on 3840x2160 frames it is 23% slower at 1.3 pairs per Gaussian,
# 2% slower at 6.9 and level (within 2%) from 10 to 50.
ES_THRESHOLD_M = 2_000_000
ES_RATIO_THRESH = 10.0


# Tile count up to which "..._adaptive" takes Morton-bucket keys (bucket
# 1280, zig-zag 792) rather than count keys zig-zagged by 264
# (example.py:53, :234-243).
ADAPTIVE_TRUCK_TILE_THRESHOLD = 16384


def _int_or(text, default):
    return int(text) if text else default


def tile_order_plan(variant, tiles):
    """(renderer, zigzag group, Morton bucket) of a tile-reordering variant,
    or None: renderer "order" (plain over tile_order), "desc" (plain over
    packed descriptors) or "persistent" (persistent over packed
    descriptors).  Suffix grammar and defaults of example.py:304-389."""
    def adaptive():
        if tiles <= ADAPTIVE_TRUCK_TILE_THRESHOLD:
            return ("persistent", 792, 1280)
        return ("persistent", 264, 0)

    def morton_suffix(rest):
        if "_zigzag" in rest:
            b, g = rest.split("_zigzag", 1)
            return ("persistent", _int_or(g, 792), _int_or(b, 1280))
        return ("persistent", 0, _int_or(rest, 1280))

    if variant == "reordered":
        return ("order", 0, 0)
    if variant == "reordered_v2":
        return ("desc", 0, 0)
    if variant.startswith("reordered_zigzag"):
        return ("desc", _int_or(variant[len("reordered_zigzag"):], 132), 0)
    if variant == "reordered_persistent_adaptive":
        return adaptive()
    if variant.startswith("reordered_persistent"):
        rest = variant[len("reordered_persistent"):]
        if rest == "":
            return ("persistent", 0, 0)
        if rest.startswith("_zigzag"):
            return ("persistent", _int_or(rest[len("_zigzag"):], 264), 0)
        if rest.startswith("_morton"):
            return morton_suffix(rest[len("_morton"):])
        return None
    if variant.startswith("preranges_smem_persistent") and "_lite" not in variant:
        rest = variant[len("preranges_smem_persistent"):]
        if rest in ("", "_adaptive"):
            return adaptive()
        if rest.startswith("_zigzag"):
            return ("persistent", _int_or(rest[len("_zigzag"):], 264), 0)
        if rest.startswith("_morton"):
            return morton_suffix(rest[len("_morton"):])
    return None


def _env_flag(name):
    return os.environ.get(name, "0") not in ("0", "", "false", "False")


class Scene:
    """Gaussians resident on one device (example.py:12-30)."""

    def __init__(self, device):
        self.device = torch.device(device)
        self.num_vertex = 0
        self.position = None
        self.shs = None
        self.shs_half = None
        self.opacity = None
        self.cov3d = None
        self._shs_bf16 = None

    def loadPly(self, scene_path):
        """Loads an INRIA-layout PLY (pybind.cpp:58-120)."""
        n, pos, shs, op, cov = ops.loadPly(scene_path)
        self.set_arrays(pos, shs, op, cov)
        print("num_vertex = %d" % self.num_vertex)

    def set_arrays(self, position, shs, opacity, cov3d):
        self.num_vertex = int(opacity.shape[0])
        self.position = position.to(self.device, torch.float32).contiguous()
        self.shs = shs.reshape(self.num_vertex, 48).to(self.device, torch.float32).contiguous()
        self.shs_half = self.shs.half()
        self.opacity = opacity.to(self.device, torch.float32).contiguous()
        self.cov3d = cov3d.to(self.device, torch.float32).contiguous()
        self._shs_bf16 = None
        return self

    @property
    def shs_bf16(self):
        if self._shs_bf16 is None:
            self._shs_bf16 = self.shs.to(torch.bfloat16)
        return self._shs_bf16

    @classmethod
    def from_arrays(cls, arrays, device):
        s = cls(device)
        return s.set_arrays(arrays.position, arrays.shs, arrays.opacity, arrays.cov3d)


class Camera:
    """Camera from a cameras.json entry (example.py:33-44)."""

    def __init__(self, camera_json):
        self.id = camera_json["id"]
        self.img_name = camera_json["img_name"]
        self.width = camera_json["width"]
        self.height = camera_json["height"]
        self.position = torch.tensor(camera_json["position"], dtype=torch.float32)
        self.rotation = torch.tensor(camera_json["rotation"], dtype=torch.float32)
        self.focal_x = camera_json["fx"]
        self.focal_y = camera_json["fy"]
        self.zFar = 100.0
        self.zNear = 0.01


class Rasterizer:
    """Owns every per-frame device buffer (example.py:56-139); forward()
    allocates nothing except the output image on a resolution change."""

    def __init__(self, scene, MAX_NUM_RENDERED, MAX_NUM_TILES, enable_es=True,
                 feature_dtype="fp32"):
        if feature_dtype not in _PREC:
            raise ValueError(f"feature_dtype must be one of {sorted(_PREC)}")
        dev = scene.device
        self.device = dev
        self.MAX_NUM_RENDERED = int(MAX_NUM_RENDERED)
        self.MAX_NUM_TILES = int(MAX_NUM_TILES)
        self.feature_dtype = feature_dtype
        self.enable_es = enable_es
        self._es_last_M = None  # previous frame's pair count, for "auto"
        self._es_threshold_M = int(os.environ.get("POLYSPLAT_ES_THRESHOLD_M", ES_THRESHOLD_M))
        self._es_ratio_thresh = float(os.environ.get("POLYSPLAT_ES_RATIO_THRESH",
                                                     ES_RATIO_THRESH))
        self._es_force_two_pass = _env_flag("POLYSPLAT_ES_FORCE_TWO_PASS")
        M = self.MAX_NUM_RENDERED
        P = scene.num_vertex
        self.gaussian_keys_unsorted = torch.zeros(M, device=dev, dtype=torch.int64)
        self.gaussian_values_unsorted = torch.zeros(M, device=dev, dtype=torch.int32)
        self.gaussian_keys_sorted = torch.zeros(M, device=dev, dtype=torch.int64)
        self.gaussian_values_sorted = torch.zeros(M, device=dev, dtype=torch.int32)
        self.SORT_BUFFER_SIZE = ops.get_sort_buffer_size(M)
        self.list_sorting_space = torch.zeros(self.SORT_BUFFER_SIZE, device=dev, dtype=torch.int8)
        self.curr_offset = torch.zeros(1, device=dev, dtype=torch.int32)
        self.points_xy = torch.zeros((P, 2), device=dev, dtype=torch.float32)
        self.rgb_depth = torch.zeros((P, 4), device=dev, dtype=torch.float32)
        self.conic_opacity = torch.zeros((P, 4), device=dev, dtype=torch.float32)
        self.records = torch.zeros(24 * P if feature_dtype != "fp32" else 0, device=dev,
                                   dtype=torch.uint8)
        self.tile_ranges_buf = torch.zeros((self.MAX_NUM_TILES, 2), device=dev, dtype=torch.int32)
        # Zero-filled once: the persistent queue words inside must start at 0
        # and every launch leaves them at 0 again.
        self.workspace = torch.zeros(ops.forward_workspace_size(P), device=dev, dtype=torch.uint8)
        self._empty = torch.zeros(0, device=dev, dtype=torch.uint8)
        self._out_color = None
        # The depth-first pipeline's 32-bit tile ids live in the unsorted
        # keys' memory: unsorted ids in its first half, sorted in its second.
        tk = self.gaussian_keys_unsorted.view(torch.int32)
        self._es_tile_keys_unsorted = tk[:M]
        self._es_tile_keys_sorted = tk[M:]
        # Tile reordering buffers (example.py:128-136).
        T = self.MAX_NUM_TILES
        self.tile_order = torch.zeros(T, device=dev, dtype=torch.int32)
        self.tile_counts_buf = torch.zeros(T, device=dev, dtype=torch.int32)
        self.tile_ids_buf = torch.zeros(T, device=dev, dtype=torch.int32)
        self.tile_order_sort_temp = torch.zeros(ops.get_tile_order_sort_temp_size(T), device=dev,
                                                dtype=torch.int8)
        self.tile_desc_buf = torch.zeros((T, 4), device=dev, dtype=torch.int32)
        # Top-k feature cache (example.py:120-123).
        self.smem_topk_lookup = torch.full((max(P, 1),), -1, device=dev, dtype=torch.int16)
        self.smem_topk_xy = torch.zeros((SMEM_TOPK_GAUSSIANS, 2), device=dev, dtype=torch.float32)
        self.smem_topk_rgb_depth = torch.zeros((SMEM_TOPK_GAUSSIANS, 4), device=dev,
                                               dtype=torch.float32)
        self.smem_topk_conic = torch.zeros((SMEM_TOPK_GAUSSIANS, 4), device=dev,
                                           dtype=torch.float32)

    def _get_out_color(self, height, width, device):
        if (self._out_color is None or self._out_color.shape[0] != height
                or self._out_color.shape[1] != width):
            self._out_color = torch.zeros((height, width, 3), device=device, dtype=torch.uint8)
        return self._out_color

    def _check_tiles(self, camera, block_x, block_y):
        tiles = ((camera.width + block_x - 1) // block_x) * ((camera.height + block_y - 1) // block_y)
        if tiles > self.MAX_NUM_TILES:
            raise ValueError(f"{tiles} tiles exceed MAX_NUM_TILES={self.MAX_NUM_TILES}")

    def _shs_for(self, scene, half):
        if not half:
            return scene.shs
        return scene.shs_bf16 if self.feature_dtype == "bf16" else scene.shs_half

    def _decide_use_es(self, scene, use_es):
        """Depth-first pipeline or not (example.py:246-291, with this
        build's own cost model: see the module docstring)."""
        if use_es not in (True, False, "auto"):
            raise ValueError(f"use_es must be True, False or 'auto' (got {use_es!r})")
        if not self.enable_es:
            return False
        if use_es is True or use_es is False:
            return use_es
        if self._es_force_two_pass:
            return True
        if self._es_last_M is None:
            return False
        M = self._es_last_M
        return M > self._es_threshold_M and M > self._es_ratio_thresh * scene.num_vertex

    def forward(self, scene, camera, bg_color, block_x=16, block_y=16, render_variant="default",
                measure_preprocess=False, measure_sort=False, measure_render=False,
                return_stats=False, use_es="auto"):
        requested = render_variant
        if render_variant == "auto":
            render_variant = AUTO_CHOICE
        if block_x != 16 or block_y != 16:
            raise ValueError("only 16x16 tiles are implemented")
        measuring = measure_preprocess or measure_sort or measure_render
        if render_variant.startswith("preranges_smem_persistent_lite_e2e") and measuring:
            # E2E calls cannot expose per-stage events (example.py:362-368).
            render_variant = "preranges_smem_persistent_lite"
        tiles = ((camera.width + 15) // 16) * ((camera.height + 15) // 16)
        order_plan = tile_order_plan(render_variant, tiles)
        if (render_variant not in VARIANT_IDS and render_variant not in STAGED
                and render_variant not in TOPK and order_plan is None):
            raise ValueError(f"render_variant {render_variant!r} is not implemented")
        self._check_tiles(camera, block_x, block_y)
        es = self._decide_use_es(scene, use_es)
        out_color = self._get_out_color(camera.height, camera.width, scene.device)
        times = {"preprocess_ms": None, "sort_ms": None, "render_ms": None}
        if order_plan is not None or render_variant in TOPK:
            num_rendered = self._forward_staged(scene, camera, bg_color, render_variant,
                                                out_color, measure_preprocess, measure_sort,
                                                measure_render, times, es, order_plan)
        elif render_variant in VARIANT_IDS and not (render_variant == "default" and measuring):
            num_rendered = self._forward_fused(scene, camera, bg_color, render_variant, out_color,
                                               es)
        else:
            num_rendered = self._forward_staged(scene, camera, bg_color, render_variant,
                                                out_color, measure_preprocess, measure_sort,
                                                measure_render, times, es)
        self._es_last_M = num_rendered
        if return_stats:
            return out_color, {
                "requested_render_variant": requested,
                "render_variant": render_variant,
                "resolved_block_x": block_x,
                "resolved_block_y": block_y,
                "num_rendered": num_rendered,
                "preprocess_ms": times["preprocess_ms"],
                "sort_ms": times["sort_ms"],
                "render_ms": times["render_ms"],
                "used_es": es,
            }
        sel = [times[k] for k, f in (("preprocess_ms", measure_preprocess),
                                     ("sort_ms", measure_sort),
                                     ("render_ms", measure_render)) if f]
        if sel:
            return (out_color, *sel)
        return out_color

    def _forward_fused(self, scene, camera, bg_color, render_variant, out_color, es=False):
        variant = VARIANT_IDS[render_variant]
        fp32_path = render_variant == "default"
        prec = 0 if fp32_path else _PREC[self.feature_dtype]
        shs = self._shs_for(scene, half=not fp32_path)
        ranges = self.tile_ranges_buf if variant == 1 else self._empty
        records = self.records if prec else self._empty
        return ops.forward_frame(
            scene.position, shs, scene.opacity, scene.cov3d,
            camera.width, camera.height, camera.position, camera.rotation,
            camera.focal_x, camera.focal_y, camera.zFar, camera.zNear,
            variant | (DEPTH_FIRST if es else 0), prec, self.curr_offset, self.list_sorting_space,
            self.gaussian_keys_unsorted, self.gaussian_values_unsorted,
            self.gaussian_keys_sorted, self.gaussian_values_sorted,
            self.points_xy, self.rgb_depth, self.conic_opacity, records, ranges,
            bg_color, out_color, self.workspace, True)

    def _forward_staged(self, scene, camera, bg_color, render_variant, out_color,
                        m_pre, m_sort, m_render, times, es=False, order_plan=None):
        ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
        # The plain renderers (default, and reordered over a tile order or
        # descriptors) read FP32 features; the reordered variants take FP32
        # SH (example.py:524-531), the preranges_smem_persistent ones FP16.
        fp32_path = (render_variant == "default" or render_variant in TOPK
                     or (order_plan is not None and order_plan[0] != "persistent"))
        prec = 0 if fp32_path else _PREC[self.feature_dtype]
        half_sh = not fp32_path and (order_plan is None
                                     or render_variant.startswith("preranges_smem_persistent"))
        shs = self._shs_for(scene, half=half_sh)
        args_cam = (camera.width, camera.height, 16, 16, camera.position, camera.rotation,
                    camera.focal_x, camera.focal_y, camera.zFar, camera.zNear)
        self.curr_offset.fill_(0)
        e = [ev(), ev(), ev(), ev(), ev(), ev()]
        if m_pre:
            e[0].record()
        if es:
            # count pass, depth sort and pairs in depth order
            # (example.py:494-522 and the first half of 550-576)
            ops.preprocess_es(scene.position, shs, scene.opacity, scene.cov3d, *args_cam,
                              self.points_xy, self.rgb_depth, self.conic_opacity, self.records,
                              prec, self._es_tile_keys_unsorted, self.gaussian_values_unsorted,
                              self.curr_offset, self.list_sorting_space)
        elif prec == 0:
            ops.preprocess(scene.position, shs, scene.opacity, scene.cov3d, *args_cam,
                           self.points_xy, self.rgb_depth, self.conic_opacity,
                           self.gaussian_keys_unsorted, self.gaussian_values_unsorted,
                           self.curr_offset)
        else:
            ops.preprocess_records(scene.position, shs, scene.opacity, scene.cov3d, *args_cam,
                                   self.records, prec, self.gaussian_keys_unsorted,
                                   self.gaussian_values_unsorted, self.curr_offset)
        if m_pre:
            e[1].record()
        num_rendered = int(self.curr_offset.cpu()[0])
        if num_rendered > self.MAX_NUM_RENDERED:
            raise RuntimeError("Too many k-v pairs!")
        if m_sort:
            e[2].record()
        if es:
            ops.es_tile_sort(num_rendered, camera.width, camera.height, 16, 16,
                             self._es_tile_keys_unsorted, self.gaussian_values_unsorted,
                             self._es_tile_keys_sorted, self.gaussian_values_sorted,
                             self.gaussian_keys_sorted, self.list_sorting_space)
        else:
            ops.sort_gaussian(num_rendered, camera.width, camera.height, 16, 16,
                              self.list_sorting_space, self.gaussian_keys_unsorted,
                              self.gaussian_values_unsorted, self.gaussian_keys_sorted,
                              self.gaussian_values_sorted)
        if m_sort:
            e[3].record()
        if (not fp32_path or order_plan is not None) and render_variant not in TOPK:
            ops.precompute_tile_ranges(num_rendered, camera.width, camera.height, 16, 16,
                                       self.gaussian_keys_sorted, self.tile_ranges_buf)
        if order_plan is not None:
            self._tile_order(camera, order_plan)
        num_topk = 0
        if render_variant in ("topk_smem", "topk_smem_persistent"):
            num_topk = self._prepare_topk_smem_cache(scene, num_rendered)
        if m_render:
            e[4].record()
        if render_variant in TOPK:
            fn = {"topk_smem": ops.render_16x16_topk_smem,
                  "topk_smem_persistent": ops.render_16x16_topk_smem_persistent,
                  "topk_smem_persistent_v2": ops.render_16x16_topk_smem_persistent_v2}
            fn[render_variant](num_rendered, camera.width, camera.height, self.points_xy,
                               self.rgb_depth, self.conic_opacity, self.gaussian_keys_sorted,
                               self.gaussian_values_sorted, self.smem_topk_xy,
                               self.smem_topk_rgb_depth, self.smem_topk_conic, num_topk,
                               bg_color, out_color)
        elif order_plan is not None:
            self._render_reordered(num_rendered, camera, bg_color, out_color, order_plan, prec,
                                   render_variant.endswith("_dt"))
        elif fp32_path:
            ops.render_16x16(num_rendered, camera.width, camera.height, self.points_xy,
                             self.rgb_depth, self.conic_opacity, self.gaussian_keys_sorted,
                             self.gaussian_values_sorted, bg_color, out_color)
        elif prec:
            ops.render_records(num_rendered, camera.width, camera.height, self.records, prec,
                               False, self.gaussian_values_sorted, self.tile_ranges_buf,
                               bg_color, out_color, render_variant.endswith("_dt"))
        elif render_variant.endswith("_dt"):
            ops.render_16x16_preranges_smem_persistent_lite_dt(
                num_rendered, camera.width, camera.height, self.points_xy, self.rgb_depth,
                self.conic_opacity, self.gaussian_values_sorted, self.tile_ranges_buf,
                bg_color, out_color)
        else:
            ops.render_16x16_preranges_smem_persistent_lite(
                num_rendered, camera.width, camera.height, self.points_xy, self.rgb_depth,
                self.conic_opacity, self.gaussian_values_sorted, self.tile_ranges_buf,
                bg_color, out_color)
        if m_render:
            e[5].record()
        if m_pre or m_sort or m_render:
            torch.cuda.synchronize()
        if m_pre:
            times["preprocess_ms"] = e[0].elapsed_time(e[1])
        if m_sort:
            times["sort_ms"] = e[2].elapsed_time(e[3])
        if m_render:
            times["render_ms"] = e[4].elapsed_time(e[5])
        return num_rendered

    def _tile_order(self, camera, plan):
        """Tile order or packed descriptors of the frame (example.py:617-669)."""
        kind, zigzag, bucket = plan
        gx = (camera.width + 15) // 16
        tiles = gx * ((camera.height + 15) // 16)
        if kind == "order":
            ops.compute_tile_order(self.tile_ranges_buf, tiles, self.tile_order,
                                   self.tile_counts_buf, self.tile_ids_buf,
                                   self.tile_order_sort_temp)
        elif bucket > 0:
            ops.compute_tile_order_packed_morton(self.tile_ranges_buf, tiles, gx,
                                                 self.tile_desc_buf, self.tile_order,
                                                 self.tile_counts_buf, self.tile_ids_buf,
                                                 self.tile_order_sort_temp, zigzag, bucket)
        else:
            ops.compute_tile_order_packed(self.tile_ranges_buf, tiles, gx, self.tile_desc_buf,
                                          self.tile_order, self.tile_counts_buf,
                                          self.tile_ids_buf, self.tile_order_sort_temp, zigzag)

    def _render_reordered(self, num_rendered, camera, bg_color, out_color, plan, prec, dt):
        kind = plan[0]
        W, H = camera.width, camera.height
        if kind == "order":
            ops.render_16x16_reordered(num_rendered, W, H, self.points_xy, self.rgb_depth,
                                       self.conic_opacity, self.gaussian_keys_sorted,
                                       self.gaussian_values_sorted, self.tile_ranges_buf,
                                       self.tile_order, bg_color, out_color)
        elif kind == "desc":
            ops.render_16x16_reordered_v2(num_rendered, W, H, self.points_xy, self.rgb_depth,
                                          self.conic_opacity, self.gaussian_keys_sorted,
                                          self.gaussian_values_sorted, self.tile_desc_buf,
                                          bg_color, out_color)
        elif prec:
            ops.render_records_reordered(num_rendered, W, H, self.records, prec,
                                         self.gaussian_values_sorted, self.tile_desc_buf,
                                         bg_color, out_color, dt)
        else:
            ops.render_16x16_reordered_persistent(num_rendered, W, H, self.points_xy,
                                                  self.rgb_depth, self.conic_opacity,
                                                  self.gaussian_keys_sorted,
                                                  self.gaussian_values_sorted,
                                                  self.tile_desc_buf, bg_color, out_color)

    def _prepare_topk_smem_cache(self, scene, num_rendered):
        """The SMEM_TOPK_GAUSSIANS Gaussians referenced by the most pairs:
        their FP32 features copied to the compact arrays, and every sorted
        value naming one of them replaced by 0x80000000 | its slot
        (example.py:168-197).  Which of equally referenced Gaussians are
        cached does not change the image."""
        num_topk = min(SMEM_TOPK_GAUSSIANS, scene.num_vertex)
        if num_topk <= 0 or num_rendered <= 0:
            return 0
        rendered_ids = self.gaussian_values_sorted[:num_rendered].to(torch.int64)
        counts = torch.bincount(rendered_ids, minlength=scene.num_vertex)
        topk_ids = torch.topk(counts, k=num_topk, largest=True, sorted=False).indices
        self.smem_topk_lookup.fill_(-1)
        self.smem_topk_lookup[topk_ids] = torch.arange(num_topk, device=scene.device,
                                                       dtype=torch.int16)
        vals = self.gaussian_values_sorted[:num_rendered]
        slots = self.smem_topk_lookup[vals.to(torch.int64)].to(torch.int32)
        hit = slots >= 0
        packed = vals.clone()
        packed[hit] = slots[hit] | (-(1 << 31))
        vals.copy_(packed)
        self.smem_topk_xy[:num_topk].copy_(torch.index_select(self.points_xy, 0, topk_ids))
        self.smem_topk_rgb_depth[:num_topk].copy_(torch.index_select(self.rgb_depth, 0, topk_ids))
        self.smem_topk_conic[:num_topk].copy_(torch.index_select(self.conic_opacity, 0, topk_ids))
        return num_topk
