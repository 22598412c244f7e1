/*
 * polysplat_amd: C ABI of the MI355X (gfx950) forward 3D Gaussian Splatting
 * rasterizer.
 *
 * Every entry point takes plain device pointers, sizes and a HIP stream, and
 * returns a status code.  Buffers are owned by the caller.  Nothing here
 * synchronises the stream except ps_fetch_num_rendered and
 * ps_forward(sync_count != 0).  Internal state lives only in caller-provided
 * workspaces, so concurrent frames on different streams are safe when each
 * has its own workspaces.
 *
 * Reference interface each function replaces is cited as file:line in
 * LuoLongZan/PolySplat.
 */
#ifndef POLYSPLAT_AMD_H
#define POLYSPLAT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same type as hipStream_t; NULL is the default stream. */
typedef struct ihipStream_t* ps_stream_t;

/* Status codes. */
enum {
  PS_OK = 0,
  PS_ERR_INVALID = 1,     /* bad argument (message in ps_last_error) */
  PS_ERR_HIP = 2,         /* HIP runtime error */
  PS_ERR_OVERFLOW = 3,    /* more (tile, Gaussian) pairs than the key buffers hold */
  PS_ERR_UNSUPPORTED = 4  /* configuration not implemented (e.g. non-16x16 render) */
};

/* Storage precision of SH coefficients and of the render feature records. */
enum { PS_PREC_F32 = 0, PS_PREC_F16 = 1, PS_PREC_BF16 = 2 };

/* Render variants understood by ps_forward (example.py:293-466). */
enum {
  PS_VARIANT_DEFAULT = 0,           /* "default": FP32 SH, plain per-tile render */
  PS_VARIANT_PERSISTENT_LITE = 1,   /* "preranges_smem_persistent_lite[_e2e_preranges]" */
  PS_VARIANT_PERSISTENT_FUSED = 2,  /* "preranges_smem_persistent_lite_e2e": in-kernel range search */
  PS_VARIANT_PERSISTENT_DT = 3      /* "preranges_smem_persistent_lite_dt": skip alpha < 1/255 */
};

/* OR'ed into ps_forward's variant: the depth-first ("early sorting") key
 * pipeline of preprocess_es_pass_a/_b, es_depth_sort_and_scan and
 * es_tile_sort (preprocess.cu:673-1248, sort.cu:151-285, selected in
 * example.py:246-291): the Gaussians are sorted by depth first, their
 * (tile, Gaussian) pairs emitted in that order as 32-bit tile ids and sorted
 * stably by tile id, so no per-tile depth sort is needed.  It gives the same
 * sorted keys and values, tile ranges and image as the default pipeline;
 * keys_unsorted then holds 32-bit tile ids (emitted and sorted) instead of
 * 64-bit keys.  It pays off when tiles hold long lists (many pairs per
 * tile), where the per-tile depth sort is slow. */
#define PS_FORWARD_DEPTH_FIRST 0x100
/* OR-ed into a ps_forward variant with tile ranges (PERSISTENT_LITE or _DT):
 * the persistent renderer takes the tiles heaviest first (the order of
 * ps_compute_tile_order, dealt interleaved to the XCD queues), computed on
 * the device after the sort in the sort space, which must then hold
 * ps_forward_tile_order_space(tiles) bytes.  Same image; pays on frames of
 * many long tile lists (about 1000 pairs per tile and more), where a heavy
 * tile started late sets the render time. */
#define PS_FORWARD_TILE_ORDER 0x200

/* Camera exactly as the reference bindings receive it (pybind.cpp:122-150):
 * rotation holds the 9 values of cameras.json "rotation" row by row. */
typedef struct ps_camera {
  int width, height;
  int block_x, block_y;
  float position[3];
  float rotation[9];
  float focal_x, focal_y;
  float z_far, z_near;
} ps_camera;

const char* ps_status_string(int status);
/* Detail of the last error on the calling thread ("" if none). */
const char* ps_last_error(void);
/* Library version string. */
const char* ps_version(void);

/* View/projection matrices (column-major) and tan(fov/2); host only.
 * Replaces getViewMatrix/getProjectionMatrix, preprocess.cu:643-671 and the
 * tan_fov lines at preprocess.cu:1263-1266. */
int ps_camera_matrices(const ps_camera* cam, float view[16], float proj[16], float tan_fov[2]);

/* ---------------------------------------------------------------- preprocess
 * Replaces preprocess (preprocess.cu:1252, ops.h:42) and preprocess_half_sh
 * (preprocess.cu:1292, ops.h:110); bound in the reference at
 * pybind.cpp:122-181.
 *
 * positions f32[P][3], shs [P][48] in sh_prec, opacities f32[P],
 * cov3d f32[P][6].  Outputs (each may be NULL): points_xy f32[P][2],
 * rgb_depth f32[P][4], conic_opacity f32[P][4], and feature records for the
 * reduced-precision renderer (24 bytes per Gaussian, feat_prec F16/BF16).
 * Features are written for every Gaussian that emits at least one key.
 *
 * Keys u64[capacity] (tile << 32 | depth bits) and values u32[capacity] are
 * appended at *curr_offset in a canonical order (Gaussian id, then tile
 * row-major); *curr_offset is advanced by the number of pairs.  If the pairs
 * do not fit (or *curr_offset is negative), no pair is written at or past
 * `capacity` (pairs below it may be), *curr_offset still receives the
 * required total and bit 0 of the status word is set (the device pointer
 * from ps_preprocess_status_ptr).  workspace: ps_preprocess_workspace_size(P)
 * bytes; no part of it needs clearing. */
size_t ps_preprocess_workspace_size(int P);
int ps_preprocess(int P, const float* positions, const void* shs, int sh_prec,
                  const float* opacities, const float* cov3d, const ps_camera* cam,
                  float* points_xy, float* rgb_depth, float* conic_opacity,
                  void* feat_records, int feat_prec,
                  uint64_t* keys_unsorted, uint32_t* values_unsorted, int capacity,
                  int* curr_offset, void* workspace, size_t workspace_bytes,
                  ps_stream_t stream);
/* Device pointer to the pair count the preprocess left for the later stages
 * (0 after an overflow), and to the overflow status word. */
const int* ps_preprocess_count_ptr(const void* workspace);
const int* ps_preprocess_status_ptr(const void* workspace);

/* ---------------------------------------------------------------- sort
 * Stable LSD radix sort of (u64 key, u32 value) pairs over bits
 * [0, 32 + bit_length(tiles)).  Replaces sort_gaussian (sort.cu:107-125,
 * ops.h:50) and get_sort_buffer_size (sort.cu:127-149, ops.h:56); bound at
 * pybind.cpp:333-349.  The pair count comes from num_rendered when
 * d_num_rendered is NULL, else from device memory (no host round trip).
 * Inputs are preserved; outputs must not alias inputs. */
size_t ps_get_sort_buffer_size(int max_pairs);
int ps_sort_gaussian(int num_rendered, const int* d_num_rendered, int max_pairs,
                     int width, int height, int block_x, int block_y,
                     void* sort_space, size_t sort_space_bytes,
                     const uint64_t* keys_unsorted, const uint32_t* values_unsorted,
                     uint64_t* keys_sorted, uint32_t* values_sorted, ps_stream_t stream);

/* ---------------------------------------------------------------- depth first
 * The depth-first ("ES", early sorting) key pipeline in two calls, for
 * callers that time or drive the stages themselves (ps_forward runs the same
 * kernels under PS_FORWARD_DEPTH_FIRST).
 *
 * ps_preprocess_depth_first replaces preprocess_es_pass_a[_half_sh]
 * (preprocess.cu:673-1040), es_depth_sort_and_scan (sort.cu:217-253) and
 * preprocess_es_pass_b (preprocess.cu:1042-1248); bound at
 * pybind.cpp:206-304.  Count pass (features and records as ps_preprocess
 * writes them), a stable sort of the Gaussians by depth, then every
 * (tile id, Gaussian) pair in (depth, Gaussian id) order, each Gaussian's
 * tiles row-major: tile ids u32[capacity] and values u32[capacity] from
 * offset 0.  *curr_offset receives the pair count; on overflow the same
 * rules as ps_preprocess hold.  workspace: ps_preprocess_workspace_size(P);
 * sort_space: ps_get_sort_buffer_size(capacity) bytes (the depth sort's
 * digit counts).
 *
 * ps_es_tile_sort replaces es_tile_sort (sort.cu:256-285); bound at
 * pybind.cpp:306-327.  Stable sort of those pairs by tile id, then
 * keys_sorted u64 = tile id << 32 (the low word zero, as the reference
 * rebuilds them) and, when tile_ranges is not NULL, the tile ranges (empty
 * tile -> (0, 0)).  The values come out in the canonical (tile, depth,
 * Gaussian id) order.  The pair count is num_rendered, or *d_num_rendered
 * (capped at max_pairs) when that is not NULL. */
int ps_preprocess_depth_first(int P, const float* positions, const void* shs, int sh_prec,
                              const float* opacities, const float* cov3d, const ps_camera* cam,
                              float* points_xy, float* rgb_depth, float* conic_opacity,
                              void* feat_records, int feat_prec, uint32_t* tile_ids,
                              uint32_t* values, int capacity, int* curr_offset, void* workspace,
                              size_t workspace_bytes, void* sort_space, size_t sort_space_bytes,
                              ps_stream_t stream);
int ps_es_tile_sort(int num_rendered, const int* d_num_rendered, int max_pairs, int width,
                    int height, int block_x, int block_y, void* sort_space,
                    size_t sort_space_bytes, const uint32_t* tile_ids_unsorted,
                    const uint32_t* values_unsorted, uint32_t* tile_ids_sorted,
                    uint32_t* values_sorted, uint64_t* keys_sorted, int* tile_ranges,
                    ps_stream_t stream);

/* Copies *d_curr_offset to the host through pinned memory and waits for it.
 * Replaces fetch_num_rendered_async (sort.cu:289-300, ops.h:103). */
int ps_fetch_num_rendered(const int* d_curr_offset, int* out_num_rendered, ps_stream_t stream);

/* ---------------------------------------------------------------- ranges
 * tile_ranges int32[tiles][2] = [start, end) of each tile in the sorted keys.
 * precompute_tile_ranges (render.cu:6812, binary search: empty tile ->
 * (lo, lo)) and precompute_tile_ranges_scan (render.cu:6826, boundary scan:
 * empty tile -> (0, 0)); bound at pybind.cpp:441-459. */
int ps_precompute_tile_ranges(int num_rendered, const int* d_num_rendered,
                              int width, int height, int block_x, int block_y,
                              const uint64_t* keys_sorted, int* tile_ranges, ps_stream_t stream);
int ps_precompute_tile_ranges_scan(int num_rendered, const int* d_num_rendered, int max_pairs,
                                   int width, int height, int block_x, int block_y,
                                   const uint64_t* keys_sorted, int* tile_ranges,
                                   ps_stream_t stream);

/* ---------------------------------------------------------------- gather
 * dst[i] = src[values_sorted[i]] for xy f32x2, rgb_depth f32x4,
 * conic_opacity f32x4.  Replaces gather_features (gather.cu:46-58, ops.h:105);
 * bound at pybind.cpp:351-364. */
int ps_gather_features(int num_rendered, const int* d_num_rendered, int max_pairs,
                       const uint32_t* values_sorted,
                       const float* src_xy, const float* src_rgb_depth, const float* src_conic,
                       float* dst_xy, float* dst_rgb_depth, float* dst_conic, ps_stream_t stream);
/* Same for the 24-byte reduced-precision records. */
int ps_gather_records(int num_rendered, const int* d_num_rendered, int max_pairs,
                      const uint32_t* values_sorted, const void* src_records, void* dst_records,
                      ps_stream_t stream);

/* ---------------------------------------------------------------- render
 * Output u8[height][width][3].  bg: 3 floats on the host.
 *
 * ps_render_16x16: plain per-tile renderer, one wave per 16x16 tile, range
 * found by an in-kernel search over keys_sorted.  Replaces render_16x16
 * (render.cu:6445 -> renderCUDA render.cu:2176-2621); bound at
 * pybind.cpp:616-632.
 *
 * ps_render_16x16_persistent: persistent tile-queue renderer with LDS
 * double buffering.  Ranges come from tile_ranges (keys_sorted may be NULL),
 * or, when tile_ranges is NULL, from an in-kernel search of keys_sorted.
 * Replaces render_16x16_preranges_smem_persistent_lite (render.cu:6537),
 * ..._dt (render.cu:6576) and ..._fused (render.cu:6616); bound at
 * pybind.cpp:691-745.  feat_prec selects f32 arrays (points_xy, rgb_depth,
 * conic_opacity) or 24-byte records (records).  `gathered` != 0 means the
 * features are already in sorted-pair order (index = pair position).
 * queue: 2 KiB of device memory (the tile queue's claim counters), zero
 * before the first use; every launch leaves it zero again.  dt != 0 skips
 * pixels with alpha < 1/255. */
int ps_render_16x16(int num_rendered, const int* d_num_rendered, int width, int height,
                    const float* points_xy, const float* rgb_depth, const float* conic_opacity,
                    const uint64_t* keys_sorted, const uint32_t* values_sorted,
                    const float bg[3], uint8_t* out_color, ps_stream_t stream);
int ps_render_16x16_persistent(int num_rendered, const int* d_num_rendered, int width, int height,
                               int feat_prec, const float* points_xy, const float* rgb_depth,
                               const float* conic_opacity, const void* records, int gathered,
                               const uint64_t* keys_sorted, const uint32_t* values_sorted,
                               const int* tile_ranges, int dt, const float bg[3],
                               uint8_t* out_color, unsigned* queue, ps_stream_t stream);

/* Same as ps_render_16x16_persistent without the `queue` argument: the tile
 * queue lives in 2 KiB of static device memory the library holds on each
 * device for the life of the process, so the call allocates nothing and
 * needs no workspace from the caller.  No counterpart in the reference (whose
 * renderer keeps a __device__ global counter, render.cu:1537, that callers
 * cannot reset); it is for callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops.  The queue needs no
 * more room for larger frames, so every size ps_render_16x16_persistent
 * accepts is accepted.  The queue is cleared on the stream before each
 * render (by a kernel, so the call can be captured).  All calls on one
 * device share that one queue: calls that overlap on a device race and must
 * be serialised by the caller (one stream, or events between streams). */
int ps_render_16x16_persistent_noalloc(int num_rendered, const int* d_num_rendered, int width,
                                       int height, int feat_prec, const float* points_xy,
                                       const float* rgb_depth, const float* conic_opacity,
                                       const void* records, int gathered,
                                       const uint64_t* keys_sorted, const uint32_t* values_sorted,
                                       const int* tile_ranges, int dt, const float bg[3],
                                       uint8_t* out_color, ps_stream_t stream);

/* ------------------------------------------------------- tile reordering
 * Tile order (render.cu:4411-4520, :6959-7100): the tiles sorted by their
 * pair count tile_ranges[t][1] - tile_ranges[t][0], descending, ties in
 * ascending tile id (the reference's stable descending radix sort).
 * Outputs: tile_order[p] = tile at position p; tile_counts[p] = the sort
 * key at position p; tile_ids[p] = p (may be NULL).  sort_temp:
 * ps_get_tile_order_sort_temp_size(total_tiles) bytes of device memory.
 * Replaces compute_tile_order (ops.h:258-266) and
 * get_tile_order_sort_temp_size (ops.h:268); bound at pybind.cpp:461-482. */
size_t ps_get_tile_order_sort_temp_size(int total_tiles);
int ps_compute_tile_order(const int* tile_ranges, int total_tiles, uint32_t* tile_order,
                          uint32_t* tile_counts, uint32_t* tile_ids, void* sort_temp,
                          size_t sort_temp_bytes, ps_stream_t stream);
/* Packed form: tile_desc[p] = (column, row, range start, range end) of the
 * tile at order position zigzag(p) (groups of zigzag_group_size positions,
 * every odd group reversed; 0 or 1 = no regrouping, render.cu:4838-4866).
 * morton_bucket_size > 0 sorts by (min(count / bucket, 0xffff) << 16) |
 * (0xffff - (Morton(column, row) & 0xffff)) instead of the count
 * (render.cu:4496-4518).  Replaces compute_tile_order_packed (ops.h:277-288)
 * and compute_tile_order_packed_morton (ops.h:290-302); bound at
 * pybind.cpp:506-554. */
int ps_compute_tile_order_packed(const int* tile_ranges, int total_tiles, int x_blocks,
                                 int* tile_desc, uint32_t* tile_order, uint32_t* tile_counts,
                                 uint32_t* tile_ids, void* sort_temp, size_t sort_temp_bytes,
                                 int zigzag_group_size, int morton_bucket_size,
                                 ps_stream_t stream);
/* Plain renderer over a tile order: workgroup p renders tile tile_order[p]
 * with its range from tile_ranges.  Replaces render_16x16_reordered
 * (render.cu:7101-7125); bound at pybind.cpp:484-504. */
int ps_render_16x16_reordered(int num_rendered, const int* d_num_rendered, int width, int height,
                              const float* points_xy, const float* rgb_depth,
                              const float* conic_opacity, const uint64_t* keys_sorted,
                              const uint32_t* values_sorted, const int* tile_ranges,
                              const uint32_t* tile_order, const float bg[3], uint8_t* out_color,
                              ps_stream_t stream);
/* Plain renderer over packed descriptors: workgroup p renders the tile and
 * range of tile_desc[p] (as ps_compute_tile_order_packed writes them).
 * Replaces render_16x16_reordered_v2 (render.cu:7127-7150); bound at
 * pybind.cpp:556-575. */
int ps_render_16x16_reordered_desc(int num_rendered, const int* d_num_rendered, int width,
                                   int height, const float* points_xy, const float* rgb_depth,
                                   const float* conic_opacity, const uint32_t* values_sorted,
                                   const int* tile_desc, const float bg[3], uint8_t* out_color,
                                   ps_stream_t stream);
/* Top-k cache renderers: values_sorted may hold, instead of a Gaussian id,
 * 0x80000000 | slot, naming slot `slot` (< num_topk) of compact FP32
 * feature arrays topk_xy f32[num_topk][2], topk_rgb_depth and topk_conic
 * f32[num_topk][4] (the most-referenced Gaussians' features, copied by the
 * caller); ranges by an in-kernel search of keys_sorted; saturation test
 * every 2 Gaussians.  persistent == 0: one workgroup per tile, replaces
 * render_16x16_topk_smem (render.cu:6698-6723); persistent != 0: the tile
 * queue (`queue` as in ps_render_16x16_persistent), replaces
 * render_16x16_topk_smem_persistent[_v2] (render.cu:6725-6800); bound at
 * pybind.cpp:366-437.  The reference stages the cache in shared memory; here
 * it is read through L2, where 40 bytes per slot stay resident (a per-
 * workgroup LDS copy would cut the renderer's waves per SIMD several-fold). */
int ps_render_16x16_topk(int num_rendered, const int* d_num_rendered, int width, int height,
                         const float* points_xy, const float* rgb_depth,
                         const float* conic_opacity, const uint64_t* keys_sorted,
                         const uint32_t* values_sorted, const float* topk_xy,
                         const float* topk_rgb_depth, const float* topk_conic, int num_topk,
                         int persistent, const float bg[3], uint8_t* out_color, unsigned* queue,
                         ps_stream_t stream);
/* Persistent renderer over a tile order: the tile queue hands out order
 * positions from the head, position p to the XCD queue p % 8 (environment
 * POLYSPLAT_ORDER_QUEUES=contiguous: each XCD a contiguous eighth), each tile
 * from tile_desc (packed, as ps_compute_tile_order_packed writes it) or,
 * when tile_desc is NULL, tile_order and tile_ranges.  Other arguments as
 * ps_render_16x16_persistent.  Replaces render_16x16_reordered_persistent
 * (render.cu:7177-7215); bound at pybind.cpp:577-596. */
int ps_render_16x16_reordered_persistent(int num_rendered, const int* d_num_rendered, int width,
                                         int height, int feat_prec, const float* points_xy,
                                         const float* rgb_depth, const float* conic_opacity,
                                         const void* records, const uint32_t* values_sorted,
                                         const int* tile_ranges, const uint32_t* tile_order,
                                         const int* tile_desc, int dt, const float bg[3],
                                         uint8_t* out_color, unsigned* queue, ps_stream_t stream);

/* ---------------------------------------------------------------- forward
 * One frame on one stream: preprocess -> sort -> ranges -> render, with the
 * pair count passed between stages in device memory.  Replaces forward_fused
 * (pybind.cpp:932-1012), forward_fused_e2e_preranges (pybind.cpp:1016-1080)
 * and forward_fused_e2e (pybind.cpp:1085-1149).  feat_prec F32 writes the
 * f32 feature arrays; F16/BF16 writes 24-byte records into `records`
 * instead.  sync_count 1: the call waits for the count and returns
 * PS_ERR_OVERFLOW when the pairs did not fit; *num_rendered_out receives the
 * count in that case.  sync_count 2 ("post"): the call waits for nothing;
 * the count and overflow status are copied into a pinned slot the library
 * owns, and *num_rendered_out receives a ticket for ps_forward_wait, so a
 * caller can enqueue the next frame before reading this one's count (the
 * reference reads the count mid-frame, pybind.cpp:1122-1125).  Workspace:
 * ps_forward_workspace_size(). */
typedef struct ps_frame_buffers {
  int* curr_offset;               /* i32[1] */
  void* sort_space;               /* ps_get_sort_buffer_size(capacity) bytes */
  size_t sort_space_bytes;
  uint64_t* keys_unsorted;        /* [capacity] */
  uint32_t* values_unsorted;
  uint64_t* keys_sorted;
  uint32_t* values_sorted;
  int capacity;
  float* points_xy;               /* f32 features [P] (F32 mode) */
  float* rgb_depth;
  float* conic_opacity;
  void* records;                  /* 24-byte records [P] (F16/BF16 mode) */
  int* tile_ranges;               /* i32[tiles][2], may be NULL for the fused variant */
  void* workspace;                /* ps_forward_workspace_size(P) bytes */
  size_t workspace_bytes;
} ps_frame_buffers;

size_t ps_forward_workspace_size(int P);
/* Sort-space bytes a frame of `tiles` tiles needs under PS_FORWARD_TILE_ORDER. */
size_t ps_forward_tile_order_space(int tiles);
int ps_forward(int P, const float* positions, const void* shs, int sh_prec,
               const float* opacities, const float* cov3d, const ps_camera* cam,
               int variant, int feat_prec, const ps_frame_buffers* fb, const float bg[3],
               uint8_t* out_color, int sync_count, int* num_rendered_out, ps_stream_t stream);

/* ------------------------------------------------------------ frame graphs
 * One frame's whole pipeline (what ps_forward enqueues for this scene,
 * variant, feature precision, buffers and output) captured once into a HIP
 * graph and replayed per request with a new camera pose and background.
 * The graph is tied to cam->width x cam->height and to every buffer of fb;
 * each replay takes its own camera pose, background and output image
 * (u8[height][width][3]).  The kernels read those from 512 bytes of the
 * forward workspace's queue area, which ps_frame_graph_launch rewrites on
 * the stream before each replay (with a one-wave kernel, so the call
 * allocates nothing and copies nothing from the host).  Replays of one
 * graph, and ps_forward calls with the same buffers, must not overlap:
 * issue them on one stream.  sync_count and num_rendered_out as ps_forward.
 * No counterpart in the reference (its per-frame host sync, pybind.cpp:
 * 1122-1125, rules out capture); for serving loops that replay per
 * (scene, resolution). */
typedef struct ps_frame_graph ps_frame_graph;
int ps_frame_graph_create(int P, const float* positions, const void* shs, int sh_prec,
                          const float* opacities, const float* cov3d, const ps_camera* cam,
                          int variant, int feat_prec, const ps_frame_buffers* fb,
                          ps_frame_graph** out);
int ps_frame_graph_launch(ps_frame_graph* graph, const ps_camera* cam, const float bg[3],
                          uint8_t* out_color, int sync_count, int* num_rendered_out,
                          ps_stream_t stream);
/* Waits for the device to finish its queued work, then frees the graph. */
int ps_frame_graph_destroy(ps_frame_graph* graph);

/* Count of a frame posted with ps_forward(sync_count == 2): waits for the
 * frame's copy, writes its pair count to *num_rendered_out and returns
 * PS_ERR_OVERFLOW when its pairs did not fit (nothing was rendered then).
 * Each ticket is waited for once, on the host thread that posted it; a
 * thread may have up to 64 posted frames not yet waited for.  No
 * counterpart in the reference (fetch_num_rendered_async, sort.cu:289-300,
 * always waits at once). */
int ps_forward_wait(int ticket, int* num_rendered_out);
/* 1 when the posted frame's count has arrived (ps_forward_wait will not
 * block), 0 when not yet, a negative status on error. */
int ps_forward_query(int ticket);

/* ---------------------------------------------------------------- batch
 * Several frames, each with its own scene and camera, rendered through
 * shared launches: one preprocess (count, scan, emit) over all frames' Gaussians,
 * one sort of all pairs (tile ids are offset per frame so that frames occupy
 * consecutive tile ranges), one range build and one persistent render whose
 * tile queue spans every frame.  Per frame the result equals ps_forward with
 * PS_VARIANT_PERSISTENT_LITE: keys, order and images are identical.
 * The reference renders one frame per call (forward_fused_e2e_preranges,
 * pybind.cpp:1016-1080) and has no batched form; this entry point serves
 * callers that render concurrent camera requests (a serving loop).
 *
 * Shared buffers come from fb: keys/values/sort_space sized for the sum of
 * the frames' pairs (capacity), tile_ranges for the sum of their tiles,
 * curr_offset receives the total pair count, workspace holds
 * ps_forward_batch_workspace_size(num_frames, P) bytes; fb's feature
 * pointers are ignored.  Per-frame features and images come from
 * ps_frame.  All frames share sh_prec and feat_prec.  The frame descriptors
 * are staged through pinned memory owned by the library, so the call only
 * enqueues work (it blocks only when 16 earlier batches of the same host
 * thread are still waiting for their descriptor copy).  With sync_count != 0
 * the call waits, writes each frame's pair count to frame_counts (host, may
 * be NULL) and returns PS_ERR_OVERFLOW when the pairs did not fit (then no
 * keys were written and every image shows its background).
 *
 * Status: currently switched off -- after validating its arguments the call
 * returns PS_ERR_UNSUPPORTED without touching the device. */
#define PS_MAX_BATCH_FRAMES 64

typedef struct ps_frame {
  int P;
  const float* positions;         /* f32[P][3] */
  const void* shs;                /* [P][48] in sh_prec */
  const float* opacities;         /* f32[P] */
  const float* cov3d;             /* f32[P][6] */
  ps_camera cam;                  /* 16x16 tiles */
  float bg[3];
  float* points_xy;               /* f32 features [P] (F32 mode) */
  float* rgb_depth;
  float* conic_opacity;
  void* records;                  /* 24-byte records [P] (F16/BF16 mode) */
  uint8_t* out_color;             /* u8[height][width][3] */
} ps_frame;

size_t ps_forward_batch_workspace_size(int num_frames, const int* P);
int ps_forward_batch(int num_frames, const ps_frame* frames, int sh_prec, int feat_prec,
                     const ps_frame_buffers* fb, int sync_count, int* frame_counts,
                     ps_stream_t stream);
/* Device pointer to the per-frame pair counts a batch leaves in its
 * workspace (int[num_frames], valid once the batch has run), for callers
 * that read them back asynchronously. */
const int* ps_forward_batch_counts_ptr(const void* workspace, int num_frames, const int* P);

#ifdef __cplusplus
}
#endif

#endif /* POLYSPLAT_AMD_H */
