// One-frame pipeline and the small host-side C ABI entries.
//
// ps_forward chains preprocess -> sort -> (ranges) -> render on one stream.
// Unlike the reference's fused forwards (pybind.cpp:932-1149), which copy the
// pair count to the host between preprocess and sort so CUB can size its
// launches, every stage here reads the count from device memory, so the whole
// frame is enqueued without a host round trip (and can be captured in a HIP
// graph).  The count is read back only at the end, and only when asked.

#include <stdarg.h>

#include "ps_common.h"
#include "ps_host.h"

namespace ps {

int preprocess_launch(int P, const float* positions, const void* shs, int sh_prec,
                      const float* opacities, const float* cov3d, const ps_camera* cam,
                      float* points_xy, float* rgb_depth, float* conic_opacity, void* feat_records,
                      int feat_prec, uint64_t* keys, uint32_t* values, int capacity,
                      int* curr_offset, void* workspace, size_t workspace_bytes,
                      hipStream_t stream, bool reset_offset, const CamParams* cam_dev = nullptr);
int sort_end_bit(int width, int height, int block_x, int block_y);
int preprocess_es_launch(int P, const float* positions, const void* shs, int sh_prec,
                         const float* opacities, const float* cov3d, const ps_camera* cam,
                         float* points_xy, float* rgb_depth, float* conic_opacity,
                         void* feat_records, int feat_prec, uint32_t* tiles, uint32_t* vals,
                         int capacity, int* curr_offset, void* workspace, size_t workspace_bytes,
                         void* sort_space, size_t sort_space_bytes, int sort_max_pairs,
                         hipStream_t stream, const CamParams* cam_dev = nullptr);

FrameDev make_frame_dev(int P, const float* positions, const void* shs, const float* opacities,
                        const float* cov3d, const ps_camera* cam, float* xy, float* rgbd,
                        float* conic, void* rec, int rec_prec, uint8_t* out, const float* bg);
int check_frame(const FrameDev& F, const ps_camera* cam);
size_t preprocess_batch_workspace_size(int nf, const int* P);
int preprocess_batch_launch(int nf, FrameDev* frames_host, int sh_prec, uint64_t* keys,
                            uint32_t* values, int capacity, int* total_out, void* workspace,
                            size_t workspace_bytes, int** frame_counts_dev, const int** state_count,
                            int* tiles_out, hipStream_t stream);
int render_batch_launch(const FrameDev* frames_dev, int nf, int feat_prec, const int* ranges,
                        int tiles, const uint32_t* order, unsigned* queue, hipStream_t stream);
const FrameDev* batch_frames_dev(int nf, const FrameDev* frames_host, void* workspace);
int* batch_counts_dev(int nf, const int* P, void* workspace);

namespace {

thread_local char g_last_error[512] = "";

constexpr int kPinnedInts = 256;

// Pinned slot for count readbacks (the reference keeps the same kind of slot
// per thread, sort.cu:39-69).
struct PinnedSlot {
  int* host = nullptr;
  hipEvent_t ev = nullptr;
  bool ok = false;
  PinnedSlot() {
    if (hipHostMalloc(reinterpret_cast<void**>(&host), kPinnedInts * sizeof(int),
                      hipHostMallocDefault) !=
        hipSuccess) {
      host = nullptr;
      return;
    }
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      ev = nullptr;
      return;
    }
    ok = true;
  }
  ~PinnedSlot() {
    if (ev) (void)hipEventDestroy(ev);
    if (host) (void)hipHostFree(host);
  }
};

PinnedSlot& pinned() {
  static thread_local PinnedSlot s;
  return s;
}

constexpr size_t kQueueBytes = 2048;

}  // namespace

void set_last_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

// Copies spans of device ints (ptrs[i], lens[i]) into `out`, back to back,
// through the pinned slot, and waits.
static int read_back_spans(const int* const* ptrs, const int* lens, int n, int* out,
                           hipStream_t stream) {
  PinnedSlot& s = pinned();
  int total = 0;
  for (int i = 0; i < n; ++i) total += lens[i];
  if (!s.ok || total > kPinnedInts) {
    int o = 0;
    for (int i = 0; i < n; ++i) {
      PS_HIP_CHECK(hipMemcpyAsync(out + o, ptrs[i], lens[i] * sizeof(int), hipMemcpyDeviceToHost,
                                  stream));
      o += lens[i];
    }
    PS_HIP_CHECK(hipStreamSynchronize(stream));
    return PS_OK;
  }
  int o = 0;
  for (int i = 0; i < n; ++i) {
    PS_HIP_CHECK(hipMemcpyAsync(s.host + o, ptrs[i], lens[i] * sizeof(int), hipMemcpyDeviceToHost,
                                stream));
    o += lens[i];
  }
  PS_HIP_CHECK(hipEventRecord(s.ev, stream));
  PS_HIP_CHECK(hipEventSynchronize(s.ev));
  for (int i = 0; i < total; ++i) out[i] = s.host[i];
  return PS_OK;
}

// Copies n single ints from device memory and waits.
static int read_back(const int* const* ptrs, int n, int* out, hipStream_t stream) {
  int lens[4] = {1, 1, 1, 1};
  if (n > 4) return PS_ERR_INVALID;
  return read_back_spans(ptrs, lens, n, out, stream);
}

}  // namespace ps

extern "C" const char* ps_status_string(int status) {
  switch (status) {
    case PS_OK: return "ok";
    case PS_ERR_INVALID: return "invalid argument";
    case PS_ERR_HIP: return "HIP runtime error";
    case PS_ERR_OVERFLOW: return "too many (tile, Gaussian) pairs for the key buffers";
    case PS_ERR_UNSUPPORTED: return "unsupported configuration";
    default: return "unknown status";
  }
}

extern "C" const char* ps_last_error(void) { return ps::g_last_error; }

extern "C" const char* ps_version(void) { return "polysplat_amd 0.1 (gfx950)"; }

extern "C" int ps_camera_matrices(const ps_camera* cam, float view[16], float proj[16],
                                  float tan_fov[2]) {
  PS_REQUIRE(cam && view && proj && tan_fov, "NULL argument");
  ps::camera_matrices(cam, view, proj, &tan_fov[0], &tan_fov[1]);
  return PS_OK;
}

extern "C" int ps_fetch_num_rendered(const int* d_curr_offset, int* out_num_rendered,
                                     ps_stream_t stream) {
  PS_REQUIRE(d_curr_offset && out_num_rendered, "NULL argument");
  const int* ptrs[1] = {d_curr_offset};
  return ps::read_back(ptrs, 1, out_num_rendered, reinterpret_cast<hipStream_t>(stream));
}

namespace ps {
namespace {

// Ring of pinned slots for posted count readbacks (ps_forward with
// sync_count == 2), one ring per host thread.  Each posted frame's pair
// count and status word are copied into a slot and an event is recorded
// after the copy; ps_forward_wait waits for that event.  A slot is reused
// only after the frame that last used it has been waited for.
struct PostRing {
  static constexpr int kSlots = 64;
  int* host = nullptr;  // 2 ints per slot
  hipEvent_t ev[kSlots] = {};
  bool busy[kSlots] = {};
  unsigned next = 0;
  int capacity = 0;
  bool ok = false;
  PostRing() {
    if (hipHostMalloc(reinterpret_cast<void**>(&host), 2 * kSlots * sizeof(int),
                      hipHostMallocDefault) != hipSuccess) {
      host = nullptr;
      return;
    }
    for (int i = 0; i < kSlots; ++i)
      if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) return;
    ok = true;
  }
  ~PostRing() {
    for (int i = 0; i < kSlots; ++i)
      if (ev[i]) (void)hipEventDestroy(ev[i]);
    if (host) (void)hipHostFree(host);
  }
};

PostRing& post_ring() {
  static thread_local PostRing r;
  return r;
}

// Capacity each posted slot's frame had, for the overflow message.
thread_local int g_post_capacity[PostRing::kSlots];

// A free slot of the calling thread's ring (checked before a posted frame
// launches anything, so a refused frame does no work).
int free_post_slot(int* slot) {
  PostRing& r = post_ring();
  PS_REQUIRE(r.ok, "no pinned memory for posted count readbacks");
  for (int k = 0; k < PostRing::kSlots; ++k) {
    const int i = (int)((r.next + (unsigned)k) % PostRing::kSlots);
    if (!r.busy[i]) {
      *slot = i;
      return PS_OK;
    }
  }
  set_last_error("more than %d posted frames have not been waited for", PostRing::kSlots);
  return PS_ERR_INVALID;
}

int post_count(int slot, const int* d_count, const int* d_status, int capacity,
               hipStream_t stream) {
  PostRing& r = post_ring();
  PS_HIP_CHECK(hipMemcpyAsync(r.host + 2 * slot, d_count, sizeof(int), hipMemcpyDeviceToHost,
                              stream));
  PS_HIP_CHECK(hipMemcpyAsync(r.host + 2 * slot + 1, d_status, sizeof(int),
                              hipMemcpyDeviceToHost, stream));
  PS_HIP_CHECK(hipEventRecord(r.ev[slot], stream));
  r.busy[slot] = true;
  r.next = (unsigned)(slot + 1);
  g_post_capacity[slot] = capacity;
  return PS_OK;
}

}  // namespace
}  // namespace ps

extern "C" int ps_forward_wait(int ticket, int* num_rendered_out) {
  using namespace ps;
  PostRing& r = post_ring();
  PS_REQUIRE(ticket >= 0 && ticket < PostRing::kSlots && r.busy[ticket],
             "ticket %d is not a posted frame of this thread", ticket);
  r.busy[ticket] = false;
  PS_HIP_CHECK(hipEventSynchronize(r.ev[ticket]));
  const int n = r.host[2 * ticket], status = r.host[2 * ticket + 1];
  if (num_rendered_out) *num_rendered_out = n;
  if (status & 1) {
    set_last_error("%d pairs do not fit in %d", n, g_post_capacity[ticket]);
    return PS_ERR_OVERFLOW;
  }
  return PS_OK;
}

extern "C" int ps_forward_query(int ticket) {
  using namespace ps;
  PostRing& r = post_ring();
  if (!(ticket >= 0 && ticket < PostRing::kSlots && r.busy[ticket])) {
    set_last_error("ticket %d is not a posted frame of this thread", ticket);
    return -PS_ERR_INVALID;
  }
  const hipError_t e = hipEventQuery(r.ev[ticket]);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();  // not an error: clear it
    return 0;
  }
  set_last_error("hipEventQuery failed: %s", hipGetErrorString(e));
  return -PS_ERR_HIP;
}

extern "C" int ps_preprocess_depth_first(int P, const float* positions, const void* shs,
                                         int sh_prec, const float* opacities, const float* cov3d,
                                         const ps_camera* cam, float* points_xy, float* rgb_depth,
                                         float* conic_opacity, void* feat_records, int feat_prec,
                                         uint32_t* tile_ids, uint32_t* values, int capacity,
                                         int* curr_offset, void* workspace,
                                         size_t workspace_bytes, void* sort_space,
                                         size_t sort_space_bytes, ps_stream_t stream) {
  PS_REQUIRE(cam != nullptr, "NULL camera");
  return ps::preprocess_es_launch(P, positions, shs, sh_prec, opacities, cov3d, cam, points_xy,
                                  rgb_depth, conic_opacity, feat_records, feat_prec, tile_ids,
                                  values, capacity, curr_offset, workspace, workspace_bytes,
                                  sort_space, sort_space_bytes, capacity,
                                  reinterpret_cast<hipStream_t>(stream));
}

extern "C" int ps_es_tile_sort(int num_rendered, const int* d_num_rendered, int max_pairs,
                               int width, int height, int block_x, int block_y, void* sort_space,
                               size_t sort_space_bytes, const uint32_t* tile_ids_unsorted,
                               const uint32_t* values_unsorted, uint32_t* tile_ids_sorted,
                               uint32_t* values_sorted, uint64_t* keys_sorted, int* tile_ranges,
                               ps_stream_t stream) {
  PS_REQUIRE(width > 0 && height > 0 && block_x > 0 && block_y > 0, "bad image/tile size");
  PS_REQUIRE(tile_ids_unsorted && values_unsorted && tile_ids_sorted && values_sorted,
             "NULL pair buffer");
  const int tiles = ((width + block_x - 1) / block_x) * ((height + block_y - 1) / block_y);
  return ps::tile_sort_launch(num_rendered, d_num_rendered, max_pairs, tiles, sort_space,
                              sort_space_bytes, tile_ids_unsorted, values_unsorted,
                              tile_ids_sorted, values_sorted, keys_sorted, tile_ranges,
                              reinterpret_cast<hipStream_t>(stream));
}

extern "C" size_t ps_forward_tile_order_space(int tiles) {
  return ps::align_up(ps::tile_order_temp_size(tiles < 0 ? 0 : tiles), 256) +
         16 * (size_t)(tiles < 0 ? 0 : tiles);
}

extern "C" size_t ps_forward_workspace_size(int P) {
  return ps::kQueueBytes + ps::align_up(ps_preprocess_workspace_size(P), 256);
}

namespace ps {
namespace {

// Checks and enqueues one frame (no count readback).  fc: NULL, or the
// device copy of the camera, background and image pointer that the kernels
// read instead of their arguments (a captured graph's frames).
int forward_enqueue(int P, const float* positions, const void* shs, int sh_prec,
                    const float* opacities, const float* cov3d, const ps_camera* cam, int variant,
                    int feat_prec, const ps_frame_buffers* fb, const float bg[3],
                    uint8_t* out_color, hipStream_t stream, const FrameConsts* fc) {
  const CamParams* cam_dev = fc ? &fc->cp : nullptr;
  PS_REQUIRE(fb != nullptr && cam != nullptr && bg != nullptr, "NULL argument");
  const bool depth_first = (variant & PS_FORWARD_DEPTH_FIRST) != 0;
  const bool tile_ordered = (variant & PS_FORWARD_TILE_ORDER) != 0;
  variant &= ~(PS_FORWARD_DEPTH_FIRST | PS_FORWARD_TILE_ORDER);
  PS_REQUIRE(variant >= PS_VARIANT_DEFAULT && variant <= PS_VARIANT_PERSISTENT_DT,
             "unknown render variant %d", variant);
  PS_REQUIRE(cam->block_x == 16 && cam->block_y == 16,
             "the renderers implement 16x16 tiles only (got %dx%d)", cam->block_x, cam->block_y);
  PS_REQUIRE(feat_prec == PS_PREC_F32 || variant != PS_VARIANT_DEFAULT,
             "the plain renderer reads FP32 features");
  PS_REQUIRE(fb->workspace != nullptr && fb->workspace_bytes >= ps_forward_workspace_size(P),
             "forward workspace too small");
  PS_REQUIRE(fb->curr_offset != nullptr, "curr_offset is NULL");
  const bool f32 = feat_prec == PS_PREC_F32;
  PS_REQUIRE(f32 ? (fb->points_xy && fb->rgb_depth && fb->conic_opacity) : fb->records != nullptr,
             "feature buffers for the chosen precision are NULL");
  const bool need_ranges =
      variant == PS_VARIANT_PERSISTENT_LITE || variant == PS_VARIANT_PERSISTENT_DT;
  PS_REQUIRE(!need_ranges || fb->tile_ranges != nullptr, "tile_ranges is NULL");
  const int gx = (cam->width + 15) / 16;
  PS_REQUIRE(!tile_ordered || need_ranges,
             "PS_FORWARD_TILE_ORDER needs a variant with tile ranges (PERSISTENT_LITE or _DT)");
  PS_REQUIRE(!tile_ordered || fb->sort_space_bytes >=
                                  ps_forward_tile_order_space(gx * ((cam->height + 15) / 16)),
             "sort space too small for the tile order (%zu < %zu)", fb->sort_space_bytes,
             ps_forward_tile_order_space(gx * ((cam->height + 15) / 16)));

  // The queue words sit at the start of the workspace, where no other stage
  // writes, so a workspace reused for scenes of other sizes keeps them zero.
  unsigned* queue = static_cast<unsigned*>(fb->workspace);
  char* ws = static_cast<char*>(fb->workspace) + kQueueBytes;
  const size_t pre_bytes = align_up(ps_preprocess_workspace_size(P), 256);
  int st;
  const int tiles = ((cam->width + 15) / 16) * ((cam->height + 15) / 16);
  const int* d_n = ps_preprocess_count_ptr(ws);
  if (depth_first) {
    // Depth-first pipeline: Gaussians sorted by depth, their (tile id,
    // Gaussian) pairs emitted in that order as 32-bit tile ids (first half
    // of keys_unsorted's memory), stably sorted by tile id (the sorted ids
    // in its second half); the 64-bit sorted keys (tile id << 32) and the
    // tile ranges from that sort.
    uint32_t* tk = reinterpret_cast<uint32_t*>(fb->keys_unsorted);
    uint32_t* tk_sorted = tk + fb->capacity;
    st = preprocess_es_launch(P, positions, shs, sh_prec, opacities, cov3d, cam,
                              f32 ? fb->points_xy : nullptr, f32 ? fb->rgb_depth : nullptr,
                              f32 ? fb->conic_opacity : nullptr, f32 ? nullptr : fb->records,
                              feat_prec, tk, fb->values_unsorted, fb->capacity, fb->curr_offset,
                              ws, pre_bytes, fb->sort_space, fb->sort_space_bytes, fb->capacity,
                              stream, cam_dev);
    if (st != PS_OK) return st;
    st = tile_sort_launch(0, d_n, fb->capacity, tiles, fb->sort_space, fb->sort_space_bytes, tk,
                          fb->values_unsorted, tk_sorted, fb->values_sorted, fb->keys_sorted,
                          need_ranges ? fb->tile_ranges : nullptr, stream);
    if (st != PS_OK) return st;
  } else {
    // The frame's keys start at offset 0 (the reference clears curr_offset
    // first, pybind.cpp:946); the scan kernel takes 0 as the base directly.
    st = preprocess_launch(P, positions, shs, sh_prec, opacities, cov3d, cam,
                           f32 ? fb->points_xy : nullptr, f32 ? fb->rgb_depth : nullptr,
                           f32 ? fb->conic_opacity : nullptr, f32 ? nullptr : fb->records,
                           feat_prec, fb->keys_unsorted, fb->values_unsorted, fb->capacity,
                           fb->curr_offset, ws, pre_bytes, stream, true, cam_dev);
    if (st != PS_OK) return st;
    // The sort leaves the tile ranges in tile_ranges when the renderer needs
    // them (its segment pass computes them anyway).
    st = sort_launch(0, d_n, fb->capacity, sort_end_bit(cam->width, cam->height, 16, 16),
                     fb->sort_space, fb->sort_space_bytes, fb->keys_unsorted, fb->values_unsorted,
                     fb->keys_sorted, fb->values_sorted, stream,
                     need_ranges ? fb->tile_ranges : nullptr, tiles);
    if (st != PS_OK) return st;
  }
  if (variant == PS_VARIANT_DEFAULT) {
    return render_plain_launch(0, d_n, cam->width, cam->height, fb->points_xy, fb->rgb_depth,
                               fb->conic_opacity, fb->keys_sorted, fb->values_sorted, bg,
                               out_color, stream, nullptr, nullptr, nullptr, fc);
  }
  int* desc = nullptr;
  if (tile_ordered) {
    // The sort is done with its space (stream order): the tile order's
    // scratch at its start, the packed descriptors after it.
    const size_t tb = tile_order_temp_size(tiles);
    desc = reinterpret_cast<int*>(static_cast<char*>(fb->sort_space) + align_up(tb, 256));
    st = tile_order_launch(fb->tile_ranges, tiles, gx, 0, 0, nullptr, nullptr, nullptr, desc,
                           fb->sort_space, tb, stream);
    if (st != PS_OK) return st;
  }
  return render_persistent_launch(0, d_n, cam->width, cam->height, feat_prec, fb->points_xy,
                                  fb->rgb_depth, fb->conic_opacity, fb->records, 0,
                                  fb->keys_sorted, fb->values_sorted,
                                  need_ranges ? fb->tile_ranges : nullptr,
                                  variant == PS_VARIANT_PERSISTENT_DT, bg, out_color, queue,
                                  stream, nullptr, desc, fc);
}

// The end-of-frame count: none (sync_count 0), waited for (1), or posted to
// a pinned slot whose ticket goes to *num_rendered_out (2; `slot` reserved).
int finish_count(const ps_frame_buffers* fb, int sync_count, int slot, int* num_rendered_out,
                 hipStream_t stream) {
  const int* status = ps_preprocess_status_ptr(static_cast<char*>(fb->workspace) + kQueueBytes);
  if (sync_count == 2) {
    const int st = post_count(slot, fb->curr_offset, status, fb->capacity, stream);
    if (st != PS_OK) return st;
    if (num_rendered_out) *num_rendered_out = slot;
    return PS_OK;
  }
  if (sync_count) {
    const int* ptrs[2] = {fb->curr_offset, status};
    int vals[2] = {0, 0};
    const int st = read_back(ptrs, 2, vals, stream);
    if (st != PS_OK) return st;
    if (num_rendered_out) *num_rendered_out = vals[0];
    if (vals[1] & 1) {
      set_last_error("%d pairs do not fit in %d", vals[0], fb->capacity);
      return PS_ERR_OVERFLOW;
    }
  }
  return PS_OK;
}

}  // namespace
}  // namespace ps

extern "C" int ps_forward(int P, const float* positions, const void* shs, int sh_prec,
                          const float* opacities, const float* cov3d, const ps_camera* cam,
                          int variant, int feat_prec, const ps_frame_buffers* fb,
                          const float bg[3], uint8_t* out_color, int sync_count,
                          int* num_rendered_out, ps_stream_t stream_) {
  using namespace ps;
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  PS_REQUIRE(fb != nullptr, "NULL argument");
  PS_REQUIRE(sync_count >= 0 && sync_count <= 2, "sync_count must be 0, 1 or 2");
  int slot = -1;
  if (sync_count == 2) {
    const int fs = free_post_slot(&slot);
    if (fs != PS_OK) return fs;
  }
  const int st = forward_enqueue(P, positions, shs, sh_prec, opacities, cov3d, cam, variant,
                                 feat_prec, fb, bg, out_color, stream, nullptr);
  if (st != PS_OK) return st;
  return finish_count(fb, sync_count, slot, num_rendered_out, stream);
}

// ------------------------------------------------------------ frame graphs
// A frame's whole pipeline (every kernel ps_forward enqueues) captured once
// into a HIP graph, for one scene, resolution, variant and set of buffers,
// and replayed per request.  The kernels take their camera, background and
// output image from a block of per-frame constants in the forward
// workspace's queue area (bytes 1536..2047, which the tile queue does not
// use) instead of their arguments, so a replay needs only that block
// rewritten first: a one-wave kernel whose arguments carry the new values
// does it, on the same stream.  Host work per frame: two launches instead of
// about twenty.
struct ps_frame_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int width = 0, height = 0;
  ps_frame_buffers fb;
  ps::FrameConsts* consts = nullptr;
};

namespace ps {
namespace {

constexpr size_t kFrameConstsOffset = 1536;
static_assert(kFrameConstsOffset + sizeof(FrameConsts) <= kQueueBytes,
              "frame constants fit the queue area");
static_assert(kFrameConstsOffset % alignof(FrameConsts) == 0, "aligned frame constants");

// The next replay's camera, background and image pointer into device memory.
__global__ __launch_bounds__(64) void ps_set_frame_consts(FrameConsts* __restrict__ dst,
                                                          FrameConsts v) {
  constexpr int kWords = (int)(sizeof(FrameConsts) / sizeof(uint32_t));
  static_assert(kWords <= 64 && sizeof(FrameConsts) % sizeof(uint32_t) == 0,
                "one word per lane");
  const int l = threadIdx.x;
  if (l < kWords) reinterpret_cast<uint32_t*>(dst)[l] = reinterpret_cast<const uint32_t*>(&v)[l];
}

}  // namespace
}  // namespace ps

extern "C" int ps_frame_graph_create(int P, const float* positions, const void* shs, int sh_prec,
                                     const float* opacities, const float* cov3d,
                                     const ps_camera* cam, int variant, int feat_prec,
                                     const ps_frame_buffers* fb, ps_frame_graph** out) {
  using namespace ps;
  PS_REQUIRE(out != nullptr && fb != nullptr && cam != nullptr, "NULL argument");
  *out = nullptr;
  PS_REQUIRE(fb->workspace != nullptr && fb->workspace_bytes >= ps_forward_workspace_size(P),
             "forward workspace too small");
  ps_frame_graph* g = new ps_frame_graph();
  g->width = cam->width;
  g->height = cam->height;
  g->fb = *fb;
  g->consts = reinterpret_cast<FrameConsts*>(static_cast<char*>(fb->workspace) +
                                             kFrameConstsOffset);
  // Capture on a stream of the library's own (a capture needs a created
  // stream; the caller's may be the null stream).  Default priority.
  hipStream_t cs = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete g;
    set_last_error("hipStreamCreateWithFlags failed: %s", hipGetErrorString(e));
    return PS_ERR_HIP;
  }
  int st = PS_OK;
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    set_last_error("hipStreamBeginCapture failed: %s", hipGetErrorString(e));
    st = PS_ERR_HIP;
  } else {
    // The captured arguments' background and image are placeholders: the
    // kernels read the constants block.
    const float bg0[3] = {0.f, 0.f, 0.f};
    st = forward_enqueue(P, positions, shs, sh_prec, opacities, cov3d, cam, variant, feat_prec,
                         fb, bg0, nullptr, cs, g->consts);
    char saved[512];
    snprintf(saved, sizeof(saved), "%s", ps_last_error());
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(cs, &graph);
    g->graph = graph;
    if (st != PS_OK) {
      set_last_error("%s", saved);
    } else if (e2 != hipSuccess || graph == nullptr) {
      set_last_error("hipStreamEndCapture failed: %s", hipGetErrorString(e2));
      st = PS_ERR_HIP;
    } else {
      const hipError_t e3 = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
      if (e3 != hipSuccess) {
        set_last_error("hipGraphInstantiate failed: %s", hipGetErrorString(e3));
        st = PS_ERR_HIP;
      }
    }
  }
  (void)hipStreamDestroy(cs);
  if (st != PS_OK) {
    ps_frame_graph_destroy(g);
    return st;
  }
  *out = g;
  return PS_OK;
}

extern "C" int ps_frame_graph_launch(ps_frame_graph* g, const ps_camera* cam, const float bg[3],
                                     uint8_t* out_color, int sync_count, int* num_rendered_out,
                                     ps_stream_t stream_) {
  using namespace ps;
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  PS_REQUIRE(g != nullptr && g->exec != nullptr && cam != nullptr && bg != nullptr,
             "NULL argument");
  PS_REQUIRE(out_color != nullptr, "out_color is NULL");
  PS_REQUIRE(cam->width == g->width && cam->height == g->height && cam->block_x == 16 &&
                 cam->block_y == 16,
             "the graph renders %dx%d frames with 16x16 tiles (camera: %dx%d, %dx%d tiles)",
             g->width, g->height, cam->width, cam->height, cam->block_x, cam->block_y);
  PS_REQUIRE(sync_count >= 0 && sync_count <= 2, "sync_count must be 0, 1 or 2");
  int slot = -1;
  if (sync_count == 2) {
    const int fs = free_post_slot(&slot);
    if (fs != PS_OK) return fs;
  }
  FrameConsts v;
  memset(&v, 0, sizeof(v));
  v.cp = make_cam_params(cam);
  v.bg[0] = bg[0];
  v.bg[1] = bg[1];
  v.bg[2] = bg[2];
  v.out = out_color;
  hipLaunchKernelGGL(ps_set_frame_consts, dim3(1), dim3(64), 0, stream, g->consts, v);
  PS_LAUNCH_CHECK();
  PS_HIP_CHECK(hipGraphLaunch(g->exec, stream));
  return finish_count(&g->fb, sync_count, slot, num_rendered_out, stream);
}

extern "C" int ps_frame_graph_destroy(ps_frame_graph* g) {
  if (g == nullptr) return PS_OK;
  // Replays may still be queued on any stream: let the device finish them
  // before their executable graph goes away.
  if (g->exec) (void)hipDeviceSynchronize();
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return PS_OK;
}

namespace ps {
// The batched path is compiled in but refused at run time unless the build
// defines PS_ENABLE_BATCH (the host emulation of the kernels does, to test
// it on the CPU).
static bool batch_enabled() {
#ifdef PS_ENABLE_BATCH
  return true;
#else
  return false;
#endif
}
}  // namespace ps

extern "C" size_t ps_forward_batch_workspace_size(int num_frames, const int* P) {
  if (num_frames <= 0 || P == nullptr) return 0;
  return ps::align_up(ps::preprocess_batch_workspace_size(num_frames, P), 256) + ps::kQueueBytes;
}

extern "C" const int* ps_forward_batch_counts_ptr(const void* workspace, int num_frames,
                                                  const int* P) {
  if (workspace == nullptr || P == nullptr || num_frames <= 0) return nullptr;
  return ps::batch_counts_dev(num_frames, P,
                              static_cast<char*>(const_cast<void*>(workspace)) + ps::kQueueBytes);
}

extern "C" int ps_forward_batch(int num_frames, const ps_frame* frames, int sh_prec, int feat_prec,
                                const ps_frame_buffers* fb, int sync_count, int* frame_counts,
                                ps_stream_t stream_) {
  using namespace ps;
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  PS_REQUIRE(num_frames > 0 && num_frames <= PS_MAX_BATCH_FRAMES,
             "batch of %d frames (allowed 1..%d)", num_frames, PS_MAX_BATCH_FRAMES);
  PS_REQUIRE(frames != nullptr && fb != nullptr, "NULL argument");
  PS_REQUIRE(feat_prec == PS_PREC_F32 || feat_prec == PS_PREC_F16 || feat_prec == PS_PREC_BF16,
             "unknown feature precision %d", feat_prec);
  PS_REQUIRE(fb->curr_offset != nullptr && fb->tile_ranges != nullptr, "shared buffers are NULL");
  const bool f32 = feat_prec == PS_PREC_F32;
  FrameDev host[PS_MAX_BATCH_FRAMES];
  int P[PS_MAX_BATCH_FRAMES];
  for (int f = 0; f < num_frames; ++f) {
    const ps_frame& q = frames[f];
    PS_REQUIRE(q.cam.block_x == 16 && q.cam.block_y == 16,
               "frame %d: the renderers implement 16x16 tiles only", f);
    PS_REQUIRE(q.out_color != nullptr, "frame %d: out_color is NULL", f);
    PS_REQUIRE(f32 ? (q.points_xy && q.rgb_depth && q.conic_opacity) : q.records != nullptr,
               "frame %d: feature buffers for the chosen precision are NULL", f);
    host[f] = make_frame_dev(q.P, q.positions, q.shs, q.opacities, q.cov3d, &q.cam,
                             f32 ? q.points_xy : nullptr, f32 ? q.rgb_depth : nullptr,
                             f32 ? q.conic_opacity : nullptr, f32 ? nullptr : q.records,
                             feat_prec, q.out_color, q.bg);
    const int st = check_frame(host[f], &q.cam);
    if (st != PS_OK) return st;
    P[f] = q.P;
  }
  // Switched off: a GPU run of the batched path faulted the device and the
  // cause has not been found yet.  Arguments are still validated.
  if (!batch_enabled()) {
    set_last_error("batched frames are switched off in this build; render the frames with "
                   "ps_forward");
    return PS_ERR_UNSUPPORTED;
  }
  const size_t pre_bytes = align_up(preprocess_batch_workspace_size(num_frames, P), 256);
  PS_REQUIRE(fb->workspace != nullptr && fb->workspace_bytes >= pre_bytes + kQueueBytes,
             "batch workspace too small (%zu < %zu)", fb->workspace_bytes, pre_bytes + kQueueBytes);
  // Queue words first (fixed place for any batch composition), then the
  // preprocess area.
  unsigned* queue = static_cast<unsigned*>(fb->workspace);
  void* ws = static_cast<char*>(fb->workspace) + kQueueBytes;
  int* d_frame_counts = nullptr;
  const int* d_n = nullptr;
  int tiles = 0;
  int st = preprocess_batch_launch(num_frames, host, sh_prec, fb->keys_unsorted,
                                   fb->values_unsorted, fb->capacity, fb->curr_offset,
                                   ws, pre_bytes, &d_frame_counts, &d_n, &tiles, stream);
  if (st != PS_OK) return st;
  st = sort_launch(0, d_n, fb->capacity, 32 + bit_length((unsigned)tiles), fb->sort_space,
                   fb->sort_space_bytes, fb->keys_unsorted, fb->values_unsorted, fb->keys_sorted,
                   fb->values_sorted, stream, fb->tile_ranges, tiles);
  if (st != PS_OK) return st;
  st = render_batch_launch(batch_frames_dev(num_frames, host, ws), num_frames,
                           feat_prec, fb->tile_ranges, tiles, fb->values_sorted, queue, stream);
  if (st != PS_OK) return st;
  if (sync_count) {
    // Total, status word (it follows the count in the preprocess state) and
    // the per-frame counts.
    const int* ptrs[3] = {fb->curr_offset, d_n + 1, d_frame_counts};
    const int lens[3] = {1, 1, num_frames};
    int vals[PS_MAX_BATCH_FRAMES + 2];
    st = read_back_spans(ptrs, lens, 3, vals, stream);
    if (st != PS_OK) return st;
    if (frame_counts)
      for (int f = 0; f < num_frames; ++f) frame_counts[f] = vals[2 + f];
    if (vals[1] & 1) {
      set_last_error("%d pairs do not fit in %d", vals[0], fb->capacity);
      return PS_ERR_OVERFLOW;
    }
  }
  return PS_OK;
}
