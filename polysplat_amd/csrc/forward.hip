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
                      hipStream_t stream, bool reset_offset);
int sort_launch(int n_host, const int* d_n, int max_pairs, int end_bit, void* space,
                size_t space_bytes, const uint64_t* keys_in, const uint32_t* vals_in,
                uint64_t* keys_out, uint32_t* vals_out, hipStream_t stream);
int sort_end_bit(int width, int height, int block_x, int block_y);
int tile_ranges_scan_launch(int n_host, const int* d_n, int max_pairs, int tiles,
                            const uint64_t* keys, int* ranges, hipStream_t stream);
int render_plain_launch(int n_host, const int* d_n, int W, int H, const float* xy,
                        const float* rgbd, const float* con, const uint64_t* keys,
                        const uint32_t* order, const float bg[3], uint8_t* out,
                        hipStream_t stream);
int render_persistent_launch(int n_host, const int* d_n, int W, int H, int feat_prec,
                             const float* xy, const float* rgbd, const float* con, const void* rec,
                             int gathered, const uint64_t* keys, const uint32_t* order,
                             const int* ranges, int dt, const float bg[3], uint8_t* out,
                             unsigned* queue, hipStream_t stream);

namespace {

thread_local char g_last_error[512] = "";

// Pinned slot for count readbacks (the reference keeps the same kind of slot
// per thread, sort.cu:39-69).
struct PinnedSlot {
  int* host = nullptr;
  hipEvent_t ev = nullptr;
  bool ok = false;
  PinnedSlot() {
    if (hipHostMalloc(reinterpret_cast<void**>(&host), 4 * sizeof(int), hipHostMallocDefault) !=
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

constexpr size_t kQueueBytes = 256;

}  // namespace

void set_last_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

// Copies n ints from device to the pinned slot and waits.
static int read_back(const int* const* ptrs, int n, int* out, hipStream_t stream) {
  PinnedSlot& s = pinned();
  if (!s.ok) {
    for (int i = 0; i < n; ++i)
      PS_HIP_CHECK(hipMemcpy(out + i, ptrs[i], sizeof(int), hipMemcpyDeviceToHost));
    return PS_OK;
  }
  for (int i = 0; i < n; ++i)
    PS_HIP_CHECK(hipMemcpyAsync(s.host + i, ptrs[i], sizeof(int), hipMemcpyDeviceToHost, stream));
  PS_HIP_CHECK(hipEventRecord(s.ev, stream));
  PS_HIP_CHECK(hipEventSynchronize(s.ev));
  for (int i = 0; i < n; ++i) out[i] = s.host[i];
  return PS_OK;
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

extern "C" size_t ps_forward_workspace_size(int P) {
  return ps::align_up(ps_preprocess_workspace_size(P), 256) + ps::kQueueBytes;
}

extern "C" int ps_forward(int P, const float* positions, const void* shs, int sh_prec,
                          const float* opacities, const float* cov3d, const ps_camera* cam,
                          int variant, int feat_prec, const ps_frame_buffers* fb,
                          const float bg[3], uint8_t* out_color, int sync_count,
                          int* num_rendered_out, ps_stream_t stream_) {
  using namespace ps;
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  PS_REQUIRE(fb != nullptr && cam != nullptr && bg != nullptr, "NULL argument");
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

  char* ws = static_cast<char*>(fb->workspace);
  const size_t pre_bytes = align_up(ps_preprocess_workspace_size(P), 256);
  unsigned* queue = reinterpret_cast<unsigned*>(ws + pre_bytes);
  int st;
  // The frame's keys start at offset 0 (the reference clears curr_offset
  // first, pybind.cpp:946); the scan kernel takes 0 as the base directly.
  st = preprocess_launch(P, positions, shs, sh_prec, opacities, cov3d, cam,
                         f32 ? fb->points_xy : nullptr, f32 ? fb->rgb_depth : nullptr,
                         f32 ? fb->conic_opacity : nullptr, f32 ? nullptr : fb->records,
                         feat_prec, fb->keys_unsorted, fb->values_unsorted, fb->capacity,
                         fb->curr_offset, ws, pre_bytes, stream, true);
  if (st != PS_OK) return st;
  const int* d_n = ps_preprocess_count_ptr(ws);
  st = sort_launch(0, d_n, fb->capacity, sort_end_bit(cam->width, cam->height, 16, 16),
                   fb->sort_space, fb->sort_space_bytes, fb->keys_unsorted, fb->values_unsorted,
                   fb->keys_sorted, fb->values_sorted, stream);
  if (st != PS_OK) return st;
  const int tiles = ((cam->width + 15) / 16) * ((cam->height + 15) / 16);
  if (variant == PS_VARIANT_DEFAULT) {
    st = render_plain_launch(0, d_n, cam->width, cam->height, fb->points_xy, fb->rgb_depth,
                             fb->conic_opacity, fb->keys_sorted, fb->values_sorted, bg, out_color,
                             stream);
  } else {
    if (need_ranges) {
      st = tile_ranges_scan_launch(0, d_n, fb->capacity, tiles, fb->keys_sorted, fb->tile_ranges,
                                   stream);
      if (st != PS_OK) return st;
    }
    st = render_persistent_launch(0, d_n, cam->width, cam->height, feat_prec, fb->points_xy,
                                  fb->rgb_depth, fb->conic_opacity, fb->records, 0,
                                  fb->keys_sorted, fb->values_sorted,
                                  need_ranges ? fb->tile_ranges : nullptr,
                                  variant == PS_VARIANT_PERSISTENT_DT, bg, out_color, queue,
                                  stream);
  }
  if (st != PS_OK) return st;
  if (sync_count) {
    const int* ptrs[2] = {fb->curr_offset, ps_preprocess_status_ptr(ws)};
    int vals[2] = {0, 0};
    st = read_back(ptrs, 2, vals, stream);
    if (st != PS_OK) return st;
    if (num_rendered_out) *num_rendered_out = vals[0];
    if (vals[1] & 1) {
      set_last_error("%d pairs do not fit in %d", vals[0], fb->capacity);
      return PS_ERR_OVERFLOW;
    }
  }
  return PS_OK;
}
