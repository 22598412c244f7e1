// pybind11 module `polysplat_amd._C`, exposed to Python as `polysplat_amd.ops`.
//
// It keeps the names, positional signatures and in-place output convention
// of the reference's `polysplat.ops` hot-path bindings (pybind.cpp:122-181,
// :333-364, :441-459, :616-632, :691-745, :932-1149).  Each binding only
// unpacks tensors into raw pointers and calls the C ABI in
// include/polysplat_amd.h on torch's current HIP stream; no arithmetic
// happens here.  Errors become RuntimeError (the reference raises
// RuntimeError for CUB failures, sort.cu:121-124).

#include <c10/hip/HIPStream.h>
#include <torch/extension.h>

#include <map>
#include <mutex>
#include <string>

#include "polysplat_amd.h"

namespace ps_ply {
std::tuple<int64_t, at::Tensor, at::Tensor, at::Tensor, at::Tensor> load_ply(
    const std::string& name);
}

namespace {

using at::Tensor;

ps_stream_t cur_stream() {
  return reinterpret_cast<ps_stream_t>(c10::hip::getCurrentHIPStream().stream());
}

void check(int status, const char* what) {
  if (status != PS_OK) {
    throw std::runtime_error(std::string(what) + ": " + ps_status_string(status) + " (" +
                             ps_last_error() + ")");
  }
}

void need(bool cond, const std::string& msg) {
  if (!cond) throw std::invalid_argument(msg);
}

void need_dev(const Tensor& t, const char* name, at::ScalarType dt) {
  need(t.is_cuda(), std::string(name) + " must be a GPU tensor");
  need(t.scalar_type() == dt, std::string(name) + " has dtype " +
                                  std::string(c10::toString(t.scalar_type())) + ", expected " +
                                  std::string(c10::toString(dt)));
  need(t.is_contiguous(), std::string(name) + " must be contiguous");
}

template <typename T>
T* ptr(const Tensor& t) {
  return reinterpret_cast<T*>(t.data_ptr());
}

std::vector<float> host_floats(const Tensor& t, int64_t n, const char* name) {
  Tensor c = t.detach().to(at::kCPU, at::kFloat).contiguous();
  need(c.numel() >= n, std::string(name) + " needs " + std::to_string(n) + " values");
  const float* p = c.data_ptr<float>();
  return std::vector<float>(p, p + n);
}

ps_camera make_camera(int width, int height, int block_x, int block_y, const Tensor& position,
                      const Tensor& rotation, float fx, float fy, float zfar, float znear) {
  ps_camera c;
  c.width = width;
  c.height = height;
  c.block_x = block_x;
  c.block_y = block_y;
  auto p = host_floats(position, 3, "position");
  auto r = host_floats(rotation, 9, "rotation");
  for (int i = 0; i < 3; ++i) c.position[i] = p[i];
  for (int i = 0; i < 9; ++i) c.rotation[i] = r[i];
  c.focal_x = fx;
  c.focal_y = fy;
  c.z_far = zfar;
  c.z_near = znear;
  return c;
}

int sh_prec_of(const Tensor& shs) {
  switch (shs.scalar_type()) {
    case at::kFloat: return PS_PREC_F32;
    case at::kHalf: return PS_PREC_F16;
    case at::kBFloat16: return PS_PREC_BF16;
    default: throw std::invalid_argument("shs must be float32, float16 or bfloat16");
  }
}

void check_scene(const Tensor& pts, const Tensor& shs, const Tensor& op, const Tensor& cov) {
  const int64_t P = op.size(0);
  need_dev(pts, "orig_points", at::kFloat);
  need_dev(op, "opacities", at::kFloat);
  need_dev(cov, "cov3Ds", at::kFloat);
  need(shs.is_cuda() && shs.is_contiguous(), "shs must be a contiguous GPU tensor");
  need(pts.numel() == P * 3, "orig_points must be [P,3]");
  need(shs.numel() == P * 48, "shs must be [P,48] (or [P,16,3])");
  need(cov.numel() == P * 6, "cov3Ds must be [P,6]");
}

// Scratch owned by the module for bindings whose reference signature has no
// workspace argument (preprocess workspaces, the persistent renderer's tile
// queue, the fused forwards' workspace).  One buffer per (purpose, device,
// stream): calls on different streams never share a queue or a workspace,
// and calls on one stream are ordered, so a buffer is never in use by two
// kernels at once.  Created zero-filled on the calling stream (the queue
// words must start at 0; every launch leaves them at 0) and grown on demand;
// a grown buffer replaces one that only earlier work on the same stream
// used, so torch's caching allocator hands its memory out again only to
// work ordered after it.  Steady-state calls allocate nothing.
Tensor cached_workspace(const c10::Device& dev, size_t bytes, const char* tag) {
  static std::mutex mu;
  static std::map<std::tuple<std::string, int, uintptr_t>, Tensor> cache;
  std::lock_guard<std::mutex> g(mu);
  const uintptr_t s = reinterpret_cast<uintptr_t>(cur_stream());
  auto key = std::make_tuple(std::string(tag), (int)dev.index(), s);
  auto it = cache.find(key);
  if (it == cache.end() || (size_t)it->second.numel() < bytes) {
    Tensor t = at::zeros({(int64_t)(bytes > 0 ? bytes : 1)},
                         at::TensorOptions().dtype(at::kByte).device(dev));
    cache[key] = t;
    return t;
  }
  return it->second;
}

// ------------------------------------------------------------------ preprocess
void preprocess_impl(Tensor& orig_points, Tensor& shs, Tensor& opacities, Tensor& cov3Ds,
                     int width, int height, int block_x, int block_y, Tensor& position,
                     Tensor& rotation, float focal_x, float focal_y, float zFar, float zNear,
                     Tensor& points_xy, Tensor& rgb_depth, Tensor& conic_opacity,
                     Tensor& keys_unsorted, Tensor& values_unsorted, Tensor& curr_offset) {
  check_scene(orig_points, shs, opacities, cov3Ds);
  need_dev(points_xy, "points_xy", at::kFloat);
  need_dev(rgb_depth, "rgb_depth", at::kFloat);
  need_dev(conic_opacity, "conic_opacity", at::kFloat);
  need_dev(keys_unsorted, "gaussian_keys_unsorted", at::kLong);
  need_dev(values_unsorted, "gaussian_values_unsorted", at::kInt);
  need_dev(curr_offset, "curr_offset", at::kInt);
  const int P = (int)opacities.size(0);
  need(points_xy.numel() >= 2LL * P && rgb_depth.numel() >= 4LL * P &&
           conic_opacity.numel() >= 4LL * P,
       "feature buffers must hold P entries");
  const int cap = (int)std::min<int64_t>(keys_unsorted.numel(), values_unsorted.numel());
  ps_camera cam = make_camera(width, height, block_x, block_y, position, rotation, focal_x,
                              focal_y, zFar, zNear);
  const size_t wsb = ps_preprocess_workspace_size(P);
  Tensor ws = cached_workspace(opacities.device(), wsb, "preprocess");
  check(ps_preprocess(P, ptr<float>(orig_points), shs.data_ptr(), sh_prec_of(shs),
                      ptr<float>(opacities), ptr<float>(cov3Ds), &cam, ptr<float>(points_xy),
                      ptr<float>(rgb_depth), ptr<float>(conic_opacity), nullptr, PS_PREC_F16,
                      ptr<uint64_t>(keys_unsorted), ptr<uint32_t>(values_unsorted), cap,
                      ptr<int>(curr_offset), ws.data_ptr(), wsb, cur_stream()),
        "preprocess");
}

// Same as preprocess, writing the 24-byte reduced-precision records too.
void preprocess_records(Tensor& orig_points, Tensor& shs, Tensor& opacities, Tensor& cov3Ds,
                        int width, int height, int block_x, int block_y, Tensor& position,
                        Tensor& rotation, float focal_x, float focal_y, float zFar, float zNear,
                        Tensor& records, int feat_prec, Tensor& keys_unsorted,
                        Tensor& values_unsorted, Tensor& curr_offset) {
  check_scene(orig_points, shs, opacities, cov3Ds);
  const int P = (int)opacities.size(0);
  need_dev(records, "records", at::kByte);
  need(records.numel() >= 24LL * P, "records must hold 24 bytes per Gaussian");
  need_dev(keys_unsorted, "gaussian_keys_unsorted", at::kLong);
  need_dev(values_unsorted, "gaussian_values_unsorted", at::kInt);
  need_dev(curr_offset, "curr_offset", at::kInt);
  const int cap = (int)std::min<int64_t>(keys_unsorted.numel(), values_unsorted.numel());
  ps_camera cam = make_camera(width, height, block_x, block_y, position, rotation, focal_x,
                              focal_y, zFar, zNear);
  const size_t wsb = ps_preprocess_workspace_size(P);
  Tensor ws = cached_workspace(opacities.device(), wsb, "preprocess");
  check(ps_preprocess(P, ptr<float>(orig_points), shs.data_ptr(), sh_prec_of(shs),
                      ptr<float>(opacities), ptr<float>(cov3Ds), &cam, nullptr, nullptr, nullptr,
                      records.data_ptr(), feat_prec, ptr<uint64_t>(keys_unsorted),
                      ptr<uint32_t>(values_unsorted), cap, ptr<int>(curr_offset), ws.data_ptr(),
                      wsb, cur_stream()),
        "preprocess_records");
}

// ------------------------------------------------------------------ depth first
// Depth-first preprocess: count pass, depth sort of the Gaussians and the
// (tile id, Gaussian) pairs in depth order (one call for the reference's
// preprocess_es_pass_a[_half_sh], es_depth_sort_and_scan and
// preprocess_es_pass_b, pybind.cpp:206-304).  feat_prec 0 writes the FP32
// feature arrays, 1/2 the FP16/BF16 records.  The tile ids start at 0;
// curr_offset receives the pair count.
void preprocess_es(Tensor& orig_points, Tensor& shs, Tensor& opacities, Tensor& cov3Ds, int width,
                   int height, int block_x, int block_y, Tensor& position, Tensor& rotation,
                   float focal_x, float focal_y, float zFar, float zNear, Tensor& points_xy,
                   Tensor& rgb_depth, Tensor& conic_opacity, Tensor& records, int feat_prec,
                   Tensor& tile_keys_unsorted, Tensor& values_unsorted, Tensor& curr_offset,
                   Tensor& list_sorting_space) {
  check_scene(orig_points, shs, opacities, cov3Ds);
  const int P = (int)opacities.size(0);
  need(feat_prec == PS_PREC_F32 || feat_prec == PS_PREC_F16 || feat_prec == PS_PREC_BF16,
       "feat_prec must be 0 (FP32), 1 (FP16) or 2 (BF16)");
  const bool f32 = feat_prec == PS_PREC_F32;
  if (f32) {
    need_dev(points_xy, "points_xy", at::kFloat);
    need_dev(rgb_depth, "rgb_depth", at::kFloat);
    need_dev(conic_opacity, "conic_opacity", at::kFloat);
    need(points_xy.numel() >= 2LL * P && rgb_depth.numel() >= 4LL * P &&
             conic_opacity.numel() >= 4LL * P,
         "feature buffers must hold P entries");
  } else {
    need_dev(records, "records", at::kByte);
    need(records.numel() >= 24LL * P, "records must hold 24 bytes per Gaussian");
  }
  need_dev(tile_keys_unsorted, "tile_keys_unsorted", at::kInt);
  need_dev(values_unsorted, "gaussian_values_unsorted", at::kInt);
  need_dev(curr_offset, "curr_offset", at::kInt);
  need_dev(list_sorting_space, "list_sorting_space", at::kChar);
  const int cap = (int)std::min<int64_t>(tile_keys_unsorted.numel(), values_unsorted.numel());
  ps_camera cam = make_camera(width, height, block_x, block_y, position, rotation, focal_x,
                              focal_y, zFar, zNear);
  const size_t wsb = ps_preprocess_workspace_size(P);
  Tensor ws = cached_workspace(opacities.device(), wsb, "preprocess");
  check(ps_preprocess_depth_first(
            P, ptr<float>(orig_points), shs.data_ptr(), sh_prec_of(shs), ptr<float>(opacities),
            ptr<float>(cov3Ds), &cam, f32 ? ptr<float>(points_xy) : nullptr,
            f32 ? ptr<float>(rgb_depth) : nullptr, f32 ? ptr<float>(conic_opacity) : nullptr,
            f32 ? nullptr : records.data_ptr(), feat_prec, ptr<uint32_t>(tile_keys_unsorted),
            ptr<uint32_t>(values_unsorted), cap, ptr<int>(curr_offset), ws.data_ptr(), wsb,
            list_sorting_space.data_ptr(), (size_t)list_sorting_space.numel(), cur_stream()),
        "preprocess_es");
}

// es_tile_sort(num_rendered, width, height, block_x, block_y, tile_keys_unsorted,
// gauss_values_unsorted, tile_keys_sorted, gauss_values_sorted,
// gaussian_keys_sorted_out, cub_sort_scratch) as bound at pybind.cpp:306-327;
// tile_ranges (optional) also receives the tile ranges.
void es_tile_sort(int num_rendered, int width, int height, int block_x, int block_y,
                  Tensor& tile_keys_unsorted, Tensor& values_unsorted, Tensor& tile_keys_sorted,
                  Tensor& values_sorted, Tensor& keys_sorted_out, Tensor& list_sorting_space,
                  c10::optional<Tensor> tile_ranges) {
  need_dev(tile_keys_unsorted, "tile_keys_unsorted", at::kInt);
  need_dev(values_unsorted, "gauss_values_unsorted", at::kInt);
  need_dev(tile_keys_sorted, "tile_keys_sorted", at::kInt);
  need_dev(values_sorted, "gauss_values_sorted", at::kInt);
  need_dev(keys_sorted_out, "gaussian_keys_sorted_out", at::kLong);
  need_dev(list_sorting_space, "list_sorting_space", at::kChar);
  const int cap = (int)std::min({tile_keys_unsorted.numel(), values_unsorted.numel(),
                                 tile_keys_sorted.numel(), values_sorted.numel(),
                                 keys_sorted_out.numel()});
  int max_pairs = cap;
  const size_t have = (size_t)list_sorting_space.numel();
  while (max_pairs > 0 && ps_get_sort_buffer_size(max_pairs) > have) max_pairs /= 2;
  need(num_rendered >= 0 && num_rendered <= max_pairs,
       "list_sorting_space or the pair buffers too small for num_rendered");
  int* rng = nullptr;
  if (tile_ranges.has_value()) {
    need_dev(*tile_ranges, "tile_ranges", at::kInt);
    const int64_t tiles = (int64_t)((width + block_x - 1) / block_x) *
                          ((height + block_y - 1) / block_y);
    need(tile_ranges->numel() >= 2 * tiles, "tile_ranges must hold [tiles, 2]");
    rng = ptr<int>(*tile_ranges);
  }
  check(ps_es_tile_sort(num_rendered, nullptr, max_pairs, width, height, block_x, block_y,
                        list_sorting_space.data_ptr(), have, ptr<uint32_t>(tile_keys_unsorted),
                        ptr<uint32_t>(values_unsorted), ptr<uint32_t>(tile_keys_sorted),
                        ptr<uint32_t>(values_sorted), ptr<uint64_t>(keys_sorted_out), rng,
                        cur_stream()),
        "es_tile_sort");
}

// ------------------------------------------------------------------ sort
size_t get_sort_buffer_size(int64_t num_rendered) {
  need(num_rendered >= 0 && num_rendered < (1LL << 31), "num_rendered out of range");
  return ps_get_sort_buffer_size((int)num_rendered);
}

void sort_gaussian(int num_rendered, int width, int height, int block_x, int block_y,
                   Tensor& list_sorting_space, Tensor& keys_unsorted, Tensor& values_unsorted,
                   Tensor& keys_sorted, Tensor& values_sorted) {
  need_dev(list_sorting_space, "list_sorting_space", at::kChar);
  need_dev(keys_unsorted, "gaussian_keys_unsorted", at::kLong);
  need_dev(values_unsorted, "gaussian_values_unsorted", at::kInt);
  need_dev(keys_sorted, "gaussian_keys_sorted", at::kLong);
  need_dev(values_sorted, "gaussian_values_sorted", at::kInt);
  const int cap = (int)std::min({keys_unsorted.numel(), values_unsorted.numel(),
                                 keys_sorted.numel(), values_sorted.numel()});
  // The sort space was sized for some capacity; find the largest it serves.
  int max_pairs = cap;
  const size_t have = (size_t)list_sorting_space.numel();
  while (max_pairs > 0 && ps_get_sort_buffer_size(max_pairs) > have) max_pairs /= 2;
  need(num_rendered <= max_pairs, "list_sorting_space too small for num_rendered");
  check(ps_sort_gaussian(num_rendered, nullptr, max_pairs, width, height, block_x, block_y,
                         list_sorting_space.data_ptr(), have, ptr<uint64_t>(keys_unsorted),
                         ptr<uint32_t>(values_unsorted), ptr<uint64_t>(keys_sorted),
                         ptr<uint32_t>(values_sorted), cur_stream()),
        "sort_gaussian");
}

// ------------------------------------------------------------------ gather / ranges
void gather_features(int num_rendered, Tensor& values_sorted, Tensor& src_xy, Tensor& src_rgbd,
                     Tensor& src_conic, Tensor& dst_xy, Tensor& dst_rgbd, Tensor& dst_conic) {
  need_dev(values_sorted, "sorted_values", at::kInt);
  for (auto* t : {&src_xy, &src_rgbd, &src_conic, &dst_xy, &dst_rgbd, &dst_conic})
    need_dev(*t, "feature", at::kFloat);
  need(num_rendered >= 0 && num_rendered <= values_sorted.numel(), "num_rendered out of range");
  need(dst_xy.numel() >= 2LL * num_rendered && dst_rgbd.numel() >= 4LL * num_rendered &&
           dst_conic.numel() >= 4LL * num_rendered,
       "destination buffers too small");
  check(ps_gather_features(num_rendered, nullptr, num_rendered, ptr<uint32_t>(values_sorted),
                           ptr<float>(src_xy), ptr<float>(src_rgbd), ptr<float>(src_conic),
                           ptr<float>(dst_xy), ptr<float>(dst_rgbd), ptr<float>(dst_conic),
                           cur_stream()),
        "gather_features");
}

int64_t tiles_of(int width, int height, int bx, int by) {
  return (int64_t)((width + bx - 1) / bx) * ((height + by - 1) / by);
}

void precompute_tile_ranges(int num_rendered, int width, int height, int block_x, int block_y,
                            Tensor& keys_sorted, Tensor& tile_ranges) {
  need_dev(keys_sorted, "gaussian_keys_sorted", at::kLong);
  need_dev(tile_ranges, "tile_ranges_buf", at::kInt);
  need(tile_ranges.numel() >= 2 * tiles_of(width, height, block_x, block_y),
       "tile_ranges_buf too small");
  check(ps_precompute_tile_ranges(num_rendered, nullptr, width, height, block_x, block_y,
                                  ptr<uint64_t>(keys_sorted), ptr<int>(tile_ranges), cur_stream()),
        "precompute_tile_ranges");
}

void precompute_tile_ranges_scan(int num_rendered, int width, int height, int block_x,
                                 int block_y, Tensor& keys_sorted, Tensor& tile_ranges) {
  need_dev(keys_sorted, "gaussian_keys_sorted", at::kLong);
  need_dev(tile_ranges, "tile_ranges_buf", at::kInt);
  need(tile_ranges.numel() >= 2 * tiles_of(width, height, block_x, block_y),
       "tile_ranges_buf too small");
  check(ps_precompute_tile_ranges_scan(num_rendered, nullptr, num_rendered, width, height,
                                       block_x, block_y, ptr<uint64_t>(keys_sorted),
                                       ptr<int>(tile_ranges), cur_stream()),
        "precompute_tile_ranges_scan");
}

// ------------------------------------------------------------------ render
void check_out(const Tensor& out, int width, int height) {
  need_dev(out, "out_color", at::kByte);
  need(out.numel() == 3LL * width * height, "out_color must be [H,W,3]");
}

void check_feats(const Tensor& xy, const Tensor& rgbd, const Tensor& con) {
  need_dev(xy, "points_xy", at::kFloat);
  need_dev(rgbd, "rgb_depth", at::kFloat);
  need_dev(con, "conic_opacity", at::kFloat);
}

void render_16x16(int num_rendered, int width, int height, Tensor& points_xy, Tensor& rgb_depth,
                  Tensor& conic_opacity, Tensor& keys_sorted, Tensor& values_sorted,
                  Tensor& bg_color, Tensor& out_color) {
  check_feats(points_xy, rgb_depth, conic_opacity);
  need_dev(keys_sorted, "gaussian_keys_sorted", at::kLong);
  need_dev(values_sorted, "gaussian_values_sorted", at::kInt);
  check_out(out_color, width, height);
  auto bg = host_floats(bg_color, 3, "bg_color");
  check(ps_render_16x16(num_rendered, nullptr, width, height, ptr<float>(points_xy),
                        ptr<float>(rgb_depth), ptr<float>(conic_opacity),
                        ptr<uint64_t>(keys_sorted), ptr<uint32_t>(values_sorted), bg.data(),
                        ptr<uint8_t>(out_color), cur_stream()),
        "render_16x16");
}

void persistent_common(int num_rendered, int width, int height, Tensor* xy, Tensor* rgbd,
                       Tensor* con, Tensor* records, int prec, bool gathered, Tensor* keys,
                       Tensor& values, Tensor* ranges, bool dt, Tensor& bg_color, Tensor& out,
                       const char* what) {
  if (prec == PS_PREC_F32) check_feats(*xy, *rgbd, *con);
  else need_dev(*records, "records", at::kByte);
  need_dev(values, "gaussian_values_sorted", at::kInt);
  if (keys) need_dev(*keys, "gaussian_keys_sorted", at::kLong);
  if (ranges) {
    need_dev(*ranges, "tile_ranges_buf", at::kInt);
    need(ranges->numel() >= 2 * tiles_of(width, height, 16, 16), "tile_ranges_buf too small");
  }
  check_out(out, width, height);
  auto bg = host_floats(bg_color, 3, "bg_color");
  Tensor q = cached_workspace(out.device(), 2048, "queue");
  check(ps_render_16x16_persistent(
            num_rendered, nullptr, width, height, prec, xy ? ptr<float>(*xy) : nullptr,
            rgbd ? ptr<float>(*rgbd) : nullptr, con ? ptr<float>(*con) : nullptr,
            records ? records->data_ptr() : nullptr, gathered ? 1 : 0,
            keys ? ptr<uint64_t>(*keys) : nullptr, ptr<uint32_t>(values),
            ranges ? ptr<int>(*ranges) : nullptr, dt ? 1 : 0, bg.data(), ptr<uint8_t>(out),
            ptr<unsigned>(q), cur_stream()),
        what);
}

void render_lite(int num_rendered, int width, int height, Tensor& xy, Tensor& rgbd, Tensor& con,
                 Tensor& values_sorted, Tensor& tile_ranges, Tensor& bg, Tensor& out) {
  persistent_common(num_rendered, width, height, &xy, &rgbd, &con, nullptr, PS_PREC_F32, false,
                    nullptr, values_sorted, &tile_ranges, false, bg, out,
                    "render_16x16_preranges_smem_persistent_lite");
}

void render_lite_dt(int num_rendered, int width, int height, Tensor& xy, Tensor& rgbd,
                    Tensor& con, Tensor& values_sorted, Tensor& tile_ranges, Tensor& bg,
                    Tensor& out) {
  persistent_common(num_rendered, width, height, &xy, &rgbd, &con, nullptr, PS_PREC_F32, false,
                    nullptr, values_sorted, &tile_ranges, true, bg, out,
                    "render_16x16_preranges_smem_persistent_lite_dt");
}

void render_lite_fused(int num_rendered, int width, int height, Tensor& xy, Tensor& rgbd,
                       Tensor& con, Tensor& keys_sorted, Tensor& values_sorted, Tensor& bg,
                       Tensor& out) {
  persistent_common(num_rendered, width, height, &xy, &rgbd, &con, nullptr, PS_PREC_F32, false,
                    &keys_sorted, values_sorted, nullptr, false, bg, out,
                    "render_16x16_preranges_smem_persistent_lite_fused");
}

// Persistent render of 24-byte records (FP16/BF16 features); dt skips
// pixels whose alpha is below 1/255, as the _dt FP32 variant does.
void render_records(int num_rendered, int width, int height, Tensor& records, int feat_prec,
                    bool gathered, Tensor& values_sorted, Tensor& tile_ranges, Tensor& bg,
                    Tensor& out, bool dt) {
  need(feat_prec == PS_PREC_F16 || feat_prec == PS_PREC_BF16,
       "render_records: feat_prec must be 1 (fp16) or 2 (bf16)");
  persistent_common(num_rendered, width, height, nullptr, nullptr, nullptr, &records, feat_prec,
                    gathered, nullptr, values_sorted, &tile_ranges, dt, bg, out,
                    "render_records");
}

// ------------------------------------------------------------------ tile reordering
int64_t get_tile_order_sort_temp_size(int total_tiles) {
  return (int64_t)ps_get_tile_order_sort_temp_size(total_tiles);
}

void check_order_bufs(int total_tiles, Tensor& tile_ranges, Tensor& tile_order,
                      Tensor& tile_counts, Tensor& tile_ids, Tensor& sort_temp) {
  need(total_tiles >= 0, "total_tiles must be >= 0");
  need_dev(tile_ranges, "tile_ranges_buf", at::kInt);
  need_dev(tile_order, "tile_order", at::kInt);
  need_dev(tile_counts, "tile_counts_buf", at::kInt);
  need_dev(tile_ids, "tile_ids_buf", at::kInt);
  need(sort_temp.is_cuda() && sort_temp.is_contiguous(), "sort_temp must be a GPU tensor");
  need(tile_ranges.numel() >= 2LL * total_tiles && tile_order.numel() >= total_tiles &&
           tile_counts.numel() >= total_tiles && tile_ids.numel() >= total_tiles,
       "tile buffers smaller than total_tiles");
  need((size_t)sort_temp.nbytes() >= ps_get_tile_order_sort_temp_size(total_tiles),
       "sort_temp smaller than get_tile_order_sort_temp_size(total_tiles)");
}

void compute_tile_order(Tensor& tile_ranges, int total_tiles, Tensor& tile_order,
                        Tensor& tile_counts, Tensor& tile_ids, Tensor& sort_temp) {
  check_order_bufs(total_tiles, tile_ranges, tile_order, tile_counts, tile_ids, sort_temp);
  check(ps_compute_tile_order(ptr<int>(tile_ranges), total_tiles, ptr<uint32_t>(tile_order),
                              ptr<uint32_t>(tile_counts), ptr<uint32_t>(tile_ids),
                              sort_temp.data_ptr(), (size_t)sort_temp.nbytes(), cur_stream()),
        "compute_tile_order");
}

void compute_tile_order_packed_common(Tensor& tile_ranges, int total_tiles, int x_blocks,
                                      Tensor& tile_desc, Tensor& tile_order, Tensor& tile_counts,
                                      Tensor& tile_ids, Tensor& sort_temp, int zigzag,
                                      int bucket, const char* what) {
  check_order_bufs(total_tiles, tile_ranges, tile_order, tile_counts, tile_ids, sort_temp);
  need_dev(tile_desc, "tile_desc_buf", at::kInt);
  need(tile_desc.numel() >= 4LL * total_tiles, "tile_desc_buf smaller than total_tiles x 4");
  need(x_blocks > 0, "x_blocks must be positive");
  check(ps_compute_tile_order_packed(ptr<int>(tile_ranges), total_tiles, x_blocks,
                                     ptr<int>(tile_desc), ptr<uint32_t>(tile_order),
                                     ptr<uint32_t>(tile_counts), ptr<uint32_t>(tile_ids),
                                     sort_temp.data_ptr(), (size_t)sort_temp.nbytes(), zigzag,
                                     bucket, cur_stream()),
        what);
}

void compute_tile_order_packed(Tensor& tile_ranges, int total_tiles, int x_blocks,
                               Tensor& tile_desc, Tensor& tile_order, Tensor& tile_counts,
                               Tensor& tile_ids, Tensor& sort_temp, int zigzag_group_size) {
  compute_tile_order_packed_common(tile_ranges, total_tiles, x_blocks, tile_desc, tile_order,
                                   tile_counts, tile_ids, sort_temp, zigzag_group_size, 0,
                                   "compute_tile_order_packed");
}

void compute_tile_order_packed_morton(Tensor& tile_ranges, int total_tiles, int x_blocks,
                                      Tensor& tile_desc, Tensor& tile_order, Tensor& tile_counts,
                                      Tensor& tile_ids, Tensor& sort_temp, int zigzag_group_size,
                                      int morton_bucket_size) {
  need(morton_bucket_size > 0, "morton_bucket_size must be positive");
  compute_tile_order_packed_common(tile_ranges, total_tiles, x_blocks, tile_desc, tile_order,
                                   tile_counts, tile_ids, sort_temp, zigzag_group_size,
                                   morton_bucket_size, "compute_tile_order_packed_morton");
}

void render_16x16_reordered(int num_rendered, int width, int height, Tensor& points_xy,
                            Tensor& rgb_depth, Tensor& conic_opacity, Tensor& keys_sorted,
                            Tensor& values_sorted, Tensor& tile_ranges, Tensor& tile_order,
                            Tensor& bg_color, Tensor& out_color) {
  check_feats(points_xy, rgb_depth, conic_opacity);
  need_dev(keys_sorted, "gaussian_keys_sorted", at::kLong);
  need_dev(values_sorted, "gaussian_values_sorted", at::kInt);
  need_dev(tile_ranges, "tile_ranges_buf", at::kInt);
  need_dev(tile_order, "tile_order", at::kInt);
  const int64_t T = tiles_of(width, height, 16, 16);
  need(tile_ranges.numel() >= 2 * T && tile_order.numel() >= T,
       "tile_ranges_buf or tile_order smaller than the frame's tiles");
  check_out(out_color, width, height);
  auto bg = host_floats(bg_color, 3, "bg_color");
  check(ps_render_16x16_reordered(num_rendered, nullptr, width, height, ptr<float>(points_xy),
                                  ptr<float>(rgb_depth), ptr<float>(conic_opacity),
                                  ptr<uint64_t>(keys_sorted), ptr<uint32_t>(values_sorted),
                                  ptr<int>(tile_ranges), ptr<uint32_t>(tile_order), bg.data(),
                                  ptr<uint8_t>(out_color), cur_stream()),
        "render_16x16_reordered");
}

void render_16x16_reordered_v2(int num_rendered, int width, int height, Tensor& points_xy,
                               Tensor& rgb_depth, Tensor& conic_opacity, Tensor& keys_sorted,
                               Tensor& values_sorted, Tensor& tile_desc, Tensor& bg_color,
                               Tensor& out_color) {
  check_feats(points_xy, rgb_depth, conic_opacity);
  need_dev(keys_sorted, "gaussian_keys_sorted", at::kLong);
  need_dev(values_sorted, "gaussian_values_sorted", at::kInt);
  need_dev(tile_desc, "tile_desc_buf", at::kInt);
  need(tile_desc.numel() >= 4 * tiles_of(width, height, 16, 16),
       "tile_desc_buf smaller than the frame's tiles x 4");
  check_out(out_color, width, height);
  auto bg = host_floats(bg_color, 3, "bg_color");
  check(ps_render_16x16_reordered_desc(num_rendered, nullptr, width, height,
                                       ptr<float>(points_xy), ptr<float>(rgb_depth),
                                       ptr<float>(conic_opacity), ptr<uint32_t>(values_sorted),
                                       ptr<int>(tile_desc), bg.data(), ptr<uint8_t>(out_color),
                                       cur_stream()),
        "render_16x16_reordered_v2");
}

// Top-k cache renders (pybind.cpp:366-437): values may hold 0x80000000 |
// slot for the compact top-k feature arrays.
void topk_common(int num_rendered, int width, int height, Tensor& xy, Tensor& rgbd, Tensor& con,
                 Tensor& keys_sorted, Tensor& values_sorted, Tensor& topk_xy, Tensor& topk_rgbd,
                 Tensor& topk_con, int num_topk, Tensor& bg_color, Tensor& out, bool persistent,
                 const char* what) {
  check_feats(xy, rgbd, con);
  need_dev(keys_sorted, "gaussian_keys_sorted", at::kLong);
  need_dev(values_sorted, "gaussian_values_sorted", at::kInt);
  need(num_topk >= 0, std::string(what) + ": num_topk must be >= 0");
  if (num_topk > 0) {
    check_feats(topk_xy, topk_rgbd, topk_con);
    need(topk_xy.numel() >= 2LL * num_topk && topk_rgbd.numel() >= 4LL * num_topk &&
             topk_con.numel() >= 4LL * num_topk,
         std::string(what) + ": top-k arrays smaller than num_topk");
  }
  check_out(out, width, height);
  auto bg = host_floats(bg_color, 3, "bg_color");
  Tensor q = cached_workspace(out.device(), 2048, "queue");
  check(ps_render_16x16_topk(num_rendered, nullptr, width, height, ptr<float>(xy),
                             ptr<float>(rgbd), ptr<float>(con), ptr<uint64_t>(keys_sorted),
                             ptr<uint32_t>(values_sorted),
                             num_topk > 0 ? ptr<float>(topk_xy) : nullptr,
                             num_topk > 0 ? ptr<float>(topk_rgbd) : nullptr,
                             num_topk > 0 ? ptr<float>(topk_con) : nullptr, num_topk,
                             persistent ? 1 : 0, bg.data(), ptr<uint8_t>(out), ptr<unsigned>(q),
                             cur_stream()),
        what);
}

void render_topk_smem(int num_rendered, int width, int height, Tensor& xy, Tensor& rgbd,
                      Tensor& con, Tensor& ks, Tensor& vs, Tensor& txy, Tensor& trgbd,
                      Tensor& tcon, int num_topk, Tensor& bg, Tensor& out) {
  topk_common(num_rendered, width, height, xy, rgbd, con, ks, vs, txy, trgbd, tcon, num_topk, bg,
              out, false, "render_16x16_topk_smem");
}

void render_topk_smem_persistent(int num_rendered, int width, int height, Tensor& xy,
                                 Tensor& rgbd, Tensor& con, Tensor& ks, Tensor& vs, Tensor& txy,
                                 Tensor& trgbd, Tensor& tcon, int num_topk, Tensor& bg,
                                 Tensor& out) {
  topk_common(num_rendered, width, height, xy, rgbd, con, ks, vs, txy, trgbd, tcon, num_topk, bg,
              out, true, "render_16x16_topk_smem_persistent");
}

// Persistent render over packed tile descriptors, from FP32 features
// (`records` undefined or empty) or 24-byte records (feat_prec 1 or 2).
void reordered_persistent_common(int num_rendered, int width, int height, Tensor* xy,
                                 Tensor* rgbd, Tensor* con, Tensor* records, int prec,
                                 Tensor& values_sorted, Tensor& tile_desc, bool dt, Tensor& bg_color,
                                 Tensor& out, const char* what) {
  if (prec == PS_PREC_F32) check_feats(*xy, *rgbd, *con);
  else need_dev(*records, "records", at::kByte);
  need_dev(values_sorted, "gaussian_values_sorted", at::kInt);
  need_dev(tile_desc, "tile_desc_buf", at::kInt);
  need(tile_desc.numel() >= 4 * tiles_of(width, height, 16, 16),
       "tile_desc_buf smaller than the frame's tiles x 4");
  check_out(out, width, height);
  auto bg = host_floats(bg_color, 3, "bg_color");
  Tensor q = cached_workspace(out.device(), 2048, "queue");
  check(ps_render_16x16_reordered_persistent(
            num_rendered, nullptr, width, height, prec, xy ? ptr<float>(*xy) : nullptr,
            rgbd ? ptr<float>(*rgbd) : nullptr, con ? ptr<float>(*con) : nullptr,
            records ? records->data_ptr() : nullptr, ptr<uint32_t>(values_sorted), nullptr,
            nullptr, ptr<int>(tile_desc), dt ? 1 : 0, bg.data(), ptr<uint8_t>(out),
            ptr<unsigned>(q), cur_stream()),
        what);
}

void render_16x16_reordered_persistent(int num_rendered, int width, int height,
                                       Tensor& points_xy, Tensor& rgb_depth,
                                       Tensor& conic_opacity, Tensor& keys_sorted,
                                       Tensor& values_sorted, Tensor& tile_desc, Tensor& bg_color,
                                       Tensor& out_color) {
  need_dev(keys_sorted, "gaussian_keys_sorted", at::kLong);
  reordered_persistent_common(num_rendered, width, height, &points_xy, &rgb_depth,
                              &conic_opacity, nullptr, PS_PREC_F32, values_sorted, tile_desc,
                              false, bg_color, out_color, "render_16x16_reordered_persistent");
}

void render_records_reordered(int num_rendered, int width, int height, Tensor& records,
                              int feat_prec, Tensor& values_sorted, Tensor& tile_desc, Tensor& bg,
                              Tensor& out, bool dt) {
  need(feat_prec == PS_PREC_F16 || feat_prec == PS_PREC_BF16,
       "render_records_reordered: feat_prec must be 1 (fp16) or 2 (bf16)");
  reordered_persistent_common(num_rendered, width, height, nullptr, nullptr, nullptr, &records,
                              feat_prec, values_sorted, tile_desc, dt, bg, out,
                              "render_records_reordered");
}

void gather_records(int num_rendered, Tensor& values_sorted, Tensor& src, Tensor& dst) {
  need_dev(values_sorted, "sorted_values", at::kInt);
  need_dev(src, "src", at::kByte);
  need_dev(dst, "dst", at::kByte);
  need(dst.numel() >= 24LL * num_rendered, "dst too small");
  check(ps_gather_records(num_rendered, nullptr, num_rendered, ptr<uint32_t>(values_sorted),
                          src.data_ptr(), dst.data_ptr(), cur_stream()),
        "gather_records");
}

// ------------------------------------------------------------------ forward
struct FrameArgs {
  Tensor* curr_offset;
  Tensor* sort_space;
  Tensor* ku;
  Tensor* vu;
  Tensor* ks;
  Tensor* vs;
  Tensor* xy;
  Tensor* rgbd;
  Tensor* con;
  Tensor* records;
  Tensor* ranges;
};

// The frame's buffers as ps_frame_buffers, after checking them; `ws_out`
// receives the workspace tensor used (the caller's, or the module's cached
// one when none is given and `allow_cached`).
ps_frame_buffers frame_buffers(Tensor& pts, Tensor& shs, Tensor& op, Tensor& cov, int width,
                               int height, int feat_prec, const FrameArgs& a, Tensor* workspace,
                               bool allow_cached, Tensor& ws_out, const char* what) {
  check_scene(pts, shs, op, cov);
  need_dev(*a.curr_offset, "curr_offset", at::kInt);
  need_dev(*a.sort_space, "list_sorting_space", at::kChar);
  need_dev(*a.ku, "gaussian_keys_unsorted", at::kLong);
  need_dev(*a.vu, "gaussian_values_unsorted", at::kInt);
  need_dev(*a.ks, "gaussian_keys_sorted", at::kLong);
  need_dev(*a.vs, "gaussian_values_sorted", at::kInt);
  const int P = (int)op.size(0);
  ps_frame_buffers fb;
  memset(&fb, 0, sizeof(fb));
  fb.curr_offset = ptr<int>(*a.curr_offset);
  fb.sort_space = a.sort_space->data_ptr();
  fb.sort_space_bytes = (size_t)a.sort_space->numel();
  fb.keys_unsorted = ptr<uint64_t>(*a.ku);
  fb.values_unsorted = ptr<uint32_t>(*a.vu);
  fb.keys_sorted = ptr<uint64_t>(*a.ks);
  fb.values_sorted = ptr<uint32_t>(*a.vs);
  int cap = (int)std::min({a.ku->numel(), a.vu->numel(), a.ks->numel(), a.vs->numel()});
  while (cap > 0 && ps_get_sort_buffer_size(cap) > fb.sort_space_bytes) cap /= 2;
  fb.capacity = cap;
  // The preprocess writes features of every emitting Gaussian, so the
  // feature buffers must hold all P of them.
  if (feat_prec == PS_PREC_F32) {
    check_feats(*a.xy, *a.rgbd, *a.con);
    need(a.xy->numel() >= 2LL * P && a.rgbd->numel() >= 4LL * P && a.con->numel() >= 4LL * P,
         std::string(what) + ": points_xy, rgb_depth and conic_opacity must hold P entries");
    fb.points_xy = ptr<float>(*a.xy);
    fb.rgb_depth = ptr<float>(*a.rgbd);
    fb.conic_opacity = ptr<float>(*a.con);
  } else {
    need(feat_prec == PS_PREC_F16 || feat_prec == PS_PREC_BF16,
         std::string(what) + ": unknown feature precision");
    need_dev(*a.records, "records", at::kByte);
    need(a.records->numel() >= 24LL * P,
         std::string(what) + ": records must hold 24 bytes per Gaussian");
    fb.records = a.records->data_ptr();
  }
  if (a.ranges) {
    need_dev(*a.ranges, "tile_ranges_buf", at::kInt);
    need(a.ranges->numel() >= 2 * tiles_of(width, height, 16, 16), "tile_ranges_buf too small");
    fb.tile_ranges = ptr<int>(*a.ranges);
  }
  const size_t wsb = ps_forward_workspace_size(P);
  Tensor ws;
  if (workspace && workspace->defined() && workspace->numel() > 0) {
    need_dev(*workspace, "workspace", at::kByte);
    need((size_t)workspace->numel() >= wsb, "workspace too small");
    ws = *workspace;
  } else {
    need(allow_cached, std::string(what) + ": needs a workspace of forward_workspace_size(P)");
    ws = cached_workspace(op.device(), wsb, "forward");
  }
  fb.workspace = ws.data_ptr();
  fb.workspace_bytes = (size_t)ws.numel();
  ws_out = ws;
  return fb;
}

int forward_common(Tensor& pts, Tensor& shs, Tensor& op, Tensor& cov, int width, int height,
                   int block_x, int block_y, Tensor& position, Tensor& rotation, float fx,
                   float fy, float zfar, float znear, int variant, int feat_prec,
                   const FrameArgs& a, Tensor& bg_color, Tensor& out, Tensor* workspace,
                   int sync, const char* what) {
  need(block_x == 16 && block_y == 16, "only 16x16 tiles are implemented");
  check_out(out, width, height);
  Tensor ws;
  ps_frame_buffers fb =
      frame_buffers(pts, shs, op, cov, width, height, feat_prec, a, workspace, true, ws, what);
  const int P = (int)op.size(0);
  ps_camera cam = make_camera(width, height, block_x, block_y, position, rotation, fx, fy, zfar,
                              znear);
  auto bg = host_floats(bg_color, 3, "bg_color");
  int n = 0;
  const int st = ps_forward(P, ptr<float>(pts), shs.data_ptr(), sh_prec_of(shs), ptr<float>(op),
                            ptr<float>(cov), &cam, variant, feat_prec, &fb, bg.data(),
                            ptr<uint8_t>(out), sync, &n, cur_stream());
  if (st == PS_ERR_OVERFLOW)
    throw std::runtime_error(std::string(what) + ": Too many k-v pairs! (" + ps_last_error() + ")");
  check(st, what);
  return n;
}

int forward_fused(Tensor& pts, Tensor& shs, Tensor& op, Tensor& cov, int width, int height,
                  int block_x, int block_y, Tensor& position, Tensor& rotation, float fx, float fy,
                  float zfar, float znear, const std::string& render_variant, Tensor& curr_offset,
                  Tensor& sort_space, Tensor& ku, Tensor& vu, Tensor& ks, Tensor& vs,
                  Tensor& xy, Tensor& rgbd, Tensor& con, Tensor& bg, Tensor& out) {
  need(render_variant == "default",
       "forward_fused implements render_variant 'default' only (got '" + render_variant + "')");
  FrameArgs a{&curr_offset, &sort_space, &ku, &vu, &ks, &vs, &xy, &rgbd, &con, nullptr, nullptr};
  return forward_common(pts, shs, op, cov, width, height, block_x, block_y, position, rotation,
                        fx, fy, zfar, znear, PS_VARIANT_DEFAULT, PS_PREC_F32, a, bg, out, nullptr,
                        1, "forward_fused");
}

int forward_fused_e2e(Tensor& pts, Tensor& shs_half, Tensor& op, Tensor& cov, int width,
                      int height, int block_x, int block_y, Tensor& position, Tensor& rotation,
                      float fx, float fy, float zfar, float znear, Tensor& curr_offset,
                      Tensor& sort_space, Tensor& ku, Tensor& vu, Tensor& ks, Tensor& vs,
                      Tensor& xy, Tensor& rgbd, Tensor& con, Tensor& bg, Tensor& out) {
  FrameArgs a{&curr_offset, &sort_space, &ku, &vu, &ks, &vs, &xy, &rgbd, &con, nullptr, nullptr};
  return forward_common(pts, shs_half, op, cov, width, height, block_x, block_y, position,
                        rotation, fx, fy, zfar, znear, PS_VARIANT_PERSISTENT_FUSED, PS_PREC_F32, a,
                        bg, out, nullptr, 1, "forward_fused_e2e");
}

int forward_fused_e2e_preranges(Tensor& pts, Tensor& shs_half, Tensor& op, Tensor& cov, int width,
                                int height, int block_x, int block_y, Tensor& position,
                                Tensor& rotation, float fx, float fy, float zfar, float znear,
                                Tensor& curr_offset, Tensor& sort_space, Tensor& ku, Tensor& vu,
                                Tensor& ks, Tensor& vs, Tensor& xy, Tensor& rgbd, Tensor& con,
                                Tensor& tile_ranges, Tensor& bg, Tensor& out) {
  FrameArgs a{&curr_offset, &sort_space, &ku, &vu, &ks, &vs, &xy, &rgbd, &con, nullptr,
              &tile_ranges};
  return forward_common(pts, shs_half, op, cov, width, height, block_x, block_y, position,
                        rotation, fx, fy, zfar, znear, PS_VARIANT_PERSISTENT_LITE, PS_PREC_F32, a,
                        bg, out, nullptr, 1, "forward_fused_e2e_preranges");
}

// Generic frame: any variant, any feature precision, optional caller
// workspace, optional end-of-frame count readback.
int forward_frame(Tensor& pts, Tensor& shs, Tensor& op, Tensor& cov, int width, int height,
                  Tensor& position, Tensor& rotation, float fx, float fy, float zfar, float znear,
                  int variant, int feat_prec, Tensor& curr_offset, Tensor& sort_space, Tensor& ku,
                  Tensor& vu, Tensor& ks, Tensor& vs, Tensor& xy, Tensor& rgbd, Tensor& con,
                  Tensor& records, Tensor& tile_ranges, Tensor& bg, Tensor& out,
                  Tensor& workspace, bool sync) {
  FrameArgs a{&curr_offset, &sort_space, &ku, &vu, &ks, &vs, &xy, &rgbd, &con, &records,
              tile_ranges.numel() > 0 ? &tile_ranges : nullptr};
  return forward_common(pts, shs, op, cov, width, height, 16, 16, position, rotation, fx, fy,
                        zfar, znear, variant, feat_prec, a, bg, out, &workspace, sync ? 1 : 0,
                        "forward_frame");
}

// forward_frame without waiting for the pair count: returns a ticket for
// forward_wait, which returns the count (or raises on overflow) once the
// frame has run, so frames can be enqueued back to back.
int forward_frame_post(Tensor& pts, Tensor& shs, Tensor& op, Tensor& cov, int width, int height,
                       Tensor& position, Tensor& rotation, float fx, float fy, float zfar,
                       float znear, int variant, int feat_prec, Tensor& curr_offset,
                       Tensor& sort_space, Tensor& ku, Tensor& vu, Tensor& ks, Tensor& vs,
                       Tensor& xy, Tensor& rgbd, Tensor& con, Tensor& records,
                       Tensor& tile_ranges, Tensor& bg, Tensor& out, Tensor& workspace) {
  FrameArgs a{&curr_offset, &sort_space, &ku, &vu, &ks, &vs, &xy, &rgbd, &con, &records,
              tile_ranges.numel() > 0 ? &tile_ranges : nullptr};
  return forward_common(pts, shs, op, cov, width, height, 16, 16, position, rotation, fx, fy,
                        zfar, znear, variant, feat_prec, a, bg, out, &workspace, 2,
                        "forward_frame_post");
}

// A frame's pipeline captured once as a HIP graph (ps_frame_graph_create)
// and replayed per request with a new pose, background and output image.
// Holds references to every buffer the graph points at, so none is freed
// while the graph lives.  Replays and ops calls on the same buffers must be
// issued on one stream.
class FrameGraph {
 public:
  FrameGraph(Tensor& pts, Tensor& shs, Tensor& op, Tensor& cov, int width, int height,
             Tensor& position, Tensor& rotation, float fx, float fy, float zfar, float znear,
             int variant, int feat_prec, Tensor& curr_offset, Tensor& sort_space, Tensor& ku,
             Tensor& vu, Tensor& ks, Tensor& vs, Tensor& xy, Tensor& rgbd, Tensor& con,
             Tensor& records, Tensor& tile_ranges, Tensor& workspace)
      : width_(width), height_(height) {
    FrameArgs a{&curr_offset, &sort_space, &ku, &vu, &ks, &vs, &xy, &rgbd, &con, &records,
                tile_ranges.numel() > 0 ? &tile_ranges : nullptr};
    Tensor ws;
    ps_frame_buffers fb = frame_buffers(pts, shs, op, cov, width, height, feat_prec, a,
                                        &workspace, false, ws, "FrameGraph");
    ps_camera cam = make_camera(width, height, 16, 16, position, rotation, fx, fy, zfar, znear);
    keep_ = {pts, shs, op, cov, curr_offset, sort_space, ku, vu, ks, vs, xy, rgbd, con, records,
             tile_ranges, ws};
    check(ps_frame_graph_create((int)op.size(0), ptr<float>(pts), shs.data_ptr(),
                                sh_prec_of(shs), ptr<float>(op), ptr<float>(cov), &cam, variant,
                                feat_prec, &fb, &g_),
          "FrameGraph");
  }
  ~FrameGraph() { ps_frame_graph_destroy(g_); }
  FrameGraph(const FrameGraph&) = delete;
  FrameGraph& operator=(const FrameGraph&) = delete;

  // Replays the frame on torch's current stream.  sync 1: returns the pair
  // count (RuntimeError on overflow); 0: returns -1; 2: returns a ticket for
  // forward_wait.
  int launch(Tensor& position, Tensor& rotation, float fx, float fy, float zfar, float znear,
             Tensor& bg_color, Tensor& out, int sync) {
    check_out(out, width_, height_);
    ps_camera cam = make_camera(width_, height_, 16, 16, position, rotation, fx, fy, zfar, znear);
    auto bg = host_floats(bg_color, 3, "bg_color");
    int n = -1;
    const int st = ps_frame_graph_launch(g_, &cam, bg.data(), ptr<uint8_t>(out), sync, &n,
                                         cur_stream());
    if (st == PS_ERR_OVERFLOW)
      throw std::runtime_error(std::string("FrameGraph.launch: Too many k-v pairs! (") +
                               ps_last_error() + ")");
    check(st, "FrameGraph.launch");
    return n;
  }

 private:
  ps_frame_graph* g_ = nullptr;
  int width_, height_;
  std::vector<Tensor> keep_;
};

// (pair count, whether the pairs overflowed the key buffers: nothing was
// rendered then and the frame has to be run again with more room).
py::tuple forward_wait(int ticket) {
  int n = 0, st = 0;
  {
    py::gil_scoped_release nogil;
    st = ps_forward_wait(ticket, &n);
  }
  if (st != PS_ERR_OVERFLOW) check(st, "forward_wait");
  return py::make_tuple(n, st == PS_ERR_OVERFLOW);
}

// Batch of frames through shared launches.  Each entry of `frames` is a
// tuple (orig_points, shs, opacities, cov3Ds, width, height, position,
// rotation, focal_x, focal_y, zFar, zNear, bg_color, out_color, records,
// points_xy, rgb_depth, conic_opacity); the feature tensors of the unused
// precision may be empty.  Returns per-frame pair counts when sync is set,
// else an empty list.
std::vector<int> forward_batch(const py::list& frames, int feat_prec, Tensor& curr_offset,
                               Tensor& sort_space, Tensor& ku, Tensor& vu, Tensor& ks, Tensor& vs,
                               Tensor& tile_ranges, Tensor& workspace, bool sync) {
  const int nf = (int)frames.size();
  need(nf > 0 && nf <= PS_MAX_BATCH_FRAMES, "forward_batch: bad frame count");
  std::vector<ps_frame> fr(nf);
  std::vector<int> P(nf);
  int sh_prec = -1;
  int64_t tiles = 0;
  for (int f = 0; f < nf; ++f) {
    py::tuple t = frames[f].cast<py::tuple>();
    need(t.size() == 18, "forward_batch: each frame is an 18-tuple");
    Tensor pts = t[0].cast<Tensor>(), shs = t[1].cast<Tensor>(), op = t[2].cast<Tensor>(),
           cov = t[3].cast<Tensor>();
    check_scene(pts, shs, op, cov);
    const int w = t[4].cast<int>(), h = t[5].cast<int>();
    Tensor pos = t[6].cast<Tensor>(), rot = t[7].cast<Tensor>();
    Tensor bg = t[12].cast<Tensor>(), out = t[13].cast<Tensor>();
    check_out(out, w, h);
    const int shp = sh_prec_of(shs);
    need(sh_prec < 0 || shp == sh_prec, "forward_batch: frames must share the SH precision");
    sh_prec = shp;
    ps_frame& q = fr[f];
    memset(&q, 0, sizeof(q));
    q.P = (int)op.size(0);
    P[f] = q.P;
    q.positions = ptr<float>(pts);
    q.shs = shs.data_ptr();
    q.opacities = ptr<float>(op);
    q.cov3d = ptr<float>(cov);
    q.cam = make_camera(w, h, 16, 16, pos, rot, t[8].cast<float>(), t[9].cast<float>(),
                        t[10].cast<float>(), t[11].cast<float>());
    auto b = host_floats(bg, 3, "bg_color");
    q.bg[0] = b[0];
    q.bg[1] = b[1];
    q.bg[2] = b[2];
    q.out_color = ptr<uint8_t>(out);
    if (feat_prec == PS_PREC_F32) {
      Tensor xy = t[15].cast<Tensor>(), rgbd = t[16].cast<Tensor>(), con = t[17].cast<Tensor>();
      check_feats(xy, rgbd, con);
      need(xy.numel() >= 2LL * q.P, "points_xy too small");
      q.points_xy = ptr<float>(xy);
      q.rgb_depth = ptr<float>(rgbd);
      q.conic_opacity = ptr<float>(con);
    } else {
      Tensor rec = t[14].cast<Tensor>();
      need_dev(rec, "records", at::kByte);
      need(rec.numel() >= 24LL * q.P, "records too small");
      q.records = rec.data_ptr();
    }
    tiles += tiles_of(w, h, 16, 16);
  }
  need_dev(curr_offset, "curr_offset", at::kInt);
  need_dev(sort_space, "list_sorting_space", at::kChar);
  need_dev(ku, "keys_unsorted", at::kLong);
  need_dev(vu, "values_unsorted", at::kInt);
  need_dev(ks, "keys_sorted", at::kLong);
  need_dev(vs, "values_sorted", at::kInt);
  need_dev(tile_ranges, "tile_ranges", at::kInt);
  need(tile_ranges.numel() >= 2 * tiles, "tile_ranges too small for the batch");
  ps_frame_buffers fb;
  memset(&fb, 0, sizeof(fb));
  fb.curr_offset = ptr<int>(curr_offset);
  fb.sort_space = sort_space.data_ptr();
  fb.sort_space_bytes = (size_t)sort_space.numel();
  fb.keys_unsorted = ptr<uint64_t>(ku);
  fb.values_unsorted = ptr<uint32_t>(vu);
  fb.keys_sorted = ptr<uint64_t>(ks);
  fb.values_sorted = ptr<uint32_t>(vs);
  int cap = (int)std::min({ku.numel(), vu.numel(), ks.numel(), vs.numel()});
  while (cap > 0 && ps_get_sort_buffer_size(cap) > fb.sort_space_bytes) cap /= 2;
  fb.capacity = cap;
  fb.tile_ranges = ptr<int>(tile_ranges);
  const size_t wsb = ps_forward_batch_workspace_size(nf, P.data());
  Tensor ws = workspace;
  if (!(ws.defined() && ws.numel() > 0)) ws = cached_workspace(ku.device(), wsb, "batch");
  need_dev(ws, "workspace", at::kByte);
  need((size_t)ws.numel() >= wsb, "workspace too small");
  fb.workspace = ws.data_ptr();
  fb.workspace_bytes = (size_t)ws.numel();
  std::vector<int> counts(nf, 0);
  const int st = ps_forward_batch(nf, fr.data(), sh_prec, feat_prec, &fb, sync ? 1 : 0,
                                  counts.data(), cur_stream());
  if (st == PS_ERR_OVERFLOW)
    throw std::runtime_error(std::string("forward_batch: Too many k-v pairs! (") +
                             ps_last_error() + ")");
  check(st, "forward_batch");
  if (!sync) counts.clear();
  return counts;
}

py::tuple camera_matrices(int width, int height, Tensor& position, Tensor& rotation, float fx,
                          float fy, float zfar, float znear) {
  ps_camera cam = make_camera(width, height, 16, 16, position, rotation, fx, fy, zfar, znear);
  Tensor v = at::empty({16}, at::kFloat), p = at::empty({16}, at::kFloat),
         t = at::empty({2}, at::kFloat);
  check(ps_camera_matrices(&cam, v.data_ptr<float>(), p.data_ptr<float>(), t.data_ptr<float>()),
        "camera_matrices");
  return py::make_tuple(v, p, t);
}

}  // namespace

PYBIND11_MODULE(_C, m) {
  m.doc() = "MI355X polysplat forward rasterizer (HIP, gfx950)";
  m.def("version", []() { return std::string(ps_version()); });
  auto ops = m.def_submodule("ops", "polysplat.ops-compatible operators");
  ops.def("loadPly", &ps_ply::load_ply);
  ops.def("preprocess", &preprocess_impl);
  ops.def("preprocess_half_sh", &preprocess_impl);
  ops.def("preprocess_records", &preprocess_records);
  ops.def("preprocess_es", &preprocess_es);
  ops.def("es_tile_sort", &es_tile_sort, py::arg("num_rendered"), py::arg("width"),
          py::arg("height"), py::arg("block_x"), py::arg("block_y"),
          py::arg("tile_keys_unsorted"), py::arg("gauss_values_unsorted"),
          py::arg("tile_keys_sorted"), py::arg("gauss_values_sorted"),
          py::arg("gaussian_keys_sorted_out"), py::arg("cub_sort_scratch"),
          py::arg("tile_ranges") = py::none());
  ops.def("get_sort_buffer_size", &get_sort_buffer_size);
  ops.def("sort_gaussian", &sort_gaussian);
  ops.def("gather_features", &gather_features);
  ops.def("gather_records", &gather_records);
  ops.def("precompute_tile_ranges", &precompute_tile_ranges);
  ops.def("precompute_tile_ranges_scan", &precompute_tile_ranges_scan);
  ops.def("render_16x16", &render_16x16);
  ops.def("render_16x16_preranges_smem_persistent_lite", &render_lite);
  ops.def("render_16x16_preranges_smem_persistent_lite_dt", &render_lite_dt);
  ops.def("render_16x16_preranges_smem_persistent_lite_fused", &render_lite_fused);
  ops.def("render_records", &render_records, py::arg("num_rendered"), py::arg("width"),
          py::arg("height"), py::arg("records"), py::arg("feat_prec"), py::arg("gathered"),
          py::arg("values_sorted"), py::arg("tile_ranges"), py::arg("bg_color"),
          py::arg("out_color"), py::arg("dt") = false);
  ops.def("get_tile_order_sort_temp_size", &get_tile_order_sort_temp_size);
  ops.def("compute_tile_order", &compute_tile_order);
  ops.def("compute_tile_order_packed", &compute_tile_order_packed);
  ops.def("compute_tile_order_packed_morton", &compute_tile_order_packed_morton);
  ops.def("render_16x16_reordered", &render_16x16_reordered);
  ops.def("render_16x16_reordered_v2", &render_16x16_reordered_v2);
  ops.def("render_16x16_topk_smem", &render_topk_smem);
  ops.def("render_16x16_topk_smem_persistent", &render_topk_smem_persistent);
  ops.def("render_16x16_topk_smem_persistent_v2", &render_topk_smem_persistent);
  ops.def("render_16x16_reordered_persistent", &render_16x16_reordered_persistent);
  ops.def("render_records_reordered", &render_records_reordered, py::arg("num_rendered"),
          py::arg("width"), py::arg("height"), py::arg("records"), py::arg("feat_prec"),
          py::arg("values_sorted"), py::arg("tile_desc"), py::arg("bg_color"),
          py::arg("out_color"), py::arg("dt") = false);
  ops.def("forward_fused", &forward_fused);
  ops.def("forward_fused_e2e", &forward_fused_e2e);
  ops.def("forward_fused_e2e_preranges", &forward_fused_e2e_preranges);
  ops.def("forward_frame", &forward_frame);
  ops.def("forward_frame_post", &forward_frame_post);
  ops.def("forward_wait", &forward_wait);
  py::class_<FrameGraph>(ops, "FrameGraph")
      .def(py::init<Tensor&, Tensor&, Tensor&, Tensor&, int, int, Tensor&, Tensor&, float, float,
                    float, float, int, int, Tensor&, Tensor&, Tensor&, Tensor&, Tensor&, Tensor&,
                    Tensor&, Tensor&, Tensor&, Tensor&, Tensor&, Tensor&>())
      .def("launch", &FrameGraph::launch, py::arg("position"), py::arg("rotation"),
           py::arg("focal_x"), py::arg("focal_y"), py::arg("zFar"), py::arg("zNear"),
           py::arg("bg_color"), py::arg("out_color"), py::arg("sync") = 1);
  ops.def("forward_query", [](int ticket) {
    const int r = ps_forward_query(ticket);
    if (r < 0) check(-r, "forward_query");
    return r == 1;
  });
  ops.def("camera_matrices", &camera_matrices);
  ops.def("forward_batch", &forward_batch);
  ops.def("forward_batch_workspace_size", [](std::vector<int> P) {
    return ps_forward_batch_workspace_size((int)P.size(), P.data());
  });
  ops.def("max_batch_frames", []() { return (int)PS_MAX_BATCH_FRAMES; });
  ops.def("forward_batch_counts_offset", [](std::vector<int> P) {
    // Byte offset of the per-frame counts inside a batch workspace.
    char* base = reinterpret_cast<char*>(4096);
    const int* p = ps_forward_batch_counts_ptr(base, (int)P.size(), P.data());
    return (int64_t)(reinterpret_cast<const char*>(p) - base);
  });
  ops.def("forward_workspace_size",
          [](int64_t P) { return ps_forward_workspace_size((int)P); });
  ops.def("preprocess_workspace_size",
          [](int64_t P) { return ps_preprocess_workspace_size((int)P); });
}
