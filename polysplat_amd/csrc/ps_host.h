// Host-side helpers shared by the C ABI translation units: camera matrices,
// status strings and error bookkeeping.
#pragma once

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include "../../include/polysplat_amd.h"
#include "ps_common.h"

namespace ps {

void set_last_error(const char* fmt, ...);

#define PS_HIP_CHECK(expr)                                                        \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      ::ps::set_last_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                           __FILE__, __LINE__);                                   \
      return PS_ERR_HIP;                                                          \
    }                                                                             \
  } while (0)

#define PS_LAUNCH_CHECK()                                                         \
  do {                                                                            \
    hipError_t _e = hipGetLastError();                                            \
    if (_e != hipSuccess) {                                                       \
      ::ps::set_last_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e), \
                           __FILE__, __LINE__);                                   \
      return PS_ERR_HIP;                                                          \
    }                                                                             \
  } while (0)

#define PS_REQUIRE(cond, ...)              \
  do {                                     \
    if (!(cond)) {                         \
      ::ps::set_last_error(__VA_ARGS__);   \
      return PS_ERR_INVALID;               \
    }                                      \
  } while (0)

inline int tiles_x(const ps_camera* c) { return (c->width + c->block_x - 1) / c->block_x; }
inline int tiles_y(const ps_camera* c) { return (c->height + c->block_y - 1) / c->block_y; }

// Builds the view and projection matrices the way the reference does
// (preprocess.cu:643-671, pybind.cpp:140-143): the 3x3 rotation given as
// rows r0..r8 becomes the column-major matrix whose columns are (r0,r1,r2),
// (r3,r4,r5), (r6,r7,r8); view = [R | R*(-c)]; proj = transpose(P) * view.
// All products are expanded in the same column-major order.
inline void camera_matrices(const ps_camera* cam, float view[16], float proj[16],
                            float* tan_fovx, float* tan_fovy) {
  const float* r = cam->rotation;
  const float* c = cam->position;
  float R[3][3] = {{r[0], r[1], r[2]}, {r[3], r[4], r[5]}, {r[6], r[7], r[8]}};  // R[col][row]
  float nc[3] = {-c[0], -c[1], -c[2]};
  float t[3];
  for (int k = 0; k < 3; ++k) t[k] = R[0][k] * nc[0] + R[1][k] * nc[1] + R[2][k] * nc[2];
  for (int col = 0; col < 3; ++col) {
    for (int row = 0; row < 3; ++row) view[col * 4 + row] = R[col][row];
    view[col * 4 + 3] = 0.0f;
  }
  view[12] = t[0];
  view[13] = t[1];
  view[14] = t[2];
  view[15] = 1.0f;

  const float zn = cam->z_near, zf = cam->z_far;
  const float top = (float)cam->height / (2.0f * cam->focal_y) * zn;
  const float bottom = -top;
  const float right = (float)cam->width / (2.0f * cam->focal_x) * zn;
  const float left = -right;
  float P[4][4];  // P[col][row]
  memset(P, 0, sizeof(P));
  P[0][0] = 2.0f * zn / (right - left);
  P[1][1] = 2.0f * zn / (top - bottom);
  P[0][2] = (right + left) / (right - left);
  P[1][2] = (top + bottom) / (top - bottom);
  P[3][2] = 1.0f;
  P[2][2] = 1.0f * zf / (zf - zn);
  P[2][3] = -(zf * zn) / (zf - zn);
  float A[4][4];  // transpose(P)
  for (int col = 0; col < 4; ++col)
    for (int row = 0; row < 4; ++row) A[col][row] = P[row][col];
  // proj[j] = A[0]*V[j][0] + A[1]*V[j][1] + A[2]*V[j][2] + A[3]*V[j][3]
  for (int j = 0; j < 4; ++j)
    for (int row = 0; row < 4; ++row)
      proj[j * 4 + row] = A[0][row] * view[j * 4 + 0] + A[1][row] * view[j * 4 + 1] +
                          A[2][row] * view[j * 4 + 2] + A[3][row] * view[j * 4 + 3];
  *tan_fovx = (float)cam->width / (2.0f * cam->focal_x);
  *tan_fovy = (float)cam->height / (2.0f * cam->focal_y);
}

inline CamParams make_cam_params(const ps_camera* cam) {
  CamParams cp;
  camera_matrices(cam, cp.view, cp.proj, &cp.tan_fovx, &cp.tan_fovy);
  cp.campos[0] = cam->position[0];
  cp.campos[1] = cam->position[1];
  cp.campos[2] = cam->position[2];
  cp.focal_x = cam->focal_x;
  cp.focal_y = cam->focal_y;
  cp.W = cam->width;
  cp.H = cam->height;
  cp.block_x = cam->block_x;
  cp.block_y = cam->block_y;
  cp.grid_x = tiles_x(cam);
  cp.grid_y = tiles_y(cam);
  return cp;
}

// Number of significant bits of n (sort.cu:14-29 computes the same value).
inline int bit_length(unsigned n) {
  int b = 0;
  while (n) {
    ++b;
    n >>= 1;
  }
  return b;
}

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Launchers shared between translation units.
// Plain renderer (render.hip): one workgroup per tile, ranges from
// tile_ranges or, when that is NULL, a search of keys; with a tile order,
// workgroup p renders tile tile_order[p], with packed descriptors the tile
// and range of descriptor p.
int render_plain_launch(int n_host, const int* d_n, int W, int H, const float* xy,
                        const float* rgbd, const float* con, const uint64_t* keys,
                        const uint32_t* order, const float bg[3], uint8_t* out,
                        hipStream_t stream, const int* ranges = nullptr,
                        const uint32_t* tile_order = nullptr, const int* tile_desc = nullptr,
                        const FrameConsts* fc = nullptr, const TopKCache& tk = TopKCache());
// Tile order (tile_order.hip): tiles by pair count, descending, stable;
// bucket > 0 selects the Morton-bucket keys; desc (int4 per position,
// zig-zag regrouped when zigzag > 1) is optional.
int tile_order_launch(const int* tile_ranges, int tiles, int gx, int zigzag, int bucket,
                      uint32_t* order, uint32_t* keys_out, uint32_t* ids_out, int* desc,
                      void* temp, size_t temp_bytes, hipStream_t stream);
size_t tile_order_temp_size(int tiles);
// Persistent renderer (render.hip).  tile_order: NULL for row-major tiles,
// else a permutation of the frame's tiles, rendered in that order;
// tile_desc: NULL, or int4 (column, row, start, end) per order position,
// which then replaces tile_ranges and tile_order; fc: NULL, or per-frame
// constants in device memory whose background and image pointer the kernel
// reads at run time instead of bg and out (frames replayed from a graph);
// tk: the top-k feature cache that packed values refer to (FP32 features).
int render_persistent_launch(int n_host, const int* d_n, int W, int H, int feat_prec,
                             const float* xy, const float* rgbd, const float* con, const void* rec,
                             int gathered, const uint64_t* keys, const uint32_t* order,
                             const int* ranges, int dt, const float bg[3], uint8_t* out,
                             unsigned* queue, hipStream_t stream,
                             const uint32_t* tile_order = nullptr,
                             const int* tile_desc = nullptr, const FrameConsts* fc = nullptr,
                             const TopKCache& tk = TopKCache());
// `cleared`: the ranges (and the counter words) were already zeroed on the
// stream since their last use, so no clearing launch is needed.
int tile_ranges_scan_launch(int n_host, const int* d_n, int max_pairs, int tiles,
                            const uint64_t* keys, int* ranges, hipStream_t stream,
                            uint32_t mask = 0xffffffffu, int* zero_counter = nullptr,
                            int zero_words = 1, bool cleared = false);
// Depth-first pipeline: the depth sort of the Gaussians and the tile sort of
// their (tile id, Gaussian) pairs (sort.hip).
int depth_sort_launch(const DepthSortArgs& A, int max_pairs, void* space, size_t space_bytes,
                      hipStream_t stream);
// Stable sort of the depth-first pipeline's (tile id, value) pairs by tile
// id, then the 64-bit keys (tile id << 32) and, when `tile_ranges` is given,
// the tile ranges.  The pair count is n_host, or *d_n when d_n is given.
int tile_sort_launch(int n_host, const int* d_n, int max_pairs, int tiles, void* space,
                     size_t space_bytes, const uint32_t* tkeys_in, const uint32_t* vals_in,
                     uint32_t* tkeys_out, uint32_t* vals_out, uint64_t* keys_out,
                     int* tile_ranges, hipStream_t stream);
// keys[i] = tkeys[i] << 32 and, when `ranges` is given, the boundary-scan
// tile ranges of the sorted tile ids (empty tiles (0, 0)).
int tile_ids_finish_launch(int n_host, const int* d_n, int max_pairs, int tiles,
                           const uint32_t* tkeys, uint64_t* keys, int* ranges,
                           hipStream_t stream);
// Stable radix sort of (key, value) pairs over key bits [0, end_bit).  When
// `tile_ranges` is given, the keys' high words are tile ids below `tiles`
// and the sort also leaves each tile's [start, end) there.
int sort_launch(int n_host, const int* d_n, int max_pairs, int end_bit, void* space,
                size_t space_bytes, const uint64_t* keys_in, const uint32_t* vals_in,
                uint64_t* keys_out, uint32_t* vals_out, hipStream_t stream,
                int* tile_ranges = nullptr, int tiles = 0);

}  // namespace ps
