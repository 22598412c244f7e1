// Tile-range build over the sorted keys, and the tile-sorted feature gather.
//
// Ranges: computeTileRanges / computeTileRangesScan (render.cu:3509-3557).
// Gather: gatherGaussianFeatures (gather.cu:6-22).
// Every kernel takes the pair count from device memory or from the host, and
// uses a fixed grid with a stride loop so it can run without a host round
// trip for the count.

#include "ps_common.h"
#include "ps_host.h"

namespace ps {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxGrid = 4096;

// Pair count from device memory (capped at the buffer size) or the host.
__device__ __forceinline__ int count_of(const int* d_n, int n_host, int max_pairs) {
  int n = d_n ? *d_n : n_host;
  if (d_n && n > max_pairs) n = max_pairs;
  return n < 0 ? 0 : n;
}

__device__ __forceinline__ uint32_t tile_of(const uint64_t* keys, int i) {
  return (uint32_t)(keys[i] >> 32);
}

// [lower_bound, upper_bound) of tile id `t` in the sorted high words.
__global__ __launch_bounds__(kThreads) void ps_ranges_bsearch(const uint64_t* __restrict__ keys,
                                                              const int* __restrict__ d_n,
                                                              int n_host, int max_pairs, int tiles,
                                                              int2* __restrict__ ranges) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= tiles) return;
  const int n = count_of(d_n, n_host, max_pairs);
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (tile_of(keys, mid) < (uint32_t)t) lo = mid + 1; else hi = mid;
  }
  const int start = lo;
  hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (tile_of(keys, mid) <= (uint32_t)t) lo = mid + 1; else hi = mid;
  }
  ranges[t] = make_int2(start, lo);
}

// Zeroes `count` ranges and, when given, `words` (at most kThreads) counter
// words.
__global__ __launch_bounds__(kThreads) void ps_fill_int2(int2* __restrict__ p, int count,
                                                         int* __restrict__ counter, int words) {
  if (counter != nullptr && blockIdx.x == 0 && threadIdx.x < words) counter[threadIdx.x] = 0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < count; i += gridDim.x * kThreads)
    p[i] = make_int2(0, 0);
}

// Boundary scan: writes starts/ends only where the tile id changes, so empty
// tiles keep the (0, 0) written by ps_fill_int2.  The tile id is the key's
// high word masked with `mask` (all ones for render ranges; the radix sort's
// segment pass keeps only the bits it sorted).
__global__ __launch_bounds__(kThreads) void ps_ranges_scan(const uint64_t* __restrict__ keys,
                                                           const int* __restrict__ d_n, int n_host,
                                                           int max_pairs, int tiles, uint32_t mask,
                                                           int2* __restrict__ ranges) {
  const int n = count_of(d_n, n_host, max_pairs);
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const uint32_t cur = tile_of(keys, i) & mask;
    const bool cur_ok = cur < (uint32_t)tiles;  // keys from a foreign caller may be garbage
    if (i == 0) {
      if (cur_ok) ranges[cur].x = 0;
    } else {
      const uint32_t prev = tile_of(keys, i - 1) & mask;
      if (cur != prev) {
        if (prev < (uint32_t)tiles) ranges[prev].y = i;
        if (cur_ok) ranges[cur].x = i;
      }
    }
    if (i == n - 1 && cur_ok) ranges[cur].y = n;
  }
}

// The same boundary scan two keys at a time from one 16-byte load (keys
// 16-byte aligned): key i's tile id is the high word of its 64-bit key
// (words 1 and 3 of the load), the key before the pair is read on its own
// (mostly from cache: the neighbouring lane loaded it).
__device__ __forceinline__ void mark_boundary(int i, int n, uint32_t prev, uint32_t cur, int tiles,
                                              int2* __restrict__ ranges) {
  const bool cur_ok = cur < (uint32_t)tiles;  // keys from a foreign caller may be garbage
  if (i == 0) {
    if (cur_ok) ranges[cur].x = 0;
  } else if (cur != prev) {
    if (prev < (uint32_t)tiles) ranges[prev].y = i;
    if (cur_ok) ranges[cur].x = i;
  }
  if (i == n - 1 && cur_ok) ranges[cur].y = n;
}

__global__ __launch_bounds__(kThreads) void ps_ranges_scan2(const uint64_t* __restrict__ keys,
                                                            const int* __restrict__ d_n,
                                                            int n_host, int max_pairs, int tiles,
                                                            uint32_t mask,
                                                            int2* __restrict__ ranges) {
  const int n = count_of(d_n, n_host, max_pairs);
  const int pairs = (n + 1) >> 1;
  const uint4* k4 = reinterpret_cast<const uint4*>(keys);
  for (int j = blockIdx.x * kThreads + threadIdx.x; j < pairs; j += gridDim.x * kThreads) {
    const int i = 2 * j;
    uint32_t c0, c1 = 0u;
    if (i + 1 < n) {
      const uint4 v = k4[j];
      c0 = v.y & mask;
      c1 = v.w & mask;
    } else {
      c0 = tile_of(keys, i) & mask;
    }
    const uint32_t prev = i > 0 ? (tile_of(keys, i - 1) & mask) : 0u;
    mark_boundary(i, n, prev, c0, tiles, ranges);
    if (i + 1 < n) mark_boundary(i + 1, n, c0, c1, tiles, ranges);
  }
}

// Sorted 32-bit tile ids to the 64-bit keys the renderers read (tile id in
// the high word, low word zero, as the reference's depth-first pipeline
// rebuilds them: sort.cu:205-215) and, when `ranges` is given, the tile
// ranges by the same boundary scan as ps_ranges_scan (tile ids need no
// masking).
__global__ __launch_bounds__(kThreads) void ps_tile_ids_finish(const uint32_t* __restrict__ tk,
                                                               const int* __restrict__ d_n,
                                                               int n_host, int max_pairs,
                                                               int tiles,
                                                               uint64_t* __restrict__ keys,
                                                               int2* __restrict__ ranges) {
  const int n = count_of(d_n, n_host, max_pairs);
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const uint32_t cur = tk[i];
    keys[i] = (uint64_t)cur << 32;
    if (ranges == nullptr) continue;
    const bool cur_ok = cur < (uint32_t)tiles;
    if (i == 0) {
      if (cur_ok) ranges[cur].x = 0;
    } else {
      const uint32_t prev = tk[i - 1];
      if (cur != prev) {
        if (prev < (uint32_t)tiles) ranges[prev].y = i;
        if (cur_ok) ranges[cur].x = i;
      }
    }
    if (i == n - 1 && cur_ok) ranges[cur].y = n;
  }
}

__global__ __launch_bounds__(kThreads) void ps_gather_f32(
    const int* __restrict__ d_n, int n_host, int max_pairs, const uint32_t* __restrict__ order,
    const float2* __restrict__ sxy, const float4* __restrict__ srgbd,
    const float4* __restrict__ scon, float2* __restrict__ dxy, float4* __restrict__ drgbd,
    float4* __restrict__ dcon) {
  const int n = count_of(d_n, n_host, max_pairs);
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const uint32_t g = order[i];
    dxy[i] = sxy[g];
    drgbd[i] = srgbd[g];
    dcon[i] = scon[g];
  }
}

__global__ __launch_bounds__(kThreads) void ps_gather_rec(const int* __restrict__ d_n, int n_host,
                                                          int max_pairs,
                                                          const uint32_t* __restrict__ order,
                                                          const FeatRec16* __restrict__ src,
                                                          FeatRec16* __restrict__ dst) {
  const int n = count_of(d_n, n_host, max_pairs);
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    dst[i] = src[order[i]];
}

inline int grid_for(int n) {
  int g = (n + kThreads - 1) / kThreads;
  if (g > kMaxGrid) g = kMaxGrid;
  return g < 1 ? 1 : g;
}

}  // namespace

int tile_ranges_scan_launch(int n_host, const int* d_n, int max_pairs, int tiles,
                            const uint64_t* keys, int* ranges, hipStream_t stream, uint32_t mask,
                            int* zero_counter, int zero_words, bool cleared) {
  PS_REQUIRE(zero_words >= 0 && zero_words <= kThreads, "bad counter word count");
  if (!cleared) {
    hipLaunchKernelGGL(ps_fill_int2, dim3(grid_for(tiles)), dim3(kThreads), 0, stream,
                       reinterpret_cast<int2*>(ranges), tiles, zero_counter, zero_words);
    PS_LAUNCH_CHECK();
  }
  const int n = d_n ? max_pairs : n_host;
  if (n > 0) {
    if ((reinterpret_cast<uintptr_t>(keys) & 15u) == 0u)
      hipLaunchKernelGGL(ps_ranges_scan2, dim3(grid_for((n + 1) / 2)), dim3(kThreads), 0, stream,
                         keys, d_n, n_host, max_pairs, tiles, mask,
                         reinterpret_cast<int2*>(ranges));
    else
      hipLaunchKernelGGL(ps_ranges_scan, dim3(grid_for(n)), dim3(kThreads), 0, stream, keys,
                         d_n, n_host, max_pairs, tiles, mask, reinterpret_cast<int2*>(ranges));
    PS_LAUNCH_CHECK();
  }
  return PS_OK;
}

int tile_ids_finish_launch(int n_host, const int* d_n, int max_pairs, int tiles,
                           const uint32_t* tkeys, uint64_t* keys, int* ranges,
                           hipStream_t stream) {
  if (ranges != nullptr) {
    hipLaunchKernelGGL(ps_fill_int2, dim3(grid_for(tiles)), dim3(kThreads), 0, stream,
                       reinterpret_cast<int2*>(ranges), tiles, (int*)nullptr, 0);
    PS_LAUNCH_CHECK();
  }
  const int n = d_n ? max_pairs : n_host;
  if (n > 0) {
    hipLaunchKernelGGL(ps_tile_ids_finish, dim3(grid_for(n)), dim3(kThreads), 0, stream, tkeys,
                       d_n, n_host, max_pairs, tiles, keys, reinterpret_cast<int2*>(ranges));
    PS_LAUNCH_CHECK();
  }
  return PS_OK;
}

int tile_ranges_bsearch_launch(int n_host, const int* d_n, int max_pairs, int tiles,
                               const uint64_t* keys, int* ranges, hipStream_t stream) {
  if (tiles <= 0) return PS_OK;
  hipLaunchKernelGGL(ps_ranges_bsearch, dim3((tiles + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     stream, keys, d_n, n_host, max_pairs, tiles, reinterpret_cast<int2*>(ranges));
  PS_LAUNCH_CHECK();
  return PS_OK;
}

}  // namespace ps

extern "C" int ps_precompute_tile_ranges(int num_rendered, const int* d_num_rendered, int width,
                                         int height, int block_x, int block_y,
                                         const uint64_t* keys_sorted, int* tile_ranges,
                                         ps_stream_t stream) {
  PS_REQUIRE(width > 0 && height > 0 && block_x > 0 && block_y > 0, "bad image/tile size");
  PS_REQUIRE(tile_ranges != nullptr, "tile_ranges is NULL");
  PS_REQUIRE(d_num_rendered != nullptr || num_rendered >= 0, "negative num_rendered");
  const int tiles = ((width + block_x - 1) / block_x) * ((height + block_y - 1) / block_y);
  return ps::tile_ranges_bsearch_launch(num_rendered, d_num_rendered,
                                        d_num_rendered ? 0x7fffffff : num_rendered, tiles,
                                        keys_sorted, tile_ranges,
                                        reinterpret_cast<hipStream_t>(stream));
}

extern "C" int ps_precompute_tile_ranges_scan(int num_rendered, const int* d_num_rendered,
                                              int max_pairs, int width, int height, int block_x,
                                              int block_y, const uint64_t* keys_sorted,
                                              int* tile_ranges, ps_stream_t stream) {
  PS_REQUIRE(width > 0 && height > 0 && block_x > 0 && block_y > 0, "bad image/tile size");
  PS_REQUIRE(tile_ranges != nullptr, "tile_ranges is NULL");
  PS_REQUIRE(d_num_rendered != nullptr || num_rendered >= 0, "negative num_rendered");
  const int tiles = ((width + block_x - 1) / block_x) * ((height + block_y - 1) / block_y);
  return ps::tile_ranges_scan_launch(num_rendered, d_num_rendered, max_pairs, tiles, keys_sorted,
                                     tile_ranges, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int ps_gather_features(int num_rendered, const int* d_num_rendered, int max_pairs,
                                  const uint32_t* values_sorted, const float* src_xy,
                                  const float* src_rgb_depth, const float* src_conic,
                                  float* dst_xy, float* dst_rgb_depth, float* dst_conic,
                                  ps_stream_t stream) {
  PS_REQUIRE(d_num_rendered != nullptr || num_rendered >= 0, "negative num_rendered");
  const int n = d_num_rendered ? max_pairs : num_rendered;
  if (n <= 0) return PS_OK;
  hipLaunchKernelGGL(ps::ps_gather_f32, dim3(ps::grid_for(n)), dim3(ps::kThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), d_num_rendered, num_rendered, n,
                     values_sorted, (const float2*)src_xy, (const float4*)src_rgb_depth,
                     (const float4*)src_conic, (float2*)dst_xy, (float4*)dst_rgb_depth,
                     (float4*)dst_conic);
  PS_LAUNCH_CHECK();
  return PS_OK;
}

extern "C" int ps_gather_records(int num_rendered, const int* d_num_rendered, int max_pairs,
                                 const uint32_t* values_sorted, const void* src_records,
                                 void* dst_records, ps_stream_t stream) {
  PS_REQUIRE(d_num_rendered != nullptr || num_rendered >= 0, "negative num_rendered");
  const int n = d_num_rendered ? max_pairs : num_rendered;
  if (n <= 0) return PS_OK;
  hipLaunchKernelGGL(ps::ps_gather_rec, dim3(ps::grid_for(n)), dim3(ps::kThreads), 0,
                     reinterpret_cast<hipStream_t>(stream), d_num_rendered, num_rendered, n,
                     values_sorted, (const ps::FeatRec16*)src_records,
                     (ps::FeatRec16*)dst_records);
  PS_LAUNCH_CHECK();
  return PS_OK;
}
