// Tile orders for the reordered renderers (render.cu:4411-4520, :4790-4840,
// :6959-7100): the frame's tiles sorted by their pair count, heaviest first,
// ties in tile-id order (a stable descending sort, as the reference's
// cub::DeviceRadixSort::SortPairsDescending gives), optionally as packed
// descriptors (column, row, range start, range end) in that order, with the
// reference's zig-zag regrouping, or with its Morton-bucket keys.
//
// The sort is the depth sort's machinery (sort.hip): the keys are the
// complements of the sort keys (so ascending = the reference's descending),
// 8-bit stable digit passes of which only the digits that differ among the
// frame's tiles run, planned on the device from the keys' OR and AND.  Pair
// counts below 65536 differ in their low two bytes only, so a frame's order
// costs two digit passes over its tiles.  Launches: one key kernel, the
// four digit steps (those of constant digits return at once), one finish
// kernel; nothing is read back to the host.

#include "ps_common.h"
#include "ps_host.h"
#include "ps_wave.h"

namespace ps {
namespace {

constexpr int kOrderThreads = 256;

// Bits of v (below 2^16) moved to the even bit positions.
__device__ __forceinline__ uint32_t spread_even(uint32_t v) {
  uint32_t r = 0;
#pragma unroll
  for (int b = 0; b < 16; ++b) r |= ((v >> b) & 1u) << (2 * b);
  return r;
}

// Z-order code of tile (x, y): x in the even bits, y in the odd bits
// (render.cu:4430-4441).
__device__ __forceinline__ uint32_t z_order(uint32_t x, uint32_t y) {
  return spread_even(x & 0xffffu) | (spread_even(y & 0xffffu) << 1);
}

// Sort key of tile t, descending: its pair count, or, with a Morton bucket
// size, (min(count / bucket, 0xffff) << 16) | (0xffff - (z_order & 0xffff))
// (render.cu:4496-4518).
__device__ __forceinline__ uint32_t order_key(int2 r, int t, int gx, int bucket) {
  const uint32_t count = (uint32_t)(r.y - r.x);
  if (bucket <= 0) return count;
  uint32_t b = count / (uint32_t)bucket;
  if (b > 0xffffu) b = 0xffffu;
  const uint32_t z = z_order((uint32_t)(t % gx), (uint32_t)(t / gx)) & 0xffffu;
  return (b << 16) | (0xffffu - z);
}

// Complemented keys and, per workgroup, their OR and AND in the plan slot of
// that workgroup (exactly kPlanSlots workgroups, so every slot is written
// once and nothing needs clearing first).
__global__ __launch_bounds__(kOrderThreads) void ps_tile_order_keys(
    const int2* __restrict__ ranges, int tiles, int gx, int bucket, uint32_t* __restrict__ words,
    uint32_t* __restrict__ plan) {
  __shared__ uint32_t s_or[kOrderThreads / kWave], s_and[kOrderThreads / kWave];
  uint32_t o = 0u, a = ~0u;
  for (int t = blockIdx.x * kOrderThreads + threadIdx.x; t < tiles;
       t += kPlanSlots * kOrderThreads) {
    const uint32_t w = ~order_key(ranges[t], t, gx, bucket);
    words[t] = w;
    o |= w;
    a &= w;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    o |= (uint32_t)__shfl_xor((int)o, d);
    a &= (uint32_t)__shfl_xor((int)a, d);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    s_or[wave] = o;
    s_and[wave] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kOrderThreads / kWave; ++w) {
      o |= s_or[w];
      a &= s_and[w];
    }
    plan[blockIdx.x * kPlanStride] = o;
    plan[(kPlanSlots + blockIdx.x) * kPlanStride] = a;
  }
}

// Source position of order position p after the zig-zag regrouping
// (render.cu:4838-4866): groups of g positions, every odd group reversed.
__device__ __forceinline__ int zigzag_src(int p, int tiles, int g) {
  if (g <= 1) return p;
  const int group = p / g, slot = p - group * g, base = group * g;
  int end = base + g;
  if (end > tiles) end = tiles;
  int src = (group & 1) ? base + (end - base - 1 - slot) : base + slot;
  if (src >= tiles) src = tiles - 1;
  return src;
}

// The sorted ids (tile_order), the sorted keys (counts), the input ids
// (0..tiles-1, as the reference leaves its value input) and, when asked,
// the packed descriptors in zig-zag order.
__global__ __launch_bounds__(kOrderThreads) void ps_tile_order_finish(
    const uint32_t* __restrict__ plan, const uint32_t* __restrict__ words, uint32_t* kb0,
    uint32_t* kb1, uint32_t* vb0, uint32_t* vb1, const int2* __restrict__ ranges, int tiles,
    int gx, int zigzag, uint32_t* __restrict__ order, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ ids_out, int4* __restrict__ desc) {
  __shared__ uint32_t s_k;
  if (threadIdx.x < kWave) {
    const int l = threadIdx.x;
    uint32_t o = plan[l * kPlanStride], a = plan[(kPlanSlots + l) * kPlanStride];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      o |= (uint32_t)__shfl_xor((int)o, d);
      a &= (uint32_t)__shfl_xor((int)a, d);
    }
    if (l == 0) s_k = (uint32_t)depth_digits_running(o == 0u ? 0u : (o ^ a));
  }
  __syncthreads();
  const int k = (int)s_k;
  const uint32_t* ids = k == 0 ? nullptr : ((k - 1) & 1 ? vb1 : vb0);
  const uint32_t* ks = k == 0 ? words : ((k - 1) & 1 ? kb1 : kb0);
  for (int p = blockIdx.x * kOrderThreads + threadIdx.x; p < tiles;
       p += gridDim.x * kOrderThreads) {
    const uint32_t id = ids ? ids[p] : (uint32_t)p;
    if (order) order[p] = id;
    if (keys_out) keys_out[p] = ~ks[p];
    if (ids_out) ids_out[p] = (uint32_t)p;
    if (desc) {
      const int src = zigzag_src(p, tiles, zigzag);
      const int t = ids ? (int)ids[src] : src;
      const int2 r = ranges[t];
      desc[p] = make_int4(t % gx, t / gx, r.x, r.y);
    }
  }
}

struct OrderWs {
  uint32_t* words;
  uint32_t* kb[2];
  uint32_t* vb[2];
  uint32_t* plan;
  char* sort;
  size_t sort_bytes;
};

inline size_t order_layout(int tiles, char* base, OrderWs* w) {
  const size_t n = (size_t)(tiles > 0 ? tiles : 1);
  const size_t arr = align_up(n * sizeof(uint32_t), 256);
  size_t off = 0;
  if (w) w->words = reinterpret_cast<uint32_t*>(base + off);
  off += arr;
  for (int i = 0; i < 2; ++i) {
    if (w) w->kb[i] = reinterpret_cast<uint32_t*>(base + off);
    off += arr;
    if (w) w->vb[i] = reinterpret_cast<uint32_t*>(base + off);
    off += arr;
  }
  if (w) w->plan = reinterpret_cast<uint32_t*>(base + off);
  off += align_up(2 * kPlanSlots * kPlanStride * sizeof(uint32_t), 256);
  const size_t sb = ps_get_sort_buffer_size(tiles > 0 ? tiles : 1);
  if (w) {
    w->sort = base + off;
    w->sort_bytes = sb;
  }
  off += sb;
  return off;
}

}  // namespace

int tile_order_launch(const int* tile_ranges, int tiles, int gx, int zigzag, int bucket,
                      uint32_t* order, uint32_t* keys_out, uint32_t* ids_out, int* desc,
                      void* temp, size_t temp_bytes, hipStream_t stream) {
  PS_REQUIRE(tiles >= 0, "negative tile count");
  PS_REQUIRE(gx > 0, "x_blocks must be positive");
  if (tiles == 0) return PS_OK;
  PS_REQUIRE(tile_ranges != nullptr, "tile_ranges is NULL");
  PS_REQUIRE(order != nullptr || desc != nullptr, "no output for the tile order");
  PS_REQUIRE(temp != nullptr && temp_bytes >= order_layout(tiles, nullptr, nullptr),
             "tile-order temp too small (%zu < %zu)", temp_bytes,
             order_layout(tiles, nullptr, nullptr));
  PS_REQUIRE(zigzag >= 0 && bucket >= 0, "negative zig-zag group or bucket size");
  OrderWs w;
  order_layout(tiles, static_cast<char*>(temp), &w);
  const int2* rg = reinterpret_cast<const int2*>(tile_ranges);
  hipLaunchKernelGGL(ps_tile_order_keys, dim3(kPlanSlots), dim3(kOrderThreads), 0, stream, rg,
                     tiles, gx, bucket, w.words, w.plan);
  PS_LAUNCH_CHECK();
  DepthSortArgs A;
  A.plan = w.plan;
  A.words = w.words;
  A.kb[0] = w.kb[0];
  A.kb[1] = w.kb[1];
  A.vb[0] = w.vb[0];
  A.vb[1] = w.vb[1];
  A.n = tiles;
  int st = depth_sort_launch(A, tiles, w.sort, w.sort_bytes, stream);
  if (st != PS_OK) return st;
  int blocks = (tiles + kOrderThreads - 1) / kOrderThreads;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(ps_tile_order_finish, dim3(blocks), dim3(kOrderThreads), 0, stream,
                     (const uint32_t*)w.plan, (const uint32_t*)w.words, w.kb[0], w.kb[1],
                     w.vb[0], w.vb[1], rg, tiles, gx, zigzag, order, keys_out, ids_out,
                     reinterpret_cast<int4*>(desc));
  PS_LAUNCH_CHECK();
  return PS_OK;
}

size_t tile_order_temp_size(int tiles) { return order_layout(tiles, nullptr, nullptr); }

}  // namespace ps

extern "C" size_t ps_get_tile_order_sort_temp_size(int total_tiles) {
  return ps::tile_order_temp_size(total_tiles < 0 ? 0 : total_tiles);
}

extern "C" int ps_compute_tile_order(const int* tile_ranges, int total_tiles, uint32_t* tile_order,
                                     uint32_t* tile_counts, uint32_t* tile_ids, void* sort_temp,
                                     size_t sort_temp_bytes, ps_stream_t stream) {
  PS_REQUIRE(tile_order != nullptr, "tile_order is NULL");
  return ps::tile_order_launch(tile_ranges, total_tiles, 1, 0, 0, tile_order, tile_counts,
                               tile_ids, nullptr, sort_temp, sort_temp_bytes,
                               reinterpret_cast<hipStream_t>(stream));
}

extern "C" int ps_compute_tile_order_packed(const int* tile_ranges, int total_tiles, int x_blocks,
                                            int* tile_desc, uint32_t* tile_order,
                                            uint32_t* tile_counts, uint32_t* tile_ids,
                                            void* sort_temp, size_t sort_temp_bytes,
                                            int zigzag_group_size, int morton_bucket_size,
                                            ps_stream_t stream) {
  PS_REQUIRE(tile_desc != nullptr, "tile_desc is NULL");
  return ps::tile_order_launch(tile_ranges, total_tiles, x_blocks, zigzag_group_size,
                               morton_bucket_size, tile_order, tile_counts, tile_ids, tile_desc,
                               sort_temp, sort_temp_bytes, reinterpret_cast<hipStream_t>(stream));
}
