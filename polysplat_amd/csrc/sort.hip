// Stable LSD radix sort of (u64 key, u32 value) pairs for gfx950.
//
// Replaces the reference's CUB onesweep dispatch (sort.cu:80-125).  Each
// 8-bit digit pass runs three launches with no inter-block waiting:
//   upsweep   : every block histograms its contiguous share of the keys;
//   scan      : one workgroup per digit scans that digit's counts over
//               blocks and records the digit total;
//   downsweep : every block ranks its keys tile by tile (wave64 ballot
//               match per digit), stages them in LDS grouped by digit and
//               writes each digit run contiguously.
// A fixed number of blocks (kSortBlocks) covers any pair count by giving
// each block a contiguous run of 2048-pair tiles, so every launch shape is
// independent of the count, which may live in device memory.
// Stability: within a tile, pairs keep their index order (round by round,
// wave by wave, lane by lane); tiles of a block are processed in order and
// blocks own ascending ranges, so equal digits keep their input order.

#include "ps_common.h"
#include "ps_host.h"
#include "ps_wave.h"

namespace ps {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;  // 2048 pairs
constexpr int kSortBlocks = 1024;
constexpr int kRadix = 256;

struct SortPlan {
  int n, ntiles, tiles_per_block;
};

__device__ __forceinline__ SortPlan sort_plan(const int* d_n, int n_host, int max_pairs) {
  SortPlan p;
  p.n = d_n ? *d_n : n_host;
  if (p.n < 0) p.n = 0;
  if (p.n > max_pairs) p.n = max_pairs;  // a device count never exceeds the buffers
  p.ntiles = (p.n + kSortTile - 1) / kSortTile;
  p.tiles_per_block = (p.ntiles + kSortBlocks - 1) / kSortBlocks;
  return p;
}

__device__ __forceinline__ int digit_of(uint64_t k, int shift) {
  return (int)((k >> shift) & 0xffu);
}

// counts layout: [digit][block]
__global__ __launch_bounds__(kSortThreads) void ps_sort_upsweep(const uint64_t* __restrict__ keys,
                                                                const int* __restrict__ d_n,
                                                                int n_host, int max_pairs,
                                                                int shift,
                                                                uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[4][kRadix];
  const SortPlan pl = sort_plan(d_n, n_host, max_pairs);
  const int tid = threadIdx.x, wave = tid >> 6, b = blockIdx.x;
  for (int i = tid; i < 4 * kRadix; i += kSortThreads) (&hist[0][0])[i] = 0;
  __syncthreads();
  const long long begin = (long long)b * pl.tiles_per_block * kSortTile;
  long long end = begin + (long long)pl.tiles_per_block * kSortTile;
  if (end > pl.n) end = pl.n;
  for (long long i = begin + tid; i < end; i += kSortThreads)
    atomicAdd(&hist[wave][digit_of(keys[i], shift)], 1u);
  __syncthreads();
  const uint32_t c = hist[0][tid] + hist[1][tid] + hist[2][tid] + hist[3][tid];
  counts[(size_t)tid * kSortBlocks + b] = c;
}

// One workgroup per digit: exclusive scan of counts[d][0..kSortBlocks) into
// offsets[d][...], digit total into totals[d].
__global__ __launch_bounds__(kSortThreads) void ps_sort_scan(const uint32_t* __restrict__ counts,
                                                             uint32_t* __restrict__ offsets,
                                                             uint32_t* __restrict__ totals) {
  __shared__ int scratch[4];
  constexpr int per = kSortBlocks / kSortThreads;
  const int d = blockIdx.x, tid = threadIdx.x;
  const uint32_t* row = counts + (size_t)d * kSortBlocks;
  uint32_t v[per];
  int local = 0;
#pragma unroll
  for (int k = 0; k < per; ++k) {
    v[k] = row[tid * per + k];
    local += (int)v[k];
  }
  int total;
  int run = block256_excl_scan(local, scratch, total);
  uint32_t* orow = offsets + (size_t)d * kSortBlocks;
#pragma unroll
  for (int k = 0; k < per; ++k) {
    orow[tid * per + k] = (uint32_t)run;
    run += (int)v[k];
  }
  if (tid == 0) totals[d] = (uint32_t)total;
}

__global__ __launch_bounds__(kSortThreads) void ps_sort_downsweep(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint64_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, const int* __restrict__ d_n,
    int n_host, int max_pairs, int shift, const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ totals) {
  __shared__ uint64_t skey[kSortTile];
  __shared__ uint32_t sval[kSortTile];
  __shared__ uint32_t hist[kRadix];
  __shared__ uint32_t wcnt[4][kRadix];
  __shared__ uint32_t wpre[4][kRadix];
  __shared__ uint32_t lstart[kRadix];
  __shared__ uint32_t gbase[kRadix];
  __shared__ int scratch[4];

  const SortPlan pl = sort_plan(d_n, n_host, max_pairs);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, b = blockIdx.x;
  const int t_begin = b * pl.tiles_per_block;
  int t_end = t_begin + pl.tiles_per_block;
  if (t_end > pl.ntiles) t_end = pl.ntiles;
  if (t_begin >= t_end) return;  // uniform across the block

  // Global start of digit `tid` for this block = digit prefix + block offset.
  int dtotal;
  const int dbase = block256_excl_scan((int)totals[tid], scratch, dtotal);
  gbase[tid] = (uint32_t)dbase + offsets[(size_t)tid * kSortBlocks + b];
  wcnt[0][tid] = 0;
  wcnt[1][tid] = 0;
  wcnt[2][tid] = 0;
  wcnt[3][tid] = 0;
  const uint64_t below = lanes_below(lane);

  for (int t = t_begin; t < t_end; ++t) {
    const long long tile0 = (long long)t * kSortTile;
    int tile_n = pl.n - (int)tile0;
    if (tile_n > kSortTile) tile_n = kSortTile;
    uint64_t k[kSortItems];
    uint32_t v[kSortItems];
    int dg[kSortItems];
    hist[tid] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      const int i = r * kSortThreads + tid;
      if (i < tile_n) {
        k[r] = keys_in[tile0 + i];
        v[r] = vals_in[tile0 + i];
        dg[r] = digit_of(k[r], shift);
        atomicAdd(&hist[dg[r]], 1u);
      } else {
        k[r] = 0;
        v[r] = 0;
        dg[r] = -1;
      }
    }
    __syncthreads();
    const int my_cnt = (int)hist[tid];
    int tot;
    int run = block256_excl_scan(my_cnt, scratch, tot);
    lstart[tid] = (uint32_t)run;

#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      const int d = dg[r];
      // Lanes of this wave holding the same digit as this lane.
      uint64_t peers = ballot(d >= 0);
#pragma unroll
      for (int bit = 0; bit < 8; ++bit) {
        const bool set = d >= 0 && ((d >> bit) & 1);
        const uint64_t bb = ballot(set);
        peers &= set ? bb : ~bb;
      }
      const int rank = popc64(peers & below);
      if (d >= 0 && rank == 0) wcnt[wave][d] = (uint32_t)popc64(peers);
      __syncthreads();
      {
        const uint32_t c0 = wcnt[0][tid], c1 = wcnt[1][tid], c2 = wcnt[2][tid], c3 = wcnt[3][tid];
        wpre[0][tid] = (uint32_t)run;
        wpre[1][tid] = (uint32_t)run + c0;
        wpre[2][tid] = (uint32_t)run + c0 + c1;
        wpre[3][tid] = (uint32_t)run + c0 + c1 + c2;
        run += (int)(c0 + c1 + c2 + c3);
        wcnt[0][tid] = 0;
        wcnt[1][tid] = 0;
        wcnt[2][tid] = 0;
        wcnt[3][tid] = 0;
      }
      __syncthreads();
      if (d >= 0) {
        const uint32_t pos = wpre[wave][d] + (uint32_t)rank;
        skey[pos] = k[r];
        sval[pos] = v[r];
      }
    }
    __syncthreads();
    // Digit runs are contiguous in LDS; each run lands contiguously in HBM.
    for (int i = tid; i < tile_n; i += kSortThreads) {
      const uint64_t key = skey[i];
      const int d = digit_of(key, shift);
      const uint32_t dst = gbase[d] + ((uint32_t)i - lstart[d]);
      keys_out[dst] = key;
      vals_out[dst] = sval[i];
    }
    __syncthreads();
    gbase[tid] += (uint32_t)my_cnt;
  }
}

struct SortWs {
  uint64_t* alt_keys;
  uint32_t* alt_vals;
  uint32_t* counts;
  uint32_t* offsets;
  uint32_t* totals;
};

inline size_t sort_layout(int cap, char* base, SortWs* w) {
  size_t off = 0;
  const size_t c = (size_t)(cap > 0 ? cap : 1);
  if (w) w->alt_keys = reinterpret_cast<uint64_t*>(base + off);
  off += align_up(c * sizeof(uint64_t), 256);
  if (w) w->alt_vals = reinterpret_cast<uint32_t*>(base + off);
  off += align_up(c * sizeof(uint32_t), 256);
  if (w) w->counts = reinterpret_cast<uint32_t*>(base + off);
  off += align_up((size_t)kRadix * kSortBlocks * sizeof(uint32_t), 256);
  if (w) w->offsets = reinterpret_cast<uint32_t*>(base + off);
  off += align_up((size_t)kRadix * kSortBlocks * sizeof(uint32_t), 256);
  if (w) w->totals = reinterpret_cast<uint32_t*>(base + off);
  off += align_up((size_t)kRadix * sizeof(uint32_t), 256);
  return off;
}

}  // namespace

int sort_launch(int n_host, const int* d_n, int max_pairs, int end_bit, void* space,
                size_t space_bytes, const uint64_t* keys_in, const uint32_t* vals_in,
                uint64_t* keys_out, uint32_t* vals_out, hipStream_t stream) {
  PS_REQUIRE(max_pairs >= 0, "max_pairs must be >= 0");
  PS_REQUIRE(d_n != nullptr || (n_host >= 0 && n_host <= max_pairs),
             "num_rendered %d outside [0, %d]", n_host, max_pairs);
  PS_REQUIRE(space != nullptr && space_bytes >= sort_layout(max_pairs, nullptr, nullptr),
             "sort space too small (%zu < %zu)", space_bytes,
             sort_layout(max_pairs, nullptr, nullptr));
  PS_REQUIRE(end_bit > 0 && end_bit <= 64, "bad end bit %d", end_bit);
  PS_REQUIRE(keys_in != keys_out && vals_in != vals_out, "sort outputs must not alias inputs");
  if (d_n == nullptr && n_host == 0) return PS_OK;
  SortWs w;
  sort_layout(max_pairs, static_cast<char*>(space), &w);
  const int passes = (end_bit + 7) / 8;
  const uint64_t* kin = keys_in;
  const uint32_t* vin = vals_in;
  for (int p = 0; p < passes; ++p) {
    // Ping-pong so that the last pass lands in the caller's output.
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    uint64_t* kout = to_out ? keys_out : w.alt_keys;
    uint32_t* vout = to_out ? vals_out : w.alt_vals;
    const int shift = 8 * p;
    hipLaunchKernelGGL(ps_sort_upsweep, dim3(kSortBlocks), dim3(kSortThreads), 0, stream, kin,
                       d_n, n_host, max_pairs, shift, w.counts);
    PS_LAUNCH_CHECK();
    hipLaunchKernelGGL(ps_sort_scan, dim3(kRadix), dim3(kSortThreads), 0, stream, w.counts,
                       w.offsets, w.totals);
    PS_LAUNCH_CHECK();
    hipLaunchKernelGGL(ps_sort_downsweep, dim3(kSortBlocks), dim3(kSortThreads), 0, stream, kin,
                       vin, kout, vout, d_n, n_host, max_pairs, shift, w.offsets, w.totals);
    PS_LAUNCH_CHECK();
    kin = kout;
    vin = vout;
  }
  return PS_OK;
}

int sort_end_bit(int width, int height, int block_x, int block_y) {
  const unsigned tiles = (unsigned)(((width + block_x - 1) / block_x) *
                                    ((height + block_y - 1) / block_y));
  return 32 + bit_length(tiles);
}

}  // namespace ps

extern "C" size_t ps_get_sort_buffer_size(int max_pairs) {
  return ps::sort_layout(max_pairs < 0 ? 0 : max_pairs, nullptr, nullptr);
}

extern "C" int ps_sort_gaussian(int num_rendered, const int* d_num_rendered, int max_pairs,
                                int width, int height, int block_x, int block_y, void* sort_space,
                                size_t sort_space_bytes, const uint64_t* keys_unsorted,
                                const uint32_t* values_unsorted, uint64_t* keys_sorted,
                                uint32_t* values_sorted, ps_stream_t stream) {
  PS_REQUIRE(width > 0 && height > 0 && block_x > 0 && block_y > 0, "bad image/tile size");
  return ps::sort_launch(num_rendered, d_num_rendered, max_pairs,
                         ps::sort_end_bit(width, height, block_x, block_y), sort_space,
                         sort_space_bytes, keys_unsorted, values_unsorted, keys_sorted,
                         values_sorted, reinterpret_cast<hipStream_t>(stream));
}
