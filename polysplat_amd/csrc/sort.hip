// Stable LSD radix sort of (u64 key, u32 value) pairs for gfx950.
//
// Replaces the reference's CUB onesweep dispatch (sort.cu:80-125).  Each
// 8-bit digit pass runs three launches with no inter-block waiting:
//   upsweep   : every block histograms its contiguous share of the keys;
//   scan      : one workgroup per digit scans that digit's counts over
//               blocks and records the digit total;
//   downsweep : every block ranks its keys tile by tile (wave64 ballot
//               match per digit against per-wave LDS counters), stages them
//               in LDS grouped by digit and writes each digit run
//               contiguously, prefetching the next tile meanwhile.
// A fixed number of blocks (kSortBlocks) covers any pair count by giving
// each block a contiguous run of 3840-pair tiles, so every launch shape is
// independent of the count, which may live in device memory.  (A
// decoupled look-back variant was measured slower here: its chain of
// device-scope status loads costs more than the upsweep it saves.)
// Stability: within a tile, pairs keep their index order (round by round,
// wave by wave, lane by lane); tiles of a block are processed in order and
// blocks own ascending ranges, so equal digits keep their input order.

#include "ps_common.h"
#include "ps_host.h"
#include "ps_wave.h"


namespace ps {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 15;
constexpr int kSortTile = kSortThreads * kSortItems;  // 3840 pairs
constexpr int kWaveItems = kWave * kSortItems;        // 960 pairs per wave
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

// Tile item layout: wave w owns items [w*960, w*960 + 960) of the tile; item
// r of lane l is w*960 + r*64 + l, so each load instruction of a wave reads
// 64 consecutive pairs, and item order = (wave, round, lane).
__device__ __forceinline__ int tile_item(int wave, int r, int lane) {
  return wave * kWaveItems + r * kWave + lane;
}

// counts layout: [digit][block].  Every block histograms the same contiguous
// run of tiles the downsweep will give it.
__global__ __launch_bounds__(kSortThreads) void ps_sort_upsweep(const uint64_t* __restrict__ keys,
                                                                const int* __restrict__ d_n,
                                                                int n_host, int max_pairs,
                                                                int shift,
                                                                uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[4][kRadix];
  const SortPlan pl = sort_plan(d_n, n_host, max_pairs);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, b = blockIdx.x;
  hist[0][tid] = 0;
  hist[1][tid] = 0;
  hist[2][tid] = 0;
  hist[3][tid] = 0;
  __syncthreads();
  const int t_begin = b * pl.tiles_per_block;
  int t_end = t_begin + pl.tiles_per_block;
  if (t_end > pl.ntiles) t_end = pl.ntiles;
  for (int t = t_begin; t < t_end; ++t) {
    const long long tile0 = (long long)t * kSortTile;
    const int tile_n = (pl.n - tile0 < kSortTile) ? (int)(pl.n - tile0) : kSortTile;
    int d[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {  // all loads issue before any use
      const int i = tile_item(wave, r, lane);
      d[r] = (i < tile_n) ? digit_of(keys[tile0 + i], shift) : -1;
    }
#pragma unroll
    for (int r = 0; r < kSortItems; ++r)
      if (d[r] >= 0) atomicAdd(&hist[wave][d[r]], 1u);
  }
  __syncthreads();
  counts[(size_t)tid * kSortBlocks + b] = hist[0][tid] + hist[1][tid] + hist[2][tid] + hist[3][tid];
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

__device__ __forceinline__ void load_tile(const uint64_t* __restrict__ keys,
                                          const uint32_t* __restrict__ vals, long long tile0,
                                          int tile_n, int wave, int lane, uint64_t (&k)[kSortItems],
                                          uint32_t (&v)[kSortItems]) {
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int i = tile_item(wave, r, lane);
    if (i < tile_n) {
      k[r] = keys[tile0 + i];
      v[r] = vals[tile0 + i];
    } else {
      k[r] = 0;
      v[r] = 0;
    }
  }
}

// Rank, stage and scatter.  Per tile: each wave ranks its 960 items against
// its own LDS digit counters (a wave's LDS operations execute in order, so no
// barrier is needed between rounds); one block scan turns the per-wave counts
// into tile-local starts; items are staged in LDS grouped by digit and each
// digit run is written contiguously.  The next tile's keys are loaded into
// registers while the current one is scanned, staged and written (about 210
// VGPRs: two workgroups per CU, no spills).
__global__ __launch_bounds__(kSortThreads) void ps_sort_downsweep(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint64_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, const int* __restrict__ d_n,
    int n_host, int max_pairs, int shift, const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ totals) {
  __shared__ uint64_t skey[kSortTile];
  __shared__ uint32_t sval[kSortTile];
  // wcnt: per-wave digit counters while ranking, then (same slots) the
  // tile-local start of each (wave, digit) run; wcnt[0][d] is digit d's
  // tile-local start.
  __shared__ uint32_t wcnt[4][kRadix];
  __shared__ uint32_t gbase[kRadix];
  __shared__ int scratch[4];

  const SortPlan pl = sort_plan(d_n, n_host, max_pairs);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, b = blockIdx.x;
  const int t_begin = b * pl.tiles_per_block;
  int t_end = t_begin + pl.tiles_per_block;
  if (t_end > pl.ntiles) t_end = pl.ntiles;
  if (t_begin >= t_end) return;  // uniform across the block

  int dtotal;
  const int dbase = block256_excl_scan((int)totals[tid], scratch, dtotal);
  gbase[tid] = (uint32_t)dbase + offsets[(size_t)tid * kSortBlocks + b];
  wcnt[0][tid] = 0;
  wcnt[1][tid] = 0;
  wcnt[2][tid] = 0;
  wcnt[3][tid] = 0;
  const uint64_t below = lanes_below(lane);

  uint64_t k[kSortItems];
  uint32_t v[kSortItems];
  long long tile0 = (long long)t_begin * kSortTile;
  int tile_n = (pl.n - tile0 < kSortTile) ? (int)(pl.n - tile0) : kSortTile;
  load_tile(keys_in, vals_in, tile0, tile_n, wave, lane, k, v);
  __syncthreads();

  for (int t = t_begin; t < t_end; ++t) {
    int dg[kSortItems];
    uint32_t pos[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      const int d = (tile_item(wave, r, lane) < tile_n) ? digit_of(k[r], shift) : -1;
      dg[r] = d;
      uint64_t peers = ballot(d >= 0);
#pragma unroll
      for (int bit = 0; bit < 8; ++bit) {
        const bool set = d >= 0 && ((d >> bit) & 1);
        const uint64_t bb = ballot(set);
        peers &= set ? bb : ~bb;
      }
      const int rank = popc64(peers & below);
      uint32_t base = 0;
      if (d >= 0) base = wcnt[wave][d];
      __builtin_amdgcn_wave_barrier();
      if (d >= 0 && rank == 0) wcnt[wave][d] = base + (uint32_t)popc64(peers);
      __builtin_amdgcn_wave_barrier();
      pos[r] = base + (uint32_t)rank;
    }
    // Prefetch the next tile while this one is scanned, staged and written.
    uint64_t kn[kSortItems];
    uint32_t vn[kSortItems];
    const long long ntile0 = tile0 + kSortTile;
    const int ntile_n =
        (t + 1 < t_end) ? ((pl.n - ntile0 < kSortTile) ? (int)(pl.n - ntile0) : kSortTile) : 0;
    if (ntile_n > 0) load_tile(keys_in, vals_in, ntile0, ntile_n, wave, lane, kn, vn);
    __syncthreads();
    const uint32_t c0 = wcnt[0][tid], c1 = wcnt[1][tid], c2 = wcnt[2][tid], c3 = wcnt[3][tid];
    const uint32_t tot = c0 + c1 + c2 + c3;
    int tile_total;
    const uint32_t ls = (uint32_t)block256_excl_scan((int)tot, scratch, tile_total);
    wcnt[0][tid] = ls;
    wcnt[1][tid] = ls + c0;
    wcnt[2][tid] = ls + c0 + c1;
    wcnt[3][tid] = ls + c0 + c1 + c2;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      if (dg[r] >= 0) {
        const uint32_t p = wcnt[wave][dg[r]] + pos[r];
        skey[p] = k[r];
        sval[p] = v[r];
      }
    }
    __syncthreads();
    for (int i = tid; i < tile_n; i += kSortThreads) {
      const uint64_t key = skey[i];
      const int d = digit_of(key, shift);
      const uint32_t dst = gbase[d] + ((uint32_t)i - wcnt[0][d]);
      if (dst < (uint32_t)pl.n) {
        keys_out[dst] = key;
        vals_out[dst] = sval[i];
      }
    }
    __syncthreads();
    gbase[tid] += tot;
    wcnt[0][tid] = 0;
    wcnt[1][tid] = 0;
    wcnt[2][tid] = 0;
    wcnt[3][tid] = 0;
    tile0 = ntile0;
    tile_n = ntile_n;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      k[r] = kn[r];
      v[r] = vn[r];
    }
    __syncthreads();
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
