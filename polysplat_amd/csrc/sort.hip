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

#include <stdlib.h>
#include <string.h>

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

// ---------------------------------------------------------------------------
// Onesweep variant: one histogram pass over the keys for every digit
// position, then one kernel per digit.  Tiles are claimed in ticket order,
// so a tile only ever waits for tiles claimed before it, which are already
// running; each tile publishes its digit counts before it looks back, so the
// chain of waits always ends at tile 0.  Every wait is bounded: past the
// limit (or once any wait has timed out) it gives up, flags the error word
// and continues, so the kernel always terminates; writes stay in bounds.
// ---------------------------------------------------------------------------
constexpr int kOsItems = 15;
constexpr int kOsTile = kSortThreads * kOsItems;  // 3840 pairs
constexpr int kOsWaveItems = kWave * kOsItems;    // 960 pairs per wave
constexpr int kMaxPasses = 8;
constexpr int kHistBlocks = 512;
constexpr uint32_t kFlagAgg = 1u << 30, kFlagPre = 2u << 30, kValMask = (1u << 30) - 1u;
constexpr int kSpinLimit = 1 << 18;

__device__ __forceinline__ uint32_t ld_status(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kSortThreads) void ps_os_zero(uint32_t* __restrict__ digit_counts,
                                                           uint32_t* __restrict__ tickets,
                                                           uint32_t* __restrict__ err) {
  for (int i = threadIdx.x; i < kMaxPasses * kRadix; i += kSortThreads) digit_counts[i] = 0;
  if (threadIdx.x < kMaxPasses) tickets[threadIdx.x] = 0;
  if (threadIdx.x == 0) *err = 0;
}

// Histograms of every digit position; also clears the first pass's status rows.
__global__ __launch_bounds__(kSortThreads) void ps_os_hist(
    const uint64_t* __restrict__ keys, const int* __restrict__ d_n, int n_host, int max_pairs,
    int passes, uint32_t* __restrict__ digit_counts, uint32_t* __restrict__ status0) {
  __shared__ uint32_t h[kMaxPasses][kRadix];
  const SortPlan pl = sort_plan(d_n, n_host, max_pairs);
  const int tid = threadIdx.x;
  for (int i = tid; i < kMaxPasses * kRadix; i += kSortThreads) (&h[0][0])[i] = 0;
  const int ntiles = (pl.n + kOsTile - 1) / kOsTile;
  for (long long i = (long long)blockIdx.x * kSortThreads + tid; i < (long long)ntiles * kRadix;
       i += (long long)gridDim.x * kSortThreads)
    status0[i] = 0;
  __syncthreads();
  for (long long i = (long long)blockIdx.x * kSortThreads + tid; i < pl.n;
       i += (long long)gridDim.x * kSortThreads) {
    const uint64_t k = keys[i];
    for (int p = 0; p < passes; ++p) atomicAdd(&h[p][digit_of(k, 8 * p)], 1u);
  }
  __syncthreads();
  for (int p = 0; p < passes; ++p) {
    const uint32_t c = h[p][tid];
    if (c) atomicAdd(&digit_counts[p * kRadix + tid], c);
  }
}

__global__ __launch_bounds__(kSortThreads) void ps_os_pass(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint64_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, const int* __restrict__ d_n,
    int n_host, int max_pairs, int pass, const uint32_t* __restrict__ digit_counts,
    uint32_t* __restrict__ tickets, uint32_t* __restrict__ status,
    uint32_t* __restrict__ status_next, uint32_t* __restrict__ err) {
  __shared__ uint64_t skey[kOsTile];
  __shared__ uint32_t sval[kOsTile];
  __shared__ uint32_t wcnt[4][kRadix];
  __shared__ uint32_t gbase[kRadix];
  __shared__ uint32_t lstart[kRadix];
  __shared__ int scratch[4];
  __shared__ int s_ticket;

  const SortPlan pl = sort_plan(d_n, n_host, max_pairs);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int shift = 8 * pass;
  const int ntiles = (pl.n + kOsTile - 1) / kOsTile;
  if (tid == 0) s_ticket = (int)atomicAdd(&tickets[pass], 1u);
  wcnt[0][tid] = 0;
  wcnt[1][tid] = 0;
  wcnt[2][tid] = 0;
  wcnt[3][tid] = 0;
  __syncthreads();
  const int t = s_ticket;
  if (t >= ntiles) return;  // uniform across the block
  const long long tile0 = (long long)t * kOsTile;
  int tile_n = pl.n - (int)tile0;
  if (tile_n > kOsTile) tile_n = kOsTile;

  // Wave w owns tile items [w*960, w*960 + 960): item k of lane l is w*960 + k*64 + l.
  uint64_t k[kOsItems];
  uint32_t v[kOsItems];
  int dg[kOsItems];
#pragma unroll
  for (int r = 0; r < kOsItems; ++r) {
    const int i = wave * kOsWaveItems + r * kWave + lane;
    if (i < tile_n) {
      k[r] = keys_in[tile0 + i];
      v[r] = vals_in[tile0 + i];
      dg[r] = digit_of(k[r], shift);
    } else {
      k[r] = 0;
      v[r] = 0;
      dg[r] = -1;
    }
  }
  // Per-wave stable ranking: pos[r] = position among this wave's items of the
  // same digit.  The wave's running counters live in its own LDS row, so no
  // barrier is needed inside the rounds (a wave's LDS operations execute in
  // program order).
  const uint64_t below = lanes_below(lane);
  uint32_t pos[kOsItems];
#pragma unroll
  for (int r = 0; r < kOsItems; ++r) {
    const int d = dg[r];
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
  __syncthreads();
  // Digit `tid`: tile count, tile-local start, per-wave starts.
  const uint32_t c0 = wcnt[0][tid], c1 = wcnt[1][tid], c2 = wcnt[2][tid], c3 = wcnt[3][tid];
  const uint32_t tot = c0 + c1 + c2 + c3;
  int tile_total;
  const int ls = block256_excl_scan((int)tot, scratch, tile_total);
  lstart[tid] = (uint32_t)ls;
  wcnt[0][tid] = (uint32_t)ls;
  wcnt[1][tid] = (uint32_t)ls + c0;
  wcnt[2][tid] = (uint32_t)ls + c0 + c1;
  wcnt[3][tid] = (uint32_t)ls + c0 + c1 + c2;
  // Global digit start: all keys with smaller digits, then this digit's
  // pairs in earlier tiles (decoupled look-back).
  int dtotal;
  const int dbase = block256_excl_scan((int)digit_counts[pass * kRadix + tid], scratch, dtotal);
  uint32_t* row = status + (size_t)t * kRadix;
  uint32_t excl = 0;
  if (t == 0) {
    st_status(&row[tid], kFlagPre | tot);
  } else {
    st_status(&row[tid], kFlagAgg | tot);
    int j = t - 1;
    bool gave_up = false;
    while (j >= 0) {
      uint32_t s = 0;
      int spins = 0;
      for (;;) {
        s = ld_status(&status[(size_t)j * kRadix + tid]);
        if (s & (kFlagAgg | kFlagPre)) break;
        if (++spins >= kSpinLimit ||
            ((spins & 63) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT) != 0)) {
          gave_up = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      if (gave_up) {
        atomicOr(err, 1u);
        break;
      }
      excl += s & kValMask;
      if (s & kFlagPre) break;
      --j;
    }
    st_status(&row[tid], kFlagPre | ((excl + tot) & kValMask));
  }
  gbase[tid] = (uint32_t)dbase + excl;
  __syncthreads();
  // Stage the tile grouped by digit, then write every digit run contiguously.
#pragma unroll
  for (int r = 0; r < kOsItems; ++r) {
    const int d = dg[r];
    if (d >= 0) {
      const uint32_t p = wcnt[wave][d] + pos[r];
      skey[p] = k[r];
      sval[p] = v[r];
    }
  }
  __syncthreads();
  for (int i = tid; i < tile_n; i += kSortThreads) {
    const uint64_t key = skey[i];
    const int d = digit_of(key, shift);
    const uint32_t dst = gbase[d] + ((uint32_t)i - lstart[d]);
    if (dst < (uint32_t)pl.n) {
      keys_out[dst] = key;
      vals_out[dst] = sval[i];
    }
  }
  // Clear this tile's status row for the next pass.
  if (status_next) status_next[(size_t)t * kRadix + tid] = 0;
}

struct SortWs {
  uint64_t* alt_keys;
  uint32_t* alt_vals;
  uint32_t* counts;
  uint32_t* offsets;
  uint32_t* totals;
  uint32_t* digit_counts;
  uint32_t* tickets;
  uint32_t* err;
  uint32_t* status[2];
};

inline size_t sort_layout(int cap, char* base, SortWs* w) {
  size_t off = 0;
  const size_t c = (size_t)(cap > 0 ? cap : 1);
  const size_t tiles = (c + kOsTile - 1) / kOsTile;
  auto take = [&](size_t bytes) {
    char* p = base ? base + off : nullptr;
    off += align_up(bytes, 256);
    return p;
  };
  char* p;
  p = take(c * sizeof(uint64_t));
  if (w) w->alt_keys = reinterpret_cast<uint64_t*>(p);
  p = take(c * sizeof(uint32_t));
  if (w) w->alt_vals = reinterpret_cast<uint32_t*>(p);
  p = take((size_t)kRadix * kSortBlocks * sizeof(uint32_t));
  if (w) w->counts = reinterpret_cast<uint32_t*>(p);
  p = take((size_t)kRadix * kSortBlocks * sizeof(uint32_t));
  if (w) w->offsets = reinterpret_cast<uint32_t*>(p);
  p = take((size_t)kRadix * sizeof(uint32_t));
  if (w) w->totals = reinterpret_cast<uint32_t*>(p);
  p = take((size_t)kMaxPasses * kRadix * sizeof(uint32_t));
  if (w) w->digit_counts = reinterpret_cast<uint32_t*>(p);
  p = take(64 * sizeof(uint32_t));
  if (w) {
    w->tickets = reinterpret_cast<uint32_t*>(p);
    w->err = reinterpret_cast<uint32_t*>(p) + 32;
  }
  for (int k = 0; k < 2; ++k) {
    p = take(tiles * kRadix * sizeof(uint32_t));
    if (w) w->status[k] = reinterpret_cast<uint32_t*>(p);
  }
  return off;
}

// Which sort implementation to run: onesweep unless POLYSPLAT_SORT=reduce_scan
// (read on every call so a process can compare both).
inline bool use_onesweep() {
  const char* e = getenv("POLYSPLAT_SORT");
  return !(e && strcmp(e, "reduce_scan") == 0);
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
  PS_REQUIRE(max_pairs < (1 << 30), "at most 2^30 - 1 pairs can be sorted");
  const uint64_t* kin = keys_in;
  const uint32_t* vin = vals_in;
  if (use_onesweep()) {
    const int n_grid = d_n ? max_pairs : n_host;
    const int tiles = (n_grid + kOsTile - 1) / kOsTile;
    hipLaunchKernelGGL(ps_os_zero, dim3(1), dim3(kSortThreads), 0, stream, w.digit_counts,
                       w.tickets, w.err);
    PS_LAUNCH_CHECK();
    int hb = (n_grid + kSortThreads * 16 - 1) / (kSortThreads * 16);
    hb = hb < 1 ? 1 : (hb > kHistBlocks ? kHistBlocks : hb);
    hipLaunchKernelGGL(ps_os_hist, dim3(hb), dim3(kSortThreads), 0, stream, keys_in, d_n, n_host,
                       max_pairs, passes, w.digit_counts, w.status[0]);
    PS_LAUNCH_CHECK();
    for (int p = 0; p < passes; ++p) {
      const bool to_out = ((passes - 1 - p) % 2) == 0;
      uint64_t* kout = to_out ? keys_out : w.alt_keys;
      uint32_t* vout = to_out ? vals_out : w.alt_vals;
      hipLaunchKernelGGL(ps_os_pass, dim3(tiles > 0 ? tiles : 1), dim3(kSortThreads), 0, stream,
                         kin, vin, kout, vout, d_n, n_host, max_pairs, p, w.digit_counts,
                         w.tickets, w.status[p & 1], p + 1 < passes ? w.status[(p + 1) & 1] : nullptr,
                         w.err);
      PS_LAUNCH_CHECK();
      kin = kout;
      vin = vout;
    }
    return PS_OK;
  }
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

extern "C" const unsigned* ps_sort_error_ptr(const void* sort_space, int max_pairs) {
  ps::SortWs w;
  ps::sort_layout(max_pairs < 0 ? 0 : max_pairs, static_cast<char*>(const_cast<void*>(sort_space)),
                  &w);
  return w.err;
}

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
