// Stable LSD radix sort of (u64 key, u32 value) pairs for gfx950.
//
// Replaces the reference's CUB onesweep dispatch (sort.cu:80-125).  Each
// 8-bit digit pass runs three launches with no inter-block waiting:
//   upsweep   : every block histograms its contiguous share of the keys;
//   scan      : one workgroup per digit scans that digit's counts over
//               blocks and records the digit total;
//   downsweep : every block ranks its keys tile by tile (per-wave digit
//               match through LDS lane masks, against per-wave LDS
//               counters), stages them
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

#include <stdlib.h>


namespace ps {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 15;
constexpr int kSortTile = kSortThreads * kSortItems;  // 3840 pairs
constexpr int kWaveItems = kWave * kSortItems;        // 960 pairs per wave
constexpr int kSortBlocks = 1024;
constexpr int kRadix = 256;
// Digit passes take 8-bit digits, or 9-bit ones (512 buckets) when that saves
// a pass (18 tile bits at 9216x6912: two passes instead of three).
constexpr int kMaxDigitBits = 9;
constexpr int kMaxRadix = 1 << kMaxDigitBits;

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

// A digit is `width` (1..9) key bits from `shift`, packed as
// shift | width << 8.  The last digit pass of a sort over [0, end_bit) takes
// only the bits below end_bit, so higher key bits never affect the order.
__host__ __device__ constexpr int digit_spec(int shift, int width) { return shift | (width << 8); }

__device__ __forceinline__ int digit_of(uint64_t k, int spec) {
  return (int)((k >> (spec & 63)) & ((1u << (spec >> 8)) - 1u));
}

// Tile item layout: wave w owns items [w*960, w*960 + 960) of the tile; item
// r of lane l is w*960 + r*64 + l, so each load instruction of a wave reads
// 64 consecutive pairs, and item order = (wave, round, lane).
__device__ __forceinline__ int tile_item(int wave, int r, int lane) {
  return wave * kWaveItems + r * kWave + lane;
}

// Where a digit pass reads its pairs: arrays of keys (64-bit sort keys, or
// 32-bit tile ids / depth words) and values, or, in the first pass of the
// depth sort of the depth-first pipeline, the count pass's depth words with
// each Gaussian's id as its value.
template <typename KeyT>
struct ArraySrc {
  const KeyT* k;
  const uint32_t* v;
  __device__ __forceinline__ KeyT key(long long i) const { return k[i]; }
  __device__ __forceinline__ uint32_t val(long long i) const { return v[i]; }
};
struct DepthSrc {
  const uint32_t* k;
  uint32_t subst;  // key for a Gaussian that emits nothing (depth word 0)
  __device__ __forceinline__ uint32_t key(long long i) const {
    const uint32_t x = k[i];
    return x ? x : subst;
  }
  __device__ __forceinline__ uint32_t val(long long i) const { return (uint32_t)i; }
};

// counts layout: [digit][block].  Every block histograms the same contiguous
// run of tiles the downsweep will give it.  DB: digit bits of the kernel
// (8 or 9); a digit of fewer bits uses the low buckets.
template <int DB, typename Src>
__device__ __forceinline__ void upsweep_body(const Src& src, const SortPlan& pl, int spec,
                                             uint32_t* __restrict__ counts,
                                             uint32_t (*hist)[1 << DB]) {
  using KeyT = decltype(src.key(0));
  constexpr int D = 1 << DB, DPT = D / kSortThreads;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, b = blockIdx.x;
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    hist[0][tid + q * kSortThreads] = 0;
    hist[1][tid + q * kSortThreads] = 0;
    hist[2][tid + q * kSortThreads] = 0;
    hist[3][tid + q * kSortThreads] = 0;
  }
  __syncthreads();
  const int t_begin = b * pl.tiles_per_block;
  int t_end = t_begin + pl.tiles_per_block;
  if (t_end > pl.ntiles) t_end = pl.ntiles;
  // The next tile's keys are loaded before this tile's counts are added, so
  // the loads of consecutive tiles overlap.
  KeyT k[kSortItems];
  int tile_n = 0;
  if (t_begin < t_end) {
    const long long tile0 = (long long)t_begin * kSortTile;
    tile_n = (pl.n - tile0 < kSortTile) ? (int)(pl.n - tile0) : kSortTile;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      const int i = tile_item(wave, r, lane);
      k[r] = (i < tile_n) ? src.key(tile0 + i) : (KeyT)0;
    }
  }
  for (int t = t_begin; t < t_end; ++t) {
    int d[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r)
      d[r] = (tile_item(wave, r, lane) < tile_n) ? digit_of((uint64_t)k[r], spec) : -1;
    if (t + 1 < t_end) {
      const long long ntile0 = (long long)(t + 1) * kSortTile;
      tile_n = (pl.n - ntile0 < kSortTile) ? (int)(pl.n - ntile0) : kSortTile;
#pragma unroll
      for (int r = 0; r < kSortItems; ++r) {
        const int i = tile_item(wave, r, lane);
        k[r] = (i < tile_n) ? src.key(ntile0 + i) : (KeyT)0;
      }
    }
#pragma unroll
    for (int r = 0; r < kSortItems; ++r)
      if (d[r] >= 0) atomicAdd(&hist[wave][d[r]], 1u);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const int dd = tid + q * kSortThreads;
    counts[(size_t)dd * kSortBlocks + b] = hist[0][dd] + hist[1][dd] + hist[2][dd] + hist[3][dd];
  }
}

template <int DB, typename KeyT>
__global__ __launch_bounds__(kSortThreads) void ps_sort_upsweep(const KeyT* __restrict__ keys,
                                                                const int* __restrict__ d_n,
                                                                int n_host, int max_pairs,
                                                                int spec,
                                                                uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[4][1 << DB];
  upsweep_body<DB>(ArraySrc<KeyT>{keys, nullptr}, sort_plan(d_n, n_host, max_pairs), spec, counts,
                   hist);
}

// Buffers a sort clears in passing (see ZeroReq).
struct ZeroReq {
  int2* ranges = nullptr;  // `n` (0, 0) ranges
  int n = 0;
  int* words = nullptr;  // `nwords` (<= 256) counter words
  int nwords = 0;
};

// One workgroup per digit: exclusive scan of counts[d][0..kSortBlocks) into
// offsets[d][...], digit total into totals[d].  The last digit pass's scan
// also clears the buffers of `z` (the tile ranges the boundary scan after
// the sort fills, and the segment lists' counters), in place of a launch of
// its own.
__global__ __launch_bounds__(kSortThreads) void ps_sort_scan(const uint32_t* __restrict__ counts,
                                                             uint32_t* __restrict__ offsets,
                                                             uint32_t* __restrict__ totals,
                                                             ZeroReq z) {
  __shared__ int scratch[4];
  if (z.ranges)
    for (int i = blockIdx.x * kSortThreads + threadIdx.x; i < z.n; i += gridDim.x * kSortThreads)
      z.ranges[i] = make_int2(0, 0);
  if (z.words && blockIdx.x == 0 && (int)threadIdx.x < z.nwords) z.words[threadIdx.x] = 0;
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

// The same set of lanes through LDS: every valid lane ORs its bit into its
// digit's 64-bit lane mask, reads the mask back and clears it.  Three LDS
// operations instead of eight ballots with their per-lane selects; lanes
// that share a digit contend for one word, which the LDS serialises.  `mt`
// holds two words per digit (lanes 0..31, 32..63), all zero on entry, and is
// left zero.
__device__ __forceinline__ uint64_t match_digit_lds(int d, bool ok, int lane, uint32_t* mt) {
  if (ok) atomicOr(&mt[2 * d + (lane >> 5)], 1u << (lane & 31));
  __builtin_amdgcn_wave_barrier();
  uint64_t peers = 0;
  if (ok) {
    const uint2 m = *reinterpret_cast<const uint2*>(&mt[2 * d]);
    peers = ((uint64_t)m.y << 32) | m.x;
  }
  __builtin_amdgcn_wave_barrier();
  if (ok) *reinterpret_cast<uint2*>(&mt[2 * d]) = make_uint2(0u, 0u);
  __builtin_amdgcn_wave_barrier();
  return peers;
}

template <typename Src, typename KeyT>
__device__ __forceinline__ void load_tile(const Src& src, long long tile0, int tile_n, int wave,
                                          int lane, KeyT (&k)[kSortItems],
                                          uint32_t (&v)[kSortItems]) {
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int i = tile_item(wave, r, lane);
    if (i < tile_n) {
      k[r] = src.key(tile0 + i);
      v[r] = src.val(tile0 + i);
    } else {
      k[r] = 0;
      v[r] = 0;
    }
  }
}

// LDS of one downsweep workgroup, declared by the kernel.
template <int DB, typename KeyT>
struct DownLds {
  KeyT skey[kSortTile];
  uint32_t sval[kSortTile];
  // per-wave digit counters while ranking, then (same slots) the tile-local
  // start of each (wave, digit) run; wcnt[0][d] is digit d's tile-local
  // start.  Thread t owns digits t*DPT .. t*DPT + DPT - 1.
  uint32_t wcnt[4][1 << DB];
  uint32_t gbase[1 << DB];
  uint2 s_match[4][1 << DB];  // per-wave digit lane masks (match_digit_lds)
  int scratch[4];
};

// Rank, stage and scatter.  Per tile: each wave ranks its 960 items against
// its own LDS digit counters (a wave's LDS operations execute in order, so no
// barrier is needed between rounds); one block scan turns the per-wave counts
// into tile-local starts; items are staged in LDS grouped by digit and each
// digit run is written contiguously.  The next tile's keys are loaded into
// registers while the current one is scanned, staged and written (about 210
// VGPRs with 64-bit keys: two workgroups per CU, no spills).
template <int DB, typename Src, typename KeyT>
__device__ __forceinline__ void downsweep_body(const Src& src, KeyT* __restrict__ keys_out,
                                               uint32_t* __restrict__ vals_out, const SortPlan& pl,
                                               int spec, const uint32_t* __restrict__ offsets,
                                               const uint32_t* __restrict__ totals,
                                               DownLds<DB, KeyT>& L) {
  constexpr int D = 1 << DB, DPT = D / kSortThreads;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, b = blockIdx.x;
  const int t_begin = b * pl.tiles_per_block;
  int t_end = t_begin + pl.tiles_per_block;
  if (t_end > pl.ntiles) t_end = pl.ntiles;
  if (t_begin >= t_end) return;  // uniform across the block

  {
    uint32_t tv[DPT];
    int local = 0;
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      tv[q] = totals[tid * DPT + q];
      local += (int)tv[q];
    }
    int dtotal;
    uint32_t run = (uint32_t)block256_excl_scan(local, L.scratch, dtotal);
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const int dd = tid * DPT + q;
      L.gbase[dd] = run + offsets[(size_t)dd * kSortBlocks + b];
      run += tv[q];
      L.wcnt[0][dd] = 0;
      L.wcnt[1][dd] = 0;
      L.wcnt[2][dd] = 0;
      L.wcnt[3][dd] = 0;
    }
  }
  const uint64_t below = lanes_below(lane);
  uint32_t* const mt = reinterpret_cast<uint32_t*>(L.s_match[wave]);
  for (int i = lane; i < 2 * D; i += kWave) mt[i] = 0u;
  __builtin_amdgcn_wave_barrier();

  KeyT k[kSortItems];
  uint32_t v[kSortItems];
  long long tile0 = (long long)t_begin * kSortTile;
  int tile_n = (pl.n - tile0 < kSortTile) ? (int)(pl.n - tile0) : kSortTile;
  load_tile(src, tile0, tile_n, wave, lane, k, v);
  __syncthreads();

  for (int t = t_begin; t < t_end; ++t) {
    int dg[kSortItems];
    uint32_t pos[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      const int d = (tile_item(wave, r, lane) < tile_n) ? digit_of((uint64_t)k[r], spec) : -1;
      dg[r] = d;
      const uint64_t peers = match_digit_lds(d >= 0 ? d : 0, d >= 0, lane, mt);
      const int rank = popc64(peers & below);
      uint32_t base = 0;
      if (d >= 0) base = L.wcnt[wave][d];
      __builtin_amdgcn_wave_barrier();
      if (d >= 0 && rank == 0) L.wcnt[wave][d] = base + (uint32_t)popc64(peers);
      __builtin_amdgcn_wave_barrier();
      pos[r] = base + (uint32_t)rank;
    }
    // Prefetch the next tile while this one is scanned, staged and written.
    KeyT kn[kSortItems];
    uint32_t vn[kSortItems];
    const long long ntile0 = tile0 + kSortTile;
    const int ntile_n =
        (t + 1 < t_end) ? ((pl.n - ntile0 < kSortTile) ? (int)(pl.n - ntile0) : kSortTile) : 0;
    if (ntile_n > 0) load_tile(src, ntile0, ntile_n, wave, lane, kn, vn);
    __syncthreads();
    uint32_t c0[DPT], c1[DPT], c2[DPT], c3[DPT], tot[DPT];
    int local = 0;
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const int dd = tid * DPT + q;
      c0[q] = L.wcnt[0][dd];
      c1[q] = L.wcnt[1][dd];
      c2[q] = L.wcnt[2][dd];
      c3[q] = L.wcnt[3][dd];
      tot[q] = c0[q] + c1[q] + c2[q] + c3[q];
      local += (int)tot[q];
    }
    int tile_total;
    uint32_t ls = (uint32_t)block256_excl_scan(local, L.scratch, tile_total);
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const int dd = tid * DPT + q;
      L.wcnt[0][dd] = ls;
      L.wcnt[1][dd] = ls + c0[q];
      L.wcnt[2][dd] = ls + c0[q] + c1[q];
      L.wcnt[3][dd] = ls + c0[q] + c1[q] + c2[q];
      ls += tot[q];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      if (dg[r] >= 0) {
        const uint32_t p = L.wcnt[wave][dg[r]] + pos[r];
        L.skey[p] = k[r];
        L.sval[p] = v[r];
      }
    }
    __syncthreads();
    for (int i = tid; i < tile_n; i += kSortThreads) {
      const KeyT key = L.skey[i];
      const int d = digit_of((uint64_t)key, spec);
      const uint32_t dst = L.gbase[d] + ((uint32_t)i - L.wcnt[0][d]);
      if (dst < (uint32_t)pl.n) {
        keys_out[dst] = key;
        vals_out[dst] = L.sval[i];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const int dd = tid * DPT + q;
      L.gbase[dd] += tot[q];
      L.wcnt[0][dd] = 0;
      L.wcnt[1][dd] = 0;
      L.wcnt[2][dd] = 0;
      L.wcnt[3][dd] = 0;
    }
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

template <int DB, typename KeyT>
__global__ __launch_bounds__(kSortThreads) void ps_sort_downsweep(
    const KeyT* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    KeyT* __restrict__ keys_out, uint32_t* __restrict__ vals_out, const int* __restrict__ d_n,
    int n_host, int max_pairs, int spec, const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ totals) {
  __shared__ DownLds<DB, KeyT> L;
  downsweep_body<DB>(ArraySrc<KeyT>{keys_in, vals_in}, keys_out, vals_out,
                     sort_plan(d_n, n_host, max_pairs), spec, offsets, totals, L);
}

// ------------------------------------------------------------ depth sort
// Stable sort of the Gaussians by the 32-bit words of their depths (the
// depth-first pipeline): 8-bit digit passes, of which only the digits that
// differ among the emitting Gaussians run.  The count pass leaves each
// Gaussian's depth word (0 when it emits nothing) and, in kPlanSlots slots,
// the OR and the AND of the emitting depths; every kernel of the sort
// derives the same plan from them: varying = OR ^ AND, digit q runs when its
// byte of `varying` is non-zero, reading what the previous running digit
// wrote (or, the first, the depth words with ids as values, a Gaussian that
// emits nothing taking the AND as its key) into buffer (index among running
// digits) % 2.  No digit runs when no Gaussian emits or all emitting
// depths are equal: the sorted order is then the identity.

// Reads the plan (every thread gets it through LDS).
__device__ __forceinline__ void read_depth_plan(const uint32_t* __restrict__ plan,
                                                uint32_t* s_plan) {
  const int tid = threadIdx.x;
  if (tid < kWave) {
    uint32_t o = plan[tid * kPlanStride], a = plan[(kPlanSlots + tid) * kPlanStride];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      o |= (uint32_t)__shfl_xor((int)o, d);
      a &= (uint32_t)__shfl_xor((int)a, d);
    }
    if (tid == 0) {
      s_plan[0] = o == 0u ? 0u : (o ^ a);  // varying bits (none without emitters)
      s_plan[1] = a;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ bool digit_runs(uint32_t varying, int q) {
  return ((varying >> (8 * q)) & 0xffu) != 0u;
}

__device__ __forceinline__ int digits_before(uint32_t varying, int q) {
  int k = 0;
  for (int i = 0; i < q; ++i) k += digit_runs(varying, i) ? 1 : 0;
  return k;
}

__global__ __launch_bounds__(kSortThreads) void ps_dsort_upsweep(DepthSortArgs A, int q,
                                                                 uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[4][256];
  __shared__ uint32_t s_plan[2];
  read_depth_plan(A.plan, s_plan);
  const uint32_t varying = s_plan[0];
  if (!digit_runs(varying, q)) return;  // uniform
  const int k = digits_before(varying, q);
  const SortPlan pl = sort_plan(nullptr, A.n, A.n);
  const int spec = digit_spec(8 * q, 8);
  if (k == 0)
    upsweep_body<8>(DepthSrc{A.words, s_plan[1]}, pl, spec, counts, hist);
  else
    upsweep_body<8>(ArraySrc<uint32_t>{A.kb[(k - 1) & 1], A.vb[(k - 1) & 1]}, pl, spec, counts,
                    hist);
}

__global__ __launch_bounds__(kSortThreads) void ps_dsort_downsweep(
    DepthSortArgs A, int q, const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ totals) {
  __shared__ DownLds<8, uint32_t> L;
  __shared__ uint32_t s_plan[2];
  read_depth_plan(A.plan, s_plan);
  const uint32_t varying = s_plan[0];
  if (!digit_runs(varying, q)) return;  // uniform
  const int k = digits_before(varying, q);
  const SortPlan pl = sort_plan(nullptr, A.n, A.n);
  const int spec = digit_spec(8 * q, 8);
  if (k == 0)
    downsweep_body<8>(DepthSrc{A.words, s_plan[1]}, A.kb[0], A.vb[0], pl, spec, offsets, totals,
                      L);
  else
    downsweep_body<8>(ArraySrc<uint32_t>{A.kb[(k - 1) & 1], A.vb[(k - 1) & 1]}, A.kb[k & 1],
                      A.vb[k & 1], pl, spec, offsets, totals, L);
}

// ------------------------------------------------------------ segment pass
// After the digit passes over the tile bits [32, end_bit), pairs are grouped
// by tile (segment) and keep their input order inside each segment; what is
// left is a stable sort of every segment by the low 32 bits.  That needs no
// global passes for segments of up to 32768 pairs: a wave (up to 896
// pairs), or a workgroup of 4, 8 or 16 waves (4096, 8192, 32768 pairs)
// sorts each in registers and LDS; longer segments are listed and a
// workgroup sorts each with four counting passes through global memory.  The
// result equals the full LSD sort bit for bit: sorting by (high, low) stably
// is sorting by the whole key stably.

constexpr int kSegCap = 896;    // pairs per wave-sorted segment
constexpr int kSegWaves = 4;    // waves per workgroup of the small-segment kernel

// Lanes of the wave whose 8-bit digit equals this lane's (valid lanes only).
__device__ __forceinline__ uint64_t match_digit(int d, bool ok) {
  uint64_t peers = ballot(ok);
#pragma unroll
  for (int bit = 0; bit < 8; ++bit) {
    const bool set = ok && ((d >> bit) & 1);
    const uint64_t bb = ballot(set);
    peers &= set ? bb : ~bb;
  }
  return peers;
}

// Each wave takes segments seg = wave id, wave id + total waves, ...  A
// segment longer than kSegCap, or whose keys' high words differ (bits above
// the sorted ones), goes to the list for the workgroup kernels.  The wave keeps its
// segment in registers (item j in round j / 64 of lane j % 64) and runs one
// stable counting pass per 8-bit digit of the low word: rank every item
// among the segment's items with the same digit (ballot match per round,
// per-digit counters in LDS), scatter the (low word, value) pairs through
// LDS and read them back in the new order.  Digits that are the same for
// every item of the segment are skipped (a stable pass over a constant digit
// changes nothing); for depths the top byte usually is.  Sorted in place.
__global__ __launch_bounds__(kSegWaves * kWave) void ps_seg_small(
    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals, const int2* __restrict__ ranges,
    int segs, int* __restrict__ big_list, int* __restrict__ big_count, int big_cap) {
  constexpr int R = kSegCap / kWave;  // rounds of 64 items
  __shared__ uint32_t s_lo[kSegWaves][kSegCap];
  __shared__ uint32_t s_v[kSegWaves][kSegCap];
  __shared__ uint32_t cnt[kSegWaves][kRadix];
  __shared__ uint2 s_match[kSegWaves][kRadix];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t below = lanes_below(lane);
  uint32_t* const slo = s_lo[wave];
  uint32_t* const sv = s_v[wave];
  uint32_t* const c = cnt[wave];
  uint32_t* const mt = reinterpret_cast<uint32_t*>(s_match[wave]);
  for (int i = lane; i < 2 * kRadix; i += kWave) mt[i] = 0u;
  __builtin_amdgcn_wave_barrier();
  for (int seg = blockIdx.x * kSegWaves + wave; seg < segs; seg += gridDim.x * kSegWaves) {
    const int2 rg = ranges[seg];
    const int n = rg.y - rg.x;
    if (n <= 1) continue;
    bool big = n > kSegCap;
    uint32_t hi0 = 0;
    uint32_t lo[R], v[R];
    uint32_t diff = 0;
    if (!big) {
      const uint64_t k0 = keys[rg.x];
      hi0 = (uint32_t)(k0 >> 32);
      bool same = true;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = r * kWave + lane;
        lo[r] = 0;
        v[r] = 0;
        if (j < n) {
          const uint64_t k = keys[rg.x + j];
          same = same && (uint32_t)(k >> 32) == hi0;
          lo[r] = (uint32_t)k;
          v[r] = vals[rg.x + j];
          diff |= lo[r] ^ (uint32_t)k0;
        }
      }
      big = ballot(!same) != 0;
    }
    if (big) {
      if (lane == 0) {
        const int at = atomicAdd(big_count, 1);
        if (at < big_cap) big_list[at] = seg;
      }
      continue;
    }
    const int rounds = (n + kWave - 1) / kWave;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) diff |= (uint32_t)__shfl_xor((int)diff, s);
    for (int p = 0; p < 4; ++p) {
      const int shift = 8 * p;
      if (((diff >> shift) & 0xffu) == 0) continue;  // uniform: constant digit
      for (int i = lane; i < kRadix; i += kWave) c[i] = 0;
      __builtin_amdgcn_wave_barrier();
      uint64_t pm[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (r < rounds) {
          const bool ok = r * kWave + lane < n;
          const int d = (int)((lo[r] >> shift) & 0xffu);
          pm[r] = match_digit_lds(d, ok, lane, mt);
          if (ok && (pm[r] & below) == 0) c[d] += (uint32_t)popc64(pm[r]);
          __builtin_amdgcn_wave_barrier();
        }
      }
      const uint32_t c0 = c[4 * lane], c1 = c[4 * lane + 1], c2 = c[4 * lane + 2],
                     c3 = c[4 * lane + 3];
      const int sum = (int)(c0 + c1 + c2 + c3);
      const uint32_t ex = (uint32_t)(wave_incl_scan(sum, lane) - sum);
      __builtin_amdgcn_wave_barrier();
      c[4 * lane] = ex;
      c[4 * lane + 1] = ex + c0;
      c[4 * lane + 2] = ex + c0 + c1;
      c[4 * lane + 3] = ex + c0 + c1 + c2;
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (r < rounds) {
          const bool ok = r * kWave + lane < n;
          const int d = (int)((lo[r] >> shift) & 0xffu);
          const uint32_t rank = (uint32_t)popc64(pm[r] & below);
          const uint32_t base = ok ? c[d] : 0u;
          __builtin_amdgcn_wave_barrier();
          if (ok && rank == 0) c[d] = base + (uint32_t)popc64(pm[r]);
          __builtin_amdgcn_wave_barrier();
          if (ok) {
            slo[base + rank] = lo[r];
            sv[base + rank] = v[r];
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = r * kWave + lane;
        if (r < rounds && j < n) {
          lo[r] = slo[j];
          v[r] = sv[j];
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    const uint64_t hi = (uint64_t)hi0 << 32;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = r * kWave + lane;
      if (j < n) {
        keys[rg.x + j] = hi | lo[r];
        vals[rg.x + j] = v[r];
      }
    }
  }
}

// Segments whose keys agree above the sorted bits, sorted by one workgroup
// of W waves each (ps_seg_group<W, RD>: up to W * RD * 64 pairs).
// Wave w takes the contiguous chunk of `rounds` rounds of 64 items starting at
// item w * rounds * 64 (rounds = n / (64 W) rounded up); items are held
// compactly in registers: the low key word, and the item's index in the
// segment (15 bits) packed with its destination in the current pass (15
// bits).  The wave's last rounds are padded up to whole rounds with items
// whose low word is the segment's largest; a stable sort places them after
// every real item (their indices are larger), so the passes need no per-item
// masks.  One stable counting pass per non-constant 8-bit digit of the low
// word: each wave ranks its items against its own LDS digit counters (ballot
// match), a block scan over (digit, wave) turns the counts into starts, and
// the low words are scattered through one LDS buffer and read back in the
// new order, then the indices the same way.  At the end the values are
// gathered by index through LDS (every read of the segment's old values
// completes before the first write) and written with the sorted keys.
// Longer (or mixed) segments go on to the next list: 4 waves x 16 rounds
// (4096 pairs), 8 waves x 24 rounds (12288), then 16 waves x 32 rounds
// (32768, most of a CU's LDS), then ps_seg_big.  The 8-wave route's 24
// rounds (two workgroups per CU) keep the runs just past 8192 pairs out of
// the one-per-CU 16-wave route (30M Gaussians at 1920x1080, 2828 runs of
// 8193-9131 pairs: whole sort 1959 -> 1357 us; 32 rounds spill and were
// slower).  Fewer waves per segment mean fewer waves
// waiting at each barrier and more segments in flight per CU (4 waves
// instead of 8 for segments of 897..4096 pairs: 893 -> 563 us for a
// 30M-Gaussian frame at 3840x2160); a 2-wave route for up to 2048 pairs
// was measured slower (one more launch, no faster per segment).
constexpr int kGroupRounds = 16;  // 4 waves: 4096 pairs
constexpr int kMidRounds = 24;    // 8 waves: 12288 pairs
constexpr int kLargeRounds = 32;  // 16 waves: 32768 pairs

// Appends a segment to the next kernel's list (dropped past its capacity,
// which only inputs with more segments than the workspace is sized for reach).
__device__ __forceinline__ void pass_on(int* list, int* count, int cap, int seg) {
  const int at = atomicAdd(count, 1);
  if (at < cap) list[at] = seg;
}

// WPS: waves per SIMD the kernel is compiled for (its register budget).
template <int W, int RD, int WPS>
__global__ __launch_bounds__(W * kWave, WPS) void ps_seg_group(
    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals, const int2* __restrict__ ranges,
    const int* __restrict__ big_list, const int* __restrict__ big_count, int total,
    int* __restrict__ next_list, int* __restrict__ next_count, int next_cap) {
  constexpr int kThreads = W * kWave;
  constexpr int DPT = kThreads >= kRadix ? 1 : kRadix / kThreads;  // digits per owner
  constexpr int kOwners = kRadix / DPT;
  static_assert(kOwners * DPT == kRadix && kOwners <= kThreads, "digit owners");
  constexpr int kCap = kThreads * RD;
  // The digit match's lane-mask tables (match_digit_lds) live in the
  // exchange buffer, which holds nothing while a pass ranks its items.
  static_assert(kCap >= W * 2 * kRadix, "lane-mask tables fit in the exchange buffer");
  __shared__ uint32_t s_buf[kCap];
  __shared__ uint32_t wcnt[W][kRadix];
  __shared__ int s_wsum[W];
  __shared__ uint32_t s_diff, s_max;
  __shared__ int s_mixed;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t below = lanes_below(lane);
  uint32_t* const mt = s_buf + wave * 2 * kRadix;
  const int nseg = min(*big_count, total);
  for (int b = blockIdx.x; b < nseg; b += gridDim.x) {
    const int seg = big_list[b];
    const int2 rg = ranges[seg];
    const int n = rg.y - rg.x;
    if (n > kCap) {  // uniform
      if (tid == 0) pass_on(next_list, next_count, next_cap, seg);
      continue;
    }
    const uint64_t k0 = keys[rg.x];
    const uint32_t hi0 = (uint32_t)(k0 >> 32);
    if (tid == 0) {
      s_diff = 0u;
      s_max = 0u;
      s_mixed = 0;
    }
    __syncthreads();
    bool same = true;
    uint32_t diff = 0, mx = 0;
    // Strided copies go kBatch items per thread at a time, so that that many
    // global accesses are in flight.
    constexpr int kBatch = 8;
    for (int i0 = tid; i0 < n; i0 += kBatch * kThreads) {
      uint64_t k[kBatch];
#pragma unroll
      for (int q = 0; q < kBatch; ++q) {
        const int i = i0 + q * kThreads;
        k[q] = i < n ? keys[rg.x + i] : k0;
      }
#pragma unroll
      for (int q = 0; q < kBatch; ++q) {
        const int i = i0 + q * kThreads;
        same = same && (uint32_t)(k[q] >> 32) == hi0;
        diff |= (uint32_t)k[q] ^ (uint32_t)k0;
        mx = max(mx, (uint32_t)k[q]);
        if (i < n) s_buf[i] = (uint32_t)k[q];
      }
    }
    if (!same) s_mixed = 1;
    if (diff) atomicOr(&s_diff, diff);
    if (mx) atomicMax(&s_max, mx);
    __syncthreads();
    const bool mixed = s_mixed != 0;
    const uint32_t all_diff = s_diff, pad = s_max;
    __syncthreads();  // s_mixed / s_diff / s_max are rewritten for the next segment
    if (mixed) {  // uniform
      if (tid == 0) pass_on(next_list, next_count, next_cap, seg);
      continue;
    }
    const int rounds = (n + kThreads - 1) / kThreads;  // uniform
    // Item r of this lane is item mine + 64 r of the segment.
    const int mine = wave * rounds * kWave + lane;
    const int lim = n - mine;  // round r holds a real item when 64 r < lim
    uint32_t* const my_buf = s_buf + mine;
    // lo: low key word; pk: index in the segment (bits 0..14), destination of
    // the current pass (bits 16..30).
    uint32_t lo[RD], pk[RD];
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      pk[r] = (uint32_t)(mine + r * kWave);
      lo[r] = r < rounds && r * kWave < lim ? my_buf[r * kWave] : pad;
    }
    __syncthreads();
    for (int p = 0; p < 4; ++p) {
      const int shift = 8 * p;
      if (((all_diff >> shift) & 0xffu) == 0) continue;  // uniform: constant digit
      for (int i = lane; i < kRadix; i += kWave) wcnt[wave][i] = 0;
      for (int i = lane; i < 2 * kRadix; i += kWave) mt[i] = 0u;
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int r = 0; r < RD; ++r) {
        if (r >= rounds) continue;  // uniform
        const int d = (int)((lo[r] >> shift) & 0xffu);
        const uint64_t peers = match_digit_lds(d, true, lane, mt);
        const uint32_t rank = (uint32_t)popc64(peers & below);
        const uint32_t base = wcnt[wave][d];
        __builtin_amdgcn_wave_barrier();
        if (rank == 0) wcnt[wave][d] = base + (uint32_t)popc64(peers);
        __builtin_amdgcn_wave_barrier();
        pk[r] = (pk[r] & 0x7fffu) | ((base + rank) << 16);
      }
      __syncthreads();
      // Digit starts over (digit, wave): thread t < kOwners owns digits
      // DPT t .. DPT t + DPT - 1.
      uint32_t c[DPT][W];
      uint32_t tot = 0;
      if (tid < kOwners) {
#pragma unroll
        for (int k = 0; k < DPT; ++k)
#pragma unroll
          for (int w = 0; w < W; ++w) {
            c[k][w] = wcnt[w][DPT * tid + k];
            tot += c[k][w];
          }
      }
      const int incl = wave_incl_scan((int)tot, lane);
      if (lane == 63) s_wsum[wave] = incl;
      __syncthreads();
      if (tid < kOwners) {
        int before = 0;
        for (int w = 0; w < wave; ++w) before += s_wsum[w];
        uint32_t start = (uint32_t)(before + incl) - tot;
#pragma unroll
        for (int k = 0; k < DPT; ++k)
#pragma unroll
          for (int w = 0; w < W; ++w) {
            wcnt[w][DPT * tid + k] = start;
            start += c[k][w];
          }
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RD; ++r) {
        if (r >= rounds) continue;  // uniform
        const uint32_t at = wcnt[wave][(lo[r] >> shift) & 0xffu] + (pk[r] >> 16);
        pk[r] = (pk[r] & 0x7fffu) | (at << 16);
        s_buf[at] = lo[r];
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RD; ++r)
        if (r < rounds) lo[r] = my_buf[r * kWave];
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RD; ++r)
        if (r < rounds) s_buf[pk[r] >> 16] = pk[r] & 0x7fffu;
      __syncthreads();
#pragma unroll
      for (int r = 0; r < RD; ++r)
        if (r < rounds) pk[r] = my_buf[r * kWave];
      __syncthreads();
    }
    const uint64_t hi = (uint64_t)hi0 << 32;
    uint64_t* const my_keys = keys + rg.x + mine;
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      if (r < rounds && r * kWave < lim) {
        my_keys[r * kWave] = hi | lo[r];
        my_buf[r * kWave] = pk[r] & 0x7fffu;
      }
    }
    __syncthreads();
    for (int i0 = tid; i0 < n; i0 += kBatch * kThreads) {
      uint32_t g[kBatch];
#pragma unroll
      for (int q = 0; q < kBatch; ++q) {
        const int i = i0 + q * kThreads;
        g[q] = i < n ? vals[rg.x + (int)s_buf[i]] : 0u;
      }
#pragma unroll
      for (int q = 0; q < kBatch; ++q) {
        const int i = i0 + q * kThreads;
        if (i < n) s_buf[i] = g[q];
      }
    }
    __syncthreads();
    for (int i = tid; i < n; i += kThreads) vals[rg.x + i] = s_buf[i];
    __syncthreads();  // s_buf is rewritten for the next segment
  }
}

// Listed segments: a workgroup per segment, four stable counting passes by
// the low 32 bits through global memory (keys <-> tmp, ending in keys).
// Chunks of 1024 pairs: wave w ranks items [256w, 256w + 256) of the chunk
// against its own LDS counters, then the per-wave counts are combined.
__global__ __launch_bounds__(kSortThreads) void ps_seg_big(
    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals, uint64_t* __restrict__ tkeys,
    uint32_t* __restrict__ tvals, const int2* __restrict__ ranges,
    const int* __restrict__ big_list, const int* __restrict__ big_count, int big_cap) {
  constexpr int kPer = 4;  // rounds of 64 per wave and chunk
  constexpr int kChunk = kSortThreads * kPer;
  __shared__ uint32_t wcnt[4][kRadix];
  __shared__ uint32_t base[kRadix];
  __shared__ int scratch[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t below = lanes_below(lane);
  const int nbig = min(*big_count, big_cap);
  for (int b = blockIdx.x; b < nbig; b += gridDim.x) {
    const int2 r = ranges[big_list[b]];
    const int n = r.y - r.x;
    for (int p = 0; p < 4; ++p) {
      const int spec = digit_spec(8 * p, 8);
      const uint64_t* sk = ((p & 1) ? tkeys : keys) + r.x;
      const uint32_t* sv = ((p & 1) ? tvals : vals) + r.x;
      uint64_t* dk = ((p & 1) ? keys : tkeys) + r.x;
      uint32_t* dv = ((p & 1) ? vals : tvals) + r.x;
      // digit histogram
      wcnt[0][tid] = 0;
      wcnt[1][tid] = 0;
      wcnt[2][tid] = 0;
      wcnt[3][tid] = 0;
      __syncthreads();
      for (int i = tid; i < n; i += kSortThreads)
        atomicAdd(&wcnt[wave][digit_of(sk[i], spec)], 1u);
      __syncthreads();
      int total;
      const int ex = block256_excl_scan(
          (int)(wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid]), scratch, total);
      base[tid] = (uint32_t)ex;
      wcnt[0][tid] = 0;
      wcnt[1][tid] = 0;
      wcnt[2][tid] = 0;
      wcnt[3][tid] = 0;
      __syncthreads();
      // Chunk c + 1 is loaded while chunk c is ranked and written.
      uint64_t k[kPer];
      uint32_t v[kPer];
#pragma unroll
      for (int q = 0; q < kPer; ++q) {
        const int i = wave * (kWave * kPer) + q * kWave + lane;
        k[q] = i < n ? sk[i] : 0ull;
        v[q] = i < n ? sv[i] : 0u;
      }
      for (int c0 = 0; c0 < n; c0 += kChunk) {
        uint32_t pos[kPer];
        int dg[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
          const int i = c0 + wave * (kWave * kPer) + q * kWave + lane;
          dg[q] = i < n ? digit_of(k[q], spec) : -1;
        }
        uint64_t kn[kPer];
        uint32_t vn[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
          const int i = c0 + kChunk + wave * (kWave * kPer) + q * kWave + lane;
          kn[q] = i < n ? sk[i] : 0ull;
          vn[q] = i < n ? sv[i] : 0u;
        }
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
          const bool ok = dg[q] >= 0;
          const int d = ok ? dg[q] : 0;
          const uint64_t peers = match_digit(d, ok);
          const uint32_t rank = (uint32_t)popc64(peers & below);
          const uint32_t bw = ok ? wcnt[wave][d] : 0u;
          __builtin_amdgcn_wave_barrier();
          if (ok && rank == 0) wcnt[wave][d] = bw + (uint32_t)popc64(peers);
          __builtin_amdgcn_wave_barrier();
          pos[q] = bw + rank;
        }
        __syncthreads();
        // per digit: counts of the four waves -> offsets of each wave's run
        const uint32_t w0 = wcnt[0][tid], w1 = wcnt[1][tid], w2 = wcnt[2][tid], w3 = wcnt[3][tid];
        const uint32_t b0 = base[tid];
        __syncthreads();
        wcnt[0][tid] = b0;
        wcnt[1][tid] = b0 + w0;
        wcnt[2][tid] = b0 + w0 + w1;
        wcnt[3][tid] = b0 + w0 + w1 + w2;
        base[tid] = b0 + w0 + w1 + w2 + w3;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
          if (dg[q] >= 0) {
            const uint32_t dst = wcnt[wave][dg[q]] + pos[q];
            dk[dst] = k[q];
            dv[dst] = v[q];
          }
        }
        __syncthreads();
        wcnt[0][tid] = 0;
        wcnt[1][tid] = 0;
        wcnt[2][tid] = 0;
        wcnt[3][tid] = 0;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
          k[q] = kn[q];
          v[q] = vn[q];
        }
        __syncthreads();
      }
      // the pass's writes must be visible to the next pass's reads
      __threadfence_block();
      __syncthreads();
    }
  }
}

struct SortWs {
  uint64_t* alt_keys;
  uint32_t* alt_vals;
  uint32_t* counts;
  uint32_t* offsets;
  uint32_t* totals;
  int2* seg_ranges;  // segment ranges when the caller gives none
  int* big_list;     // segments longer than kSegCap (or mixed high words)
  int* big_count;    // entries of big_list, big_list2, big_list3, big_list (second use)
  int* big_list2;    // segments for the 512-thread ps_seg_group
  int* big_list3;    // segments for the 1024-thread ps_seg_group
};

// Segments (values of the sorted high bits) the internal range table holds.
constexpr int kSegMaxSegs = 1 << 18;

// Entries the list of segments for ps_seg_big can take: every segment longer
// than kSegCap (at most pairs / (kSegCap + 1)), plus every internal segment
// whose keys differ above the sorted bits.
inline size_t big_list_cap(size_t pairs) { return pairs / (kSegCap + 1) + 1 + kSegMaxSegs; }

inline size_t sort_layout(int cap, char* base, SortWs* w) {
  size_t off = 0;
  const size_t c = (size_t)(cap > 0 ? cap : 1);
  if (w) w->alt_keys = reinterpret_cast<uint64_t*>(base + off);
  off += align_up(c * sizeof(uint64_t), 256);
  if (w) w->alt_vals = reinterpret_cast<uint32_t*>(base + off);
  off += align_up(c * sizeof(uint32_t), 256);
  if (w) w->counts = reinterpret_cast<uint32_t*>(base + off);
  off += align_up((size_t)kMaxRadix * kSortBlocks * sizeof(uint32_t), 256);
  if (w) w->offsets = reinterpret_cast<uint32_t*>(base + off);
  off += align_up((size_t)kMaxRadix * kSortBlocks * sizeof(uint32_t), 256);
  if (w) w->totals = reinterpret_cast<uint32_t*>(base + off);
  off += align_up((size_t)kMaxRadix * sizeof(uint32_t), 256);
  if (w) w->seg_ranges = reinterpret_cast<int2*>(base + off);
  off += align_up((size_t)kSegMaxSegs * sizeof(int2), 256);
  if (w) w->big_count = reinterpret_cast<int*>(base + off);
  off += 256;
  if (w) w->big_list = reinterpret_cast<int*>(base + off);
  off += align_up(big_list_cap(c) * sizeof(int), 256);
  if (w) w->big_list2 = reinterpret_cast<int*>(base + off);
  off += align_up(big_list_cap(c) * sizeof(int), 256);
  if (w) w->big_list3 = reinterpret_cast<int*>(base + off);
  off += align_up(big_list_cap(c) * sizeof(int), 256);
  return off;
}

// Sort algorithm: full LSD over every key bit, or digit passes over the
// tile bits followed by the per-segment pass.  The segmented form is the
// default; POLYSPLAT_SORT=lsd selects the full LSD (both are exact).
inline bool use_segmented() {
  const char* e = getenv("POLYSPLAT_SORT");
  return !(e && e[0] == 'l');
}

template <int DB, typename KeyT>
void launch_digit_pass(const KeyT* kin, const uint32_t* vin, KeyT* kout, uint32_t* vout,
                       int n_host, const int* d_n, int max_pairs, int spec, const SortWs& w,
                       hipStream_t stream, const ZeroReq& z) {
  hipLaunchKernelGGL((ps_sort_upsweep<DB, KeyT>), dim3(kSortBlocks), dim3(kSortThreads), 0,
                     stream, kin, d_n, n_host, max_pairs, spec, w.counts);
  hipLaunchKernelGGL(ps_sort_scan, dim3(1 << DB), dim3(kSortThreads), 0, stream, w.counts,
                     w.offsets, w.totals, z);
  hipLaunchKernelGGL((ps_sort_downsweep<DB, KeyT>), dim3(kSortBlocks), dim3(kSortThreads), 0,
                     stream, kin, vin, kout, vout, d_n, n_host, max_pairs, spec, w.offsets,
                     w.totals);
}

// Stable LSD passes over key bits [bit0, bit1): 8-bit digits, or 9-bit ones
// when that needs fewer passes; the bits are spread evenly over the passes.
// `alt_keys`/`alt_vals` hold the pairs between passes; the last pass's scan
// clears what `z` names.
template <typename KeyT>
int digit_passes(int n_host, const int* d_n, int max_pairs, int bit0, int bit1, const SortWs& w,
                 const KeyT* keys_in, const uint32_t* vals_in, KeyT* keys_out,
                 uint32_t* vals_out, hipStream_t stream, const ZeroReq& z = ZeroReq()) {
  const int bits = bit1 - bit0;
  const int p8 = (bits + 7) / 8, p9 = (bits + kMaxDigitBits - 1) / kMaxDigitBits;
  const int passes = p9 < p8 ? p9 : p8;
  const KeyT* kin = keys_in;
  const uint32_t* vin = vals_in;
  KeyT* const alt_keys = reinterpret_cast<KeyT*>(w.alt_keys);
  int lo = bit0;
  for (int p = 0; p < passes; ++p) {
    // Ping-pong so that the last pass lands in the caller's output.
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    KeyT* kout = to_out ? keys_out : alt_keys;
    uint32_t* vout = to_out ? vals_out : w.alt_vals;
    // The narrower digit first (C3's 15 tile bits as 7 + 8): measured 1-2%
    // faster whole sorts than the wider one first on every odd split.
    const int width = (bit1 - lo) / (passes - p);  // <= 9
    const int spec = digit_spec(lo, width);
    const ZeroReq zp = (p == passes - 1) ? z : ZeroReq();
    if (width > 8)
      launch_digit_pass<9, KeyT>(kin, vin, kout, vout, n_host, d_n, max_pairs, spec, w, stream,
                                 zp);
    else
      launch_digit_pass<8, KeyT>(kin, vin, kout, vout, n_host, d_n, max_pairs, spec, w, stream,
                                 zp);
    PS_LAUNCH_CHECK();
    lo += width;
    kin = kout;
    vin = vout;
  }
  return PS_OK;
}

}  // namespace

int sort_launch(int n_host, const int* d_n, int max_pairs, int end_bit, void* space,
                size_t space_bytes, const uint64_t* keys_in, const uint32_t* vals_in,
                uint64_t* keys_out, uint32_t* vals_out, hipStream_t stream, int* tile_ranges,
                int tiles) {
  PS_REQUIRE(max_pairs >= 0, "max_pairs must be >= 0");
  PS_REQUIRE(d_n != nullptr || (n_host >= 0 && n_host <= max_pairs),
             "num_rendered %d outside [0, %d]", n_host, max_pairs);
  PS_REQUIRE(space != nullptr && space_bytes >= sort_layout(max_pairs, nullptr, nullptr),
             "sort space too small (%zu < %zu)", space_bytes,
             sort_layout(max_pairs, nullptr, nullptr));
  PS_REQUIRE(end_bit > 0 && end_bit <= 64, "bad end bit %d", end_bit);
  PS_REQUIRE(keys_in != keys_out && vals_in != vals_out, "sort outputs must not alias inputs");
  PS_REQUIRE(tile_ranges == nullptr || tiles > 0, "tile ranges need a tile count");
  SortWs w;
  sort_layout(max_pairs, static_cast<char*>(space), &w);
  const int seg_bits = end_bit - 32;
  const bool segmented = use_segmented() && seg_bits > 0 &&
                         (tile_ranges != nullptr || seg_bits <= 18);
  if (!segmented) {
    if (d_n == nullptr && n_host == 0) {
      if (tile_ranges) return tile_ranges_scan_launch(0, nullptr, 0, tiles, keys_out,
                                                      tile_ranges, stream);
      return PS_OK;
    }
    ZeroReq z;
    z.ranges = reinterpret_cast<int2*>(tile_ranges);
    z.n = tile_ranges ? tiles : 0;
    int st = digit_passes(n_host, d_n, max_pairs, 0, end_bit, w, keys_in, vals_in, keys_out,
                          vals_out, stream, z);
    if (st != PS_OK || tile_ranges == nullptr) return st;
    return tile_ranges_scan_launch(n_host, d_n, max_pairs, tiles, keys_out, tile_ranges, stream,
                                   0xffffffffu, nullptr, 0, /*cleared=*/true);
  }
  // Digit passes over the tile bits, then segment ranges, then the segment
  // pass in place in the output.
  int* ranges = tile_ranges;
  int segs = tiles;
  uint32_t mask = 0xffffffffu;
  if (ranges == nullptr) {
    ranges = reinterpret_cast<int*>(w.seg_ranges);
    segs = 1 << seg_bits;
    mask = (uint32_t)segs - 1u;
  }
  int st = PS_OK;
  // The last digit pass clears the segment ranges and the four list
  // counters; without digit passes (nothing to sort) the boundary scan's own
  // clearing launch does.
  const bool passes = d_n != nullptr || n_host > 0;
  if (passes) {
    ZeroReq z;
    z.ranges = reinterpret_cast<int2*>(ranges);
    z.n = segs;
    z.words = w.big_count;
    z.nwords = 4;
    st = digit_passes(n_host, d_n, max_pairs, 32, end_bit, w, keys_in, vals_in, keys_out,
                      vals_out, stream, z);
  }
  if (st != PS_OK) return st;
  st = tile_ranges_scan_launch(n_host, d_n, max_pairs, segs, keys_out, ranges, stream, mask,
                               w.big_count, 4, /*cleared=*/passes);
  if (st != PS_OK) return st;
  if (d_n == nullptr && n_host == 0) return PS_OK;
  const int big_cap = (int)big_list_cap((size_t)(max_pairs > 0 ? max_pairs : 1));
  int grid = (segs + kSegWaves - 1) / kSegWaves;
  if (grid > 16384) grid = 16384;
  hipLaunchKernelGGL(ps_seg_small, dim3(grid), dim3(kSegWaves * kWave), 0, stream, keys_out,
                     vals_out, reinterpret_cast<const int2*>(ranges), segs, w.big_list,
                     w.big_count, big_cap);
  PS_LAUNCH_CHECK();
  // Listed segments go through the workgroup routes in turn; each passes on
  // what it cannot hold.  Lists: big_list -> big_list2 -> big_list3 ->
  // big_list -> ps_seg_big (a list is reused once its reader is done),
  // counters 0..3.
  const int2* rg = reinterpret_cast<const int2*>(ranges);
  // 256-thread workgroups, five waves per SIMD.
  hipLaunchKernelGGL((ps_seg_group<4, kGroupRounds, 5>), dim3(2048), dim3(4 * kWave), 0, stream,
                     keys_out, vals_out, rg, w.big_list, w.big_count, big_cap, w.big_list2,
                     w.big_count + 1, big_cap);
  PS_LAUNCH_CHECK();
  // Two 512-thread workgroups per CU.
  hipLaunchKernelGGL((ps_seg_group<8, kMidRounds, 4>), dim3(1024), dim3(8 * kWave), 0, stream,
                     keys_out, vals_out, rg, w.big_list2, w.big_count + 1, big_cap, w.big_list3,
                     w.big_count + 2, big_cap);
  PS_LAUNCH_CHECK();
  // One 1024-thread workgroup per CU (its LDS buffer takes most of the CU's
  // 160 KiB).
  hipLaunchKernelGGL((ps_seg_group<16, kLargeRounds, 4>), dim3(256), dim3(16 * kWave), 0, stream,
                     keys_out, vals_out, rg, w.big_list3, w.big_count + 2, big_cap, w.big_list,
                     w.big_count + 3, big_cap);
  PS_LAUNCH_CHECK();
  // One workgroup per listed segment at a time; enough of them to fill the
  // device (about seven per CU fit), since each walks its segment serially.
  hipLaunchKernelGGL(ps_seg_big, dim3(2048), dim3(kSortThreads), 0, stream, keys_out, vals_out,
                     w.alt_keys, w.alt_vals, rg, w.big_list, w.big_count + 3, big_cap);
  PS_LAUNCH_CHECK();
  return PS_OK;
}

// Depth sort of the depth-first pipeline (see DepthSortArgs): four 8-bit
// digit launches each; a digit that does not differ among the emitting
// depths returns at once.  The sort space supplies the digit counts only.
int depth_sort_launch(const DepthSortArgs& A, int max_pairs, void* space, size_t space_bytes,
                      hipStream_t stream) {
  PS_REQUIRE(space != nullptr && space_bytes >= sort_layout(max_pairs, nullptr, nullptr),
             "sort space too small");
  if (A.n <= 0) return PS_OK;
  SortWs w;
  sort_layout(max_pairs, static_cast<char*>(space), &w);
  for (int q = 0; q < 4; ++q) {
    hipLaunchKernelGGL(ps_dsort_upsweep, dim3(kSortBlocks), dim3(kSortThreads), 0, stream, A, q,
                       w.counts);
    hipLaunchKernelGGL(ps_sort_scan, dim3(256), dim3(kSortThreads), 0, stream, w.counts,
                       w.offsets, w.totals, ZeroReq());
    hipLaunchKernelGGL(ps_dsort_downsweep, dim3(kSortBlocks), dim3(kSortThreads), 0, stream, A,
                       q, (const uint32_t*)w.offsets, (const uint32_t*)w.totals);
    PS_LAUNCH_CHECK();
  }
  return PS_OK;
}

// Tile sort of the depth-first pipeline: stable sort of the (tile id, value)
// pairs, emitted in depth order, by their 32-bit tile ids (bits
// [0, bit_length(tiles))), then the 64-bit keys and the tile ranges.
int tile_sort_launch(int n_host, const int* d_n, int max_pairs, int tiles, void* space,
                     size_t space_bytes, const uint32_t* tkeys_in, const uint32_t* vals_in,
                     uint32_t* tkeys_out, uint32_t* vals_out, uint64_t* keys_out,
                     int* tile_ranges, hipStream_t stream) {
  PS_REQUIRE(max_pairs >= 0, "max_pairs must be >= 0");
  PS_REQUIRE(d_n != nullptr || (n_host >= 0 && n_host <= max_pairs),
             "num_rendered %d outside [0, %d]", n_host, max_pairs);
  PS_REQUIRE(space != nullptr && space_bytes >= sort_layout(max_pairs, nullptr, nullptr),
             "sort space too small");
  PS_REQUIRE(tiles > 0, "tile ranges need a tile count");
  PS_REQUIRE(tkeys_in != tkeys_out && vals_in != vals_out, "sort outputs must not alias inputs");
  PS_REQUIRE(keys_out != nullptr, "keys_sorted is NULL");
  SortWs w;
  sort_layout(max_pairs, static_cast<char*>(space), &w);
  const int st = digit_passes<uint32_t>(n_host, d_n, max_pairs, 0, bit_length((unsigned)tiles),
                                        w, tkeys_in, vals_in, tkeys_out, vals_out, stream);
  if (st != PS_OK) return st;
  return tile_ids_finish_launch(n_host, d_n, max_pairs, tiles, tkeys_out, keys_out, tile_ranges,
                                stream);
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
