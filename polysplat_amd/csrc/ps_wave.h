// Wave64 building blocks for gfx950 (CDNA4): lane masks, ballots, uniform
// broadcasts and per-lane permutes, plus 256-thread block scans.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ps {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() {
  return (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p ? 1 : 0); }
__device__ __forceinline__ int popc64(uint64_t m) { return __popcll(m); }
__device__ __forceinline__ uint64_t lanes_below(int lane) {
  return (lane >= 64) ? ~0ull : ((1ull << lane) - 1ull);
}
// Bits [lo, hi) of a 64-bit lane mask (0 <= lo <= hi <= 64).
__device__ __forceinline__ uint64_t lane_range(int lo, int hi) {
  return lanes_below(hi) & ~lanes_below(lo);
}

// Broadcast from a wave-uniform lane index (v_readlane).
__device__ __forceinline__ int bcast(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ float bcast(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
// Per-lane gather from an arbitrary lane (ds_bpermute).  Call with the whole
// wave active.
__device__ __forceinline__ int permute(int v, int src) {
  return __builtin_amdgcn_ds_bpermute(src << 2, v);
}
__device__ __forceinline__ float permute(float v, int src) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}

// Inclusive prefix sum across the wave.
__device__ __forceinline__ int wave_incl_scan(int x, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int n = permute(x, lane - d);
    if (lane >= d) x += n;
  }
  return x;
}

__device__ __forceinline__ int wave_sum(int x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d);
  return x;
}

// Exclusive scan over a 256-thread block.  `scratch` needs 4 ints of LDS.
// Returns the exclusive prefix; `total` receives the block sum.
__device__ __forceinline__ int block256_excl_scan(int x, int* scratch, int& total) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int incl = wave_incl_scan(x, lane);
  if (lane == 63) scratch[wave] = incl;
  __syncthreads();
  int w0 = scratch[0], w1 = scratch[1], w2 = scratch[2], w3 = scratch[3];
  int before = (wave > 0 ? w0 : 0) + (wave > 1 ? w1 : 0) + (wave > 2 ? w2 : 0);
  total = w0 + w1 + w2 + w3;
  __syncthreads();
  return before + incl - x;
}

}  // namespace ps
