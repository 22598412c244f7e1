// Preprocess for the MI355X forward rasterizer: projection, EWA conic, SH
// colour and deterministic tile-key emission.
//
// The reference (preprocess.cu:313-641) emits keys with one global atomic
// counter, so the unsorted order -- and the order of equal keys after the
// stable sort -- changes from run to run.  Here the key array comes out in a
// canonical order (Gaussian id, then tile row-major), with the exact pair
// count known on the device before any key is written:
//   1. ps_pre_count    : project every Gaussian, count the tiles its ellipse
//                        really touches, write its render features and an
//                        emit record (rect, depth bits, hit mask), and add the
//                        256-Gaussian block's count into its chunk sum;
//   2. ps_pre_emit_rec : each workgroup's offset from the chunk sums and the
//                        block sums of its chunk; keys written from the hit
//                        masks (a Gaussian whose rect is too large for a mask
//                        is projected and swept again, bit-identically).
// POLYSPLAT_EMIT=3pass selects the earlier form: count, one scan kernel over
// the block sums, and an emit pass that projects and sweeps every emitting
// Gaussian again.
//
// Tile enumeration inside a wave routes each Gaussian by its candidate-tile
// count (the rect area):
//   single     (area == 1)        : the owning lane tests and writes directly;
//   fragmented (2 <= area < 64)   : the wave flattens all such rects into one
//                                   task stream, 64 tile tests per step;
//   massive    (area >= 64)       : the whole wave sweeps that one rect, 64
//                                   tiles per step.
// The reference routes per warp with `multi_count <= 2 || max_tile_count >= 32`
// (preprocess.cu:444).  With 64-lane waves a rect of 64+ candidates alone
// fills every lane of a sweep step, so the saturation threshold becomes 64
// and the decision is made per Gaussian rather than per wave.
//
// Compile with -ffp-contract=off: the key-deciding arithmetic must match the
// host restatement bit for bit.

#include "ps_common.h"
#include "ps_host.h"
#include "ps_wave.h"

#include <stdlib.h>
#include <string.h>

#include <type_traits>

namespace ps {
namespace {

constexpr int kPreThreads = 256;
constexpr int kMassive = 64;

// What the count pass leaves for the emit pass about a Gaussian that emits,
// indexing its candidate rect's tiles row-major (local index t):
// * rect of at most kMaskTiles tiles: x0w = x0 | width << 24, y0, the depth
//   bits of its keys, and mask = the hit mask (bit t: tile t hit);
// * larger rect: x0w = x0 (width field 0), y0 = y0 | width << 16, depth, and
//   mask = the first of its ceil(area / 64) words in the hit pool (word s,
//   bit l: tile 64 s + l hit), the ballots of the count pass's tile sweep; or
//   kNoPool when the pool was full, and the emit pass projects and sweeps
//   that Gaussian again.
struct alignas(16) EmitRec {
  uint32_t x0w;
  uint32_t y0;
  uint32_t depth;
  uint32_t mask;
};
// state[0] = offset the frame's keys start at, state[1] = pair count for the
// later stages (0 after overflow), state[2] = status bits (bit 0: overflow).
// state[4..5]: words of the hit pool taken so far (a 64-bit counter).
constexpr int kStateBase = 0, kStateCount = 1, kStateFlags = 2, kStatePool = 4, kStateWords = 64;

constexpr int kMaskTiles = 32;
// Longest hit mask a lane of the emit pass walks by itself (see
// emit_records_block).
constexpr int kEmitLoopMax = 8;
constexpr uint32_t kNoPool = 0xffffffffu;
// Hit pool: each count-pass workgroup owns kPoolPerBlock words at
// vb * kPoolPerBlock (its Gaussians take them by a block-wide prefix over
// their needs, with no atomics); a workgroup that needs more takes its words
// from a shared overflow area behind them by one atomic add, and if that is
// full too its large rects are swept again by the emit pass.  About one
// Gaussian in a hundred has a rect of more than 32 tiles at the benchmark
// workloads, most of them one or two words.
constexpr int kPoolPerBlock = 64;
inline size_t pool_overflow_words(int nb) { return (size_t)nb * 16 + 16384; }
inline size_t pool_words_for(int nb) {
  return (size_t)nb * kPoolPerBlock + pool_overflow_words(nb);
}
// The count pass adds each workgroup's pair total into the sum of its chunk
// of kChunkBlocks workgroups, so an emit workgroup finds its offset from at
// most chunks + kChunkBlocks sums.
constexpr int kChunkBlocks = 256;
inline int chunks_for(int nb) { return (nb + kChunkBlocks - 1) / kChunkBlocks; }

// A lane with no Gaussian (past P, culled, or emitting nothing).
__device__ __forceinline__ Proj no_proj() {
  Proj p;
  p.xy_x = p.xy_y = p.con_a = p.con_b = p.con_c = 0.0f;
  p.power = p.log2_opacity = p.depth = 0.0f;
  p.rmin_x = p.rmin_y = p.rmax_x = p.rmax_y = 0;
  p.valid = false;
  return p;
}

// Per-lane view of a projected Gaussian used by the sweeps.
struct Cand {
  float x, y, a, b, c, power, depth;
  int rx0, ry0, rx1, ry1;
};

__device__ __forceinline__ Cand cand_from(const Proj& p) {
  Cand c;
  c.x = p.xy_x;
  c.y = p.xy_y;
  c.a = p.con_a;
  c.b = p.con_b;
  c.c = p.con_c;
  c.power = p.power;
  c.depth = p.depth;
  c.rx0 = p.rmin_x;
  c.ry0 = p.rmin_y;
  c.rx1 = p.rmax_x;
  c.ry1 = p.rmax_y;
  return c;
}

__device__ __forceinline__ Proj proj_from(const Cand& c) {
  Proj p;
  p.xy_x = c.x;
  p.xy_y = c.y;
  p.con_a = c.a;
  p.con_b = c.b;
  p.con_c = c.c;
  p.power = c.power;
  p.depth = c.depth;
  p.rmin_x = c.rx0;
  p.rmin_y = c.ry0;
  p.rmax_x = c.rx1;
  p.rmax_y = c.ry1;
  p.valid = true;
  return p;
}

__device__ __forceinline__ Cand bcast_cand(const Cand& v, int i) {
  Cand c;
  c.x = bcast(v.x, i);
  c.y = bcast(v.y, i);
  c.a = bcast(v.a, i);
  c.b = bcast(v.b, i);
  c.c = bcast(v.c, i);
  c.power = bcast(v.power, i);
  c.depth = bcast(v.depth, i);
  c.rx0 = bcast(v.rx0, i);
  c.ry0 = bcast(v.ry0, i);
  c.rx1 = bcast(v.rx1, i);
  c.ry1 = bcast(v.ry1, i);
  return c;
}

__device__ __forceinline__ Cand permute_cand(const Cand& v, int src) {
  Cand c;
  c.x = permute(v.x, src);
  c.y = permute(v.y, src);
  c.a = permute(v.a, src);
  c.b = permute(v.b, src);
  c.c = permute(v.c, src);
  c.power = permute(v.power, src);
  c.depth = permute(v.depth, src);
  c.rx0 = permute(v.rx0, src);
  c.ry0 = permute(v.ry0, src);
  c.rx1 = permute(v.rx1, src);
  c.ry1 = permute(v.ry1, src);
  return c;
}

// The sweeps report each step to a sink, called by the whole wave once per
// step: `ok` is this lane's tile-test result, `m` the step's ballot of them,
// `tile` this lane's frame-local tile id, `owner` the lane that owns this
// lane's Gaussian and `k` the index of this hit among that Gaussian's hits
// (valid where ok).
struct NullSink {
  __device__ __forceinline__ void step(bool, uint64_t, int, int, int) {}
  // Massive route: step s of lane i's rect gave ballot m (tiles 64 s ..).
  __device__ __forceinline__ void ballot(int, int, uint64_t) {}
};

// Count pass: keeps the massive route's ballots of each Gaussian that has a
// hit-pool area (base >= 0) as its pool words.
struct PoolSink {
  uint64_t* pool;
  int base, lane;
  __device__ __forceinline__ void step(bool, uint64_t, int, int, int) {}
  __device__ __forceinline__ void ballot(int i, int s, uint64_t m) {
    const int b = bcast(base, i);
    if (lane == 0 && b >= 0) pool[b + s] = m;
  }
};

// Writes keys straight to global memory at off(owner) + k, never past the
// owner's count.
struct GlobalSink {
  int off, cnt, gid, tile0;
  float depth;
  uint64_t* keys;
  uint32_t* vals;
  __device__ __forceinline__ void step(bool ok, uint64_t, int tile, int owner, int k) {
    const int o_off = permute(off, owner), o_cnt = permute(cnt, owner);
    const int o_gid = permute(gid, owner);
    const float o_depth = permute(depth, owner);
    if (ok && k < o_cnt) {
      keys[o_off + k] = make_key(tile0 + tile, o_depth);
      vals[o_off + k] = (uint32_t)o_gid;
    }
  }
  __device__ __forceinline__ void ballot(int, int, uint64_t) {}
};

// Massive Gaussians: the wave walks each one's rect 64 tiles at a time in
// row-major order.  Returns this lane's own tile count.
template <class Sink>
__device__ int sweep_massive(const Cand& me, bool is_massive, int lane, int bx, int by,
                             int grid_x, Sink& sink) {
  uint64_t todo = ballot(is_massive);
  int mine = 0;
  const uint64_t below = lanes_below(lane);
  while (todo) {
    const int i = __builtin_ctzll(todo);
    todo &= todo - 1;
    const Cand q = bcast_cand(me, i);
    const Proj pq = proj_from(q);
    const int w = q.rx1 - q.rx0;
    const int total = w * (q.ry1 - q.ry0);
    int run = 0;
    for (int t0 = 0; t0 < total; t0 += kWave) {
      const int t = t0 + lane;
      bool ok = false;
      int tx = 0, ty = 0;
      if (t < total) {
        ty = t / w;
        tx = t - ty * w;
        ok = tile_hits(q.rx0 + tx, q.ry0 + ty, pq, bx, by);
      }
      const uint64_t m = ballot(ok);
      sink.step(ok, m, (q.ry0 + ty) * grid_x + (q.rx0 + tx), i, run + popc64(m & below));
      sink.ballot(i, t0 / kWave, m);
      run += popc64(m);
    }
    if (lane == i) mine = run;
  }
  return mine;
}

// Fragmented Gaussians: all lanes' small rects become one stream of
// (Gaussian, tile) tasks, 64 per step; the owner of task t is the first lane
// whose inclusive prefix of areas exceeds t.  Returns this lane's own count.
template <class Sink>
__device__ int sweep_fragmented(const Cand& me, bool is_frag, int lane, int bx, int by,
                                int grid_x, Sink& sink, uint64_t& mask) {
  const int area = is_frag ? (me.rx1 - me.rx0) * (me.ry1 - me.ry0) : 0;
  const int incl = wave_incl_scan(area, lane);
  const int excl = incl - area;
  const int total = bcast(incl, 63);
  const uint64_t below = lanes_below(lane);
  int mine = 0;
  for (int t0 = 0; t0 < total; t0 += kWave) {
    const int t = t0 + lane;
    const bool tv = t < total;
    int lo = 0, hi = 63;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const int mid = (lo + hi) >> 1;
      const int v = permute(incl, mid);
      if (v > t) hi = mid; else lo = mid + 1;
    }
    const int o = tv ? lo : lane;
    const Cand q = permute_cand(me, o);
    const int o_excl = permute(excl, o);
    const int o_done = permute(mine, o);
    bool ok = false;
    int tx = 0, ty = 0;
    if (tv) {
      const int w = q.rx1 - q.rx0;
      const int local = t - o_excl;
      // local / w for local < 64, w < 64: the reciprocal is within one ulp
      // of 1 / w, so local * rcp(w) is off by under 2^-16, and the 1/128
      // bias lifts exact quotients without crossing the next integer (the
      // fraction of an inexact one is at most 1 - 1/63); checked for every
      // (local, w) and both one-ulp neighbours of the reciprocal.
      ty = (int)((float)local * __builtin_amdgcn_rcpf((float)w) + (1.0f / 128.0f));
      tx = local - ty * w;
      ok = tile_hits(q.rx0 + tx, q.ry0 + ty, proj_from(q), bx, by);
    }
    const uint64_t m = ballot(ok);
    const int seg_lo = imax(o_excl - t0, 0);
    sink.step(ok, m, (q.ry0 + ty) * grid_x + (q.rx0 + tx), o,
              o_done + popc64(m & below & ~lanes_below(seg_lo)));
    // Each owner adds the passing tasks of its own segment in this step, and
    // their bits to its mask (rect-local index of lane j: t0 + j - excl).
    if (is_frag) {
      const int lo_l = imax(excl - t0, 0), hi_l = imin(incl - t0, kWave);
      if (lo_l < hi_l) {
        const uint64_t seg = m & lane_range(lo_l, hi_l);
        mine += popc64(seg);
        // rect-local indices stay below the area, itself below kMassive = 64
        const int first_local = t0 + lo_l - excl;
        mask |= (seg >> lo_l) << first_local;
      }
    }
  }
  return mine;
}

// All of a wave's (Gaussian, tile) hits, in the three routes: single-tile
// rects (tested by their own lane), massive and fragmented.  Returns this
// lane's count and, for rects of fewer than kMassive tiles, its hit mask
// (bit t: rect-local tile t hit).  `single_ok` is this lane's result for a
// single-tile rect.
template <class Sink>
__device__ __forceinline__ int sweep_all(const Cand& me, int area, bool single_ok, int lane,
                                         int bx, int by, int grid_x, Sink& sink,
                                         uint64_t& mask) {
  sink.step(single_ok, ballot(single_ok), me.ry0 * grid_x + me.rx0, lane, 0);
  int count = single_ok ? 1 : 0;
  mask = single_ok ? 1ull : 0ull;
  count += sweep_massive(me, area >= kMassive, lane, bx, by, grid_x, sink);
  count += sweep_fragmented(me, area >= 2 && area < kMassive, lane, bx, by, grid_x, sink, mask);
  return count;
}

__device__ __forceinline__ void load_cov(const float* cov3, int g, float c[6]) {
  const float2* p = reinterpret_cast<const float2*>(cov3 + (size_t)g * 6);
  float2 a = p[0], b = p[1], d = p[2];
  c[0] = a.x; c[1] = a.y; c[2] = b.x; c[3] = b.y; c[4] = d.x; c[5] = d.y;
}

// The 96 bytes of a Gaussian's 16-bit SH as loaded (two coefficients per
// word, coefficient-major).
struct ShRaw16 {
  uint4 q[6];
};

__device__ __forceinline__ ShRaw16 load_sh_raw16(const void* shs, int g) {
  const uint4* p = reinterpret_cast<const uint4*>(static_cast<const uint16_t*>(shs) + (size_t)g * 48);
  ShRaw16 r;
#pragma unroll
  for (int k = 0; k < 6; ++k) r.q[k] = p[k];
  return r;
}

template <int SHP>
__device__ __forceinline__ void sh_from_raw16(const ShRaw16& raw, float sh[48]) {
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const uint32_t w4[4] = {raw.q[k].x, raw.q[k].y, raw.q[k].z, raw.q[k].w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint16_t lo = (uint16_t)(w4[j] & 0xffffu), hi = (uint16_t)(w4[j] >> 16);
      if (SHP == kShF16) {
        sh[8 * k + 2 * j] = __half2float(__ushort_as_half(lo));
        sh[8 * k + 2 * j + 1] = __half2float(__ushort_as_half(hi));
      } else {
        sh[8 * k + 2 * j] = bf16_to_f32(lo);
        sh[8 * k + 2 * j + 1] = bf16_to_f32(hi);
      }
    }
  }
}

__device__ __forceinline__ void load_sh_f32(const void* shs, int g, float sh[48]) {
  const float4* p = reinterpret_cast<const float4*>(static_cast<const float*>(shs) + (size_t)g * 48);
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    float4 v = p[k];
    sh[4 * k + 0] = v.x; sh[4 * k + 1] = v.y; sh[4 * k + 2] = v.z; sh[4 * k + 3] = v.w;
  }
}

__device__ __forceinline__ uint16_t to_half_bits(float f, int prec) {
  return prec == kFeatBF16 ? f32_to_bf16(f) : __half_as_ushort(__float2half_rn(f));
}

// Projects Gaussian g of frame F (no Gaussian when g >= P).
__device__ __forceinline__ Proj project_gaussian(const FrameDev& F, int g, float& px, float& py,
                                                 float& pz) {
  px = py = pz = 0.0f;
  if (g >= F.P) return no_proj();
  px = F.pos[3 * (size_t)g + 0];
  py = F.pos[3 * (size_t)g + 1];
  pz = F.pos[3 * (size_t)g + 2];
  float c[6];
  load_cov(F.cov3, g, c);
  return project(px, py, pz, F.opac[g], c, F.cp);
}

// Render features of an emitting Gaussian: SH colour, scaled conic, and the
// FP32 arrays and/or the 16-bit record, whichever the frame asks for.
template <int SHP>
__device__ __forceinline__ void write_features(const FrameDev& F, int g, const Proj& pj, float px,
                                               float py, float pz) {
  const float4 con = make_float4((-0.5f * kLog2e) * pj.con_a, -kLog2e * pj.con_b,
                                 (-0.5f * kLog2e) * pj.con_c, pj.log2_opacity);
  float sh[48];
  if (SHP == kShF32)
    load_sh_f32(F.shs, g, sh);
  else
    sh_from_raw16<SHP>(load_sh_raw16(F.shs, g), sh);
  const F3 rgb = sh_color(px, py, pz, F.cp.campos, sh);
  if (F.xy) F.xy[g] = make_float2(pj.xy_x, pj.xy_y);
  if (F.conic) F.conic[g] = con;
  if (F.rgbd) F.rgbd[g] = make_float4(rgb.x, rgb.y, rgb.z, pj.depth);
  if (F.rec) {
    FeatRec16 r;
    r.x = pj.xy_x;
    r.y = pj.xy_y;
    r.w = con.w;
    r.ca = to_half_bits(con.x * kConicScale, F.rec_prec);
    r.cb = to_half_bits(con.y * kConicScale, F.rec_prec);
    r.cc = to_half_bits(con.z * kConicScale, F.rec_prec);
    r.r = to_half_bits(rgb.x, F.rec_prec);
    r.g = to_half_bits(rgb.y, F.rec_prec);
    r.b = to_half_bits(rgb.z, F.rec_prec);
    F.rec[g] = r;
  }
}

__device__ __forceinline__ int rect_area(const Proj& pj) {
  return pj.valid ? (pj.rmax_x - pj.rmin_x) * (pj.rmax_y - pj.rmin_y) : 0;
}

// Count pass for the workgroup `vb` of a launch (vb - F.block0 is the
// frame-local workgroup).  counts/block_sums are indexed by launch position.
// With `recs`, also leaves each emitting Gaussian's EmitRec, adds the
// workgroup's total into its chunk sum and (workgroup 0) records the
// frame's starting offset in state[kStateBase].
template <int SHP>
__device__ __forceinline__ void pre_count_body(const FrameDev& F, int vb, int* __restrict__ counts,
                                               int* __restrict__ block_sums,
                                               EmitRec* __restrict__ recs,
                                               unsigned long long* __restrict__ chunk_sums,
                                               int* __restrict__ state,
                                               const int* __restrict__ curr_offset,
                                               int reset_offset, uint64_t* __restrict__ pool,
                                               long long pool_cap, long long pool_overflow,
                                               uint32_t* __restrict__ dkey,
                                               uint32_t* __restrict__ plan) {
  __shared__ int scratch[4];
  const int tid = threadIdx.x, lane = tid & 63;
  const int g = (vb - F.block0) * kPreThreads + tid;
  const CamParams& cp = F.cp;
  float px, py, pz;
  const Proj pj = project_gaussian(F, g, px, py, pz);
  const int area = rect_area(pj);
  const bool single_ok = area == 1 && tile_hits(pj.rmin_x, pj.rmin_y, pj, cp.block_x, cp.block_y);
  // Hit-pool words for rects of more than kMaskTiles tiles (see
  // kPoolPerBlock): a block-wide prefix over the Gaussians' needs.
  int pbase = -1;
  if (pool) {  // uniform
    __shared__ long long s_over;
    const int need = area > kMaskTiles ? (area + kWave - 1) / kWave : 0;
    int btot;
    const int bexcl = block256_excl_scan(need, scratch, btot);
    if (btot > 0) {  // uniform across the block
      const long long own = (long long)vb * kPoolPerBlock;
      long long start = own;
      if (btot > kPoolPerBlock) {  // uniform: the shared overflow area
        if (tid == 0) {
          const unsigned long long b0 = atomicAdd(
              reinterpret_cast<unsigned long long*>(state + kStatePool), (unsigned long long)btot);
          s_over = (long long)b0;
        }
        __syncthreads();
        const long long over0 = pool_cap - (long long)pool_overflow;
        start = (s_over + btot <= (long long)pool_overflow) ? over0 + s_over : -1;
      }
      pbase = (need > 0 && start >= 0) ? (int)(start + bexcl) : -1;
    }
  }
  PoolSink sink{pool, pbase, lane};
  uint64_t mask;
  const int count = sweep_all(cand_from(pj), area, single_ok, lane, cp.block_x, cp.block_y,
                              cp.grid_x, sink, mask);
  // Rects of kMaskTiles < area < kMassive: their one pool word is the
  // fragmented route's mask (the massive route stored its ballots).
  if (pbase >= 0 && area < kMassive) pool[pbase] = mask;
  counts[(size_t)vb * kPreThreads + tid] = count;
  if (recs && g < F.P) {
    // Written for every Gaussian (zeros when it emits nothing) so that each
    // wave stores whole cache lines.
    const bool small = area <= kMaskTiles;
    const uint32_t rw = (uint32_t)(pj.rmax_x - pj.rmin_x);
    EmitRec r;
    r.x0w = count > 0 ? ((uint32_t)pj.rmin_x | ((small ? rw : 0u) << 24)) : 0u;
    r.y0 = count > 0 ? ((uint32_t)pj.rmin_y | (small ? 0u : rw << 16)) : 0u;
    r.depth = count > 0 ? f2u(pj.depth) : 0u;
    r.mask = count > 0 ? (small ? (uint32_t)mask : (pbase >= 0 ? (uint32_t)pbase : kNoPool))
                       : 0u;
    recs[(size_t)vb * kPreThreads + tid] = r;
  }
  if (count > 0) write_features<SHP>(F, g, pj, px, py, pz);
  if (dkey) {  // uniform: the depth-first pipeline's depth words and plan
    if (g < F.P) dkey[g] = count > 0 ? f2u(pj.depth) : 0u;
    __shared__ uint32_t s_or[kPreThreads / kWave], s_and[kPreThreads / kWave];
    uint32_t o = count > 0 ? f2u(pj.depth) : 0u, a = count > 0 ? f2u(pj.depth) : ~0u;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      o |= (uint32_t)__shfl_xor((int)o, d);
      a &= (uint32_t)__shfl_xor((int)a, d);
    }
    if (lane == 0) {
      s_or[tid >> 6] = o;
      s_and[tid >> 6] = a;
    }
    __syncthreads();
    if (tid == 0) {
      o = s_or[0] | s_or[1] | s_or[2] | s_or[3];
      a = s_and[0] & s_and[1] & s_and[2] & s_and[3];
      if (o != 0u) {  // the block has an emitting Gaussian
        const int slot = vb % kPlanSlots;
        atomicOr(&plan[slot * kPlanStride], o);
        atomicAnd(&plan[(kPlanSlots + slot) * kPlanStride], a);
      }
    }
  }
  int total;
  block256_excl_scan(count, scratch, total);
  if (tid == 0) {
    block_sums[vb] = total;
    if (chunk_sums) atomicAdd(&chunk_sums[vb / kChunkBlocks], (unsigned long long)total);
    if (recs && vb == 0) state[kStateBase] = reset_offset ? 0 : *curr_offset;
  }
}

template <int SHP>
__global__ __launch_bounds__(kPreThreads) void ps_pre_count(
    FrameDev F, int* __restrict__ counts, int* __restrict__ block_sums,
    EmitRec* __restrict__ recs, unsigned long long* __restrict__ chunk_sums,
    int* __restrict__ state, const int* __restrict__ curr_offset, int reset_offset,
    uint64_t* __restrict__ pool, long long pool_cap, long long pool_overflow,
    uint32_t* __restrict__ dkey, uint32_t* __restrict__ plan) {
  bind_camera(F);
  pre_count_body<SHP>(F, blockIdx.x, counts, block_sums, recs, chunk_sums, state, curr_offset,
                      reset_offset, pool, pool_cap, pool_overflow, dkey, plan);
}

__device__ __forceinline__ int frame_of_block(const FrameDev* frames, int nf, int vb) {
  return frame_search(frames, nf, vb, [](const FrameDev& f) { return f.block0; });
}

template <int SHP>
__global__ __launch_bounds__(kPreThreads) void ps_pre_count_batch(const FrameDev* __restrict__ frames,
                                                                  int nf, int* __restrict__ counts,
                                                                  int* __restrict__ block_sums) {
  const int f = frame_of_block(frames, nf, blockIdx.x);
  pre_count_body<SHP>(frames[f], blockIdx.x, counts, block_sums, nullptr, nullptr, nullptr,
                      nullptr, 0, nullptr, 0, 0, nullptr, nullptr);
}


__global__ __launch_bounds__(1024) void ps_pre_scan(int nblocks, const int* __restrict__ block_sums,
                                                    int* __restrict__ block_offsets,
                                                    int* __restrict__ curr_offset, int capacity,
                                                    int reset_offset, int* __restrict__ state,
                                                    const FrameDev* __restrict__ frames, int nf,
                                                    int* __restrict__ frame_counts) {
  __shared__ int wsum[16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int per = (nblocks + 1023) / 1024;
  const int b0 = tid * per;
  // Each thread sums a contiguous run of block sums, 8 loads in flight at a
  // time (the scan is latency-bound, not bandwidth-bound).
  int local = 0;
  for (int k0 = 0; k0 < per; k0 += 8) {
    int v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int idx = b0 + k0 + j;
      v[j] = (k0 + j < per && idx < nblocks) ? block_sums[idx] : 0;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) local += v[j];
  }
  int incl = wave_incl_scan(local, lane);
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  int before = 0, total = 0;
  for (int w = 0; w < 16; ++w) {
    const int v = wsum[w];
    if (w < wave) before += v;
    total += v;
  }
  int run = before + incl - local;
  for (int k0 = 0; k0 < per; k0 += 8) {
    int v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int idx = b0 + k0 + j;
      v[j] = (k0 + j < per && idx < nblocks) ? block_sums[idx] : 0;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int idx = b0 + k0 + j;
      if (k0 + j < per && idx < nblocks) block_offsets[idx] = run;
      run += v[j];
    }
  }
  if (tid == 0) {
    const int base = reset_offset ? 0 : *curr_offset;
    const long long end = (long long)base + (long long)total;
    // A negative starting offset cannot be written at: reported as overflow.
    const bool overflow = base < 0 || end > (long long)capacity || end > 0x7fffffffLL;
    *curr_offset = (int)(end > 0x7fffffffLL ? 0x7fffffff : end);
    state[kStateBase] = base;
    state[kStateCount] = overflow ? 0 : (int)end;
    state[kStateFlags] = overflow ? 1 : 0;
  }
  if (frames != nullptr && frame_counts != nullptr) {
    // Pairs of each frame of a batch: difference of its first block's offset
    // and the next frame's.
    __syncthreads();
    for (int f = tid; f < nf; f += 1024) {
      const int b = frames[f].block0;
      const int bn = (f + 1 < nf) ? frames[f + 1].block0 : nblocks;
      const int start = (b < nblocks) ? block_offsets[b] : total;
      const int stop = (bn < nblocks) ? block_offsets[bn] : total;
      frame_counts[f] = stop - start;
    }
  }
}

__device__ __forceinline__ void pre_emit_body(const FrameDev& F, int vb,
                                              const int* __restrict__ counts,
                                              const int* __restrict__ block_offsets,
                                              const int* __restrict__ state,
                                              uint64_t* __restrict__ keys,
                                              uint32_t* __restrict__ vals) {
  __shared__ int scratch[4];
  if (state[kStateFlags] & 1) return;
  const int tid = threadIdx.x, lane = tid & 63;
  const int g = (vb - F.block0) * kPreThreads + tid;
  const int count = (g < F.P) ? counts[(size_t)vb * kPreThreads + tid] : 0;
  int total;
  const int excl = block256_excl_scan(count, scratch, total);
  if (total == 0) return;  // uniform across the block
  const CamParams& cp = F.cp;
  const int off = state[kStateBase] + block_offsets[vb] + excl;
  float px, py, pz;
  const Proj pj = count > 0 ? project_gaussian(F, g, px, py, pz) : no_proj();
  const int area = count > 0 ? rect_area(pj) : 0;
  const Cand me = cand_from(pj);
  GlobalSink sink{off, count, g, F.tile0, pj.depth, keys, vals};
  uint64_t mask;
  sweep_all(me, area, area == 1, lane, cp.block_x, cp.block_y, cp.grid_x, sink, mask);
}

__global__ __launch_bounds__(kPreThreads) void ps_pre_emit(FrameDev F, const int* __restrict__ counts,
                                                           const int* __restrict__ block_offsets,
                                                           const int* __restrict__ state,
                                                           uint64_t* __restrict__ keys,
                                                           uint32_t* __restrict__ vals) {
  bind_camera(F);
  pre_emit_body(F, blockIdx.x, counts, block_offsets, state, keys, vals);
}

__global__ __launch_bounds__(kPreThreads) void ps_pre_emit_batch(
    const FrameDev* __restrict__ frames, int nf, const int* __restrict__ counts,
    const int* __restrict__ block_offsets, const int* __restrict__ state,
    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int f = frame_of_block(frames, nf, blockIdx.x);
  pre_emit_body(frames[f], blockIdx.x, counts, block_offsets, state, keys, vals);
}

// ------------------------------------------------- emission from the records
__global__ __launch_bounds__(256) void ps_pre_chunk_reset(unsigned long long* __restrict__ sums,
                                                          int n,
                                                          unsigned long long* __restrict__ pool_taken,
                                                          uint32_t* __restrict__ plan) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) sums[i] = 0ull;
  if (i == 0) *pool_taken = 0ull;
  if (plan && i < kPlanSlots) {
    plan[i * kPlanStride] = 0u;                  // OR of emitting depth words
    plan[(kPlanSlots + i) * kPlanStride] = ~0u;  // AND
  }
}

__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const int lo = __shfl_xor((int)(uint32_t)(uint64_t)v, d);
    const int hi = __shfl_xor((int)(uint32_t)((uint64_t)v >> 32), d);
    v += (long long)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
  }
  return v;
}

// Where the emit passes put a (tile, Gaussian) pair: the 64-bit key
// (tile << 32 | depth word) and the id, or, in the depth-first pipeline,
// the 32-bit tile id and the id.
struct Out64 {
  uint64_t* keys;
  uint32_t* vals;
  __device__ __forceinline__ void put(int at, uint32_t tile, uint32_t depth, uint32_t gid) const {
    keys[at] = ((uint64_t)tile << 32) | (uint64_t)depth;
    vals[at] = gid;
  }
};
struct Out32 {
  uint32_t* tiles;
  uint32_t* vals;
  __device__ __forceinline__ void put(int at, uint32_t tile, uint32_t, uint32_t gid) const {
    tiles[at] = tile;
    vals[at] = gid;
  }
};

// Sweep sink writing through an Out (the re-projecting emit).
template <class Out>
struct OutSink {
  int off, cnt, gid, tile0;
  float depth;
  Out out;
  __device__ __forceinline__ void step(bool ok, uint64_t, int tile, int owner, int k) {
    const int o_off = permute(off, owner), o_cnt = permute(cnt, owner);
    const int o_gid = permute(gid, owner);
    const float o_depth = permute(depth, owner);
    if (ok && k < o_cnt) out.put(o_off + k, (uint32_t)(tile0 + tile), f2u(o_depth), (uint32_t)o_gid);
  }
  __device__ __forceinline__ void ballot(int, int, uint64_t) {}
};

// One workgroup of an emit pass over the count pass's records.  Thread tid
// handles Gaussian g (g >= F.P: none) with `count` pairs and record r; the
// workgroup's pairs start at the frame's offset plus the chunk sums before
// its chunk plus the block sums of its chunk before it, and each Gaussian's
// at the block's exclusive prefix of the counts, in thread order.  A
// Gaussian with a hit mask writes its pairs from the mask, one with pool
// words from those (the wave takes them one at a time, 64 rect tiles per
// word), and one whose rect found no room in the pool is projected and swept
// again.  Pairs are written only by workgroups whose whole range fits below
// `capacity`.  Workgroup nb - 1 publishes the total, the count for the later
// stages and the overflow bit.
template <class Out>
__device__ __forceinline__ void emit_records_block(
    const FrameDev& F, int vb, int nb, int g, int count, const EmitRec& r,
    const int* __restrict__ block_sums, const unsigned long long* __restrict__ chunk_sums,
    const uint64_t* __restrict__ pool, int* __restrict__ state, int* __restrict__ curr_offset,
    int capacity, const Out& out) {
  __shared__ int scratch[4];
  __shared__ long long s_part[kPreThreads / kWave];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int chunk = vb / kChunkBlocks;
  long long part = 0;
  for (int i = tid; i < chunk; i += kPreThreads) part += (long long)chunk_sums[i];
  const int b = chunk * kChunkBlocks + tid;
  if (b < vb) part += block_sums[b];
  part = wave_sum_i64(part);
  if (lane == 0) s_part[wave] = part;
  int total;
  const int excl = block256_excl_scan(count, scratch, total);  // its barriers publish s_part
  const long long base = state[kStateBase];
  const long long start = base + s_part[0] + s_part[1] + s_part[2] + s_part[3];
  // A negative starting offset cannot be written at: reported as overflow.
  const bool writable = base >= 0 && start + total <= (long long)capacity;
  if (vb == nb - 1 && tid == 0) {
    const long long end = start + total;
    const bool overflow = base < 0 || end > (long long)capacity || end > 0x7fffffffLL;
    *curr_offset = (int)(end > 0x7fffffffLL ? 0x7fffffffLL : end);
    state[kStateCount] = overflow ? 0 : (int)end;
    state[kStateFlags] = overflow ? 1 : 0;
  }
  if (total == 0 || !writable) return;  // uniform across the block
  const int off = (int)start + excl;
  const CamParams& cp = F.cp;
  const int w = count > 0 ? (int)(r.x0w >> 24) : 0;
  // Rects with a hit mask: pair k of a Gaussian is the k-th set bit of its
  // mask (row-major in its rect), written at off + k.  When every such rect
  // of the wave has at most kEmitLoopMax hits, each lane walks its own mask;
  // otherwise a lane with a long mask would keep the wave stepping with the
  // others idle, and the wave's pairs become one stream of tasks, 64 per
  // step, task t of lane o's run writing o's hit k = t - (o's start in the
  // stream), so consecutive lanes write consecutive pairs.  Both forms write
  // the same pairs to the same slots.
  const int c = (count > 0 && w > 0) ? imin(count, __builtin_popcount(r.mask)) : 0;
  if (ballot(c > kEmitLoopMax) == 0) {
    if (c > 0) {
      const int x0 = (int)(r.x0w & 0xffffffu), y0 = (int)r.y0;
      // t / w for t < 32 <= 2^10 and w <= 32 as (t * ceil(2^16 / w)) >> 16:
      // the rounding adds less than 2^-11 < 1/w to t / w, so the floor is
      // exact.
      const uint32_t inv_w = (65536u + (uint32_t)w - 1u) / (uint32_t)w;
      uint32_t m = r.mask;
      for (int k = 0; k < c; ++k) {
        const int t = __builtin_ctz(m);
        m &= m - 1u;
        const int ty = (int)(((uint32_t)t * inv_w) >> 16), tx = t - ty * w;
        out.put(off + k, (uint32_t)(F.tile0 + (y0 + ty) * cp.grid_x + (x0 + tx)), r.depth,
                (uint32_t)g);
      }
    }
  } else {
    const int incl = wave_incl_scan(c, lane);
    const int total_w = bcast(incl, 63);
    const uint32_t inv_w = w > 0 ? (65536u + (uint32_t)w - 1u) / (uint32_t)w : 0u;  // as above
    // Owner data, packed: rect origin (16 bits each: the grid has fewer than
    // 2^16 columns and rows), width and its reciprocal, and where the owner's
    // task 0 lands (off - its start in the stream).
    const int xy0 = (int)((r.x0w & 0xffffu) | (r.y0 << 16));
    const int wv = (int)((uint32_t)w | (inv_w << 8));
    const int excl_c = incl - c;
    const int base = off - excl_c;
    for (int t0 = 0; t0 < total_w; t0 += kWave) {  // uniform: every lane permutes
      const int t = t0 + lane;
      int lo = 0, hi = 63;
#pragma unroll
      for (int st = 0; st < 6; ++st) {
        const int mid = (lo + hi) >> 1;
        if (permute(incl, mid) > t) hi = mid; else lo = mid + 1;
      }
      const int o = lo;
      const int k = t - permute(excl_c, o);
      uint32_t m = (uint32_t)permute((int)r.mask, o);
      const uint32_t o_wv = (uint32_t)permute(wv, o), o_xy = (uint32_t)permute(xy0, o);
      const uint32_t o_depth = (uint32_t)permute((int)r.depth, o);
      const int o_g = permute(g, o), o_base = permute(base, o);
      if (t < total_w) {
        // position of the k-th set bit of m (k < popcount(m))
        int kk = k, pos = 0;
#pragma unroll
        for (int half = 16; half >= 1; half >>= 1) {
          const int below_half = __builtin_popcount(m & ((1u << half) - 1u));
          if (kk >= below_half) {
            kk -= below_half;
            m >>= half;
            pos += half;
          }
        }
        const int ow = (int)(o_wv & 0xffu);
        const int ty = (int)(((uint32_t)pos * (o_wv >> 8)) >> 16), tx = pos - ty * ow;
        const int ox0 = (int)(o_xy & 0xffffu), oy0 = (int)(o_xy >> 16);
        out.put(o_base + t, (uint32_t)(F.tile0 + (oy0 + ty) * cp.grid_x + (ox0 + tx)), o_depth,
                (uint32_t)o_g);
      }
    }
  }
  const uint64_t below = lanes_below(lane);
  const int tiles = cp.grid_x * cp.grid_y;
  uint64_t todo = ballot(count > 0 && w == 0 && r.mask != kNoPool);
  while (todo) {
    const int i = __builtin_ctzll(todo);
    todo &= todo - 1;
    const int pbase = bcast((int)r.mask, i), cnt = bcast(count, i), o_off = bcast(off, i);
    const int x0 = bcast((int)(r.x0w & 0xffffffu), i);
    const int y0 = bcast((int)(r.y0 & 0xffffu), i), rw = bcast((int)(r.y0 >> 16), i);
    const uint32_t dep = (uint32_t)bcast((int)r.depth, i);
    const uint32_t gid = (uint32_t)bcast(g, i);
    int run = 0;
    // the words hold exactly cnt hits; the bound only guards a corrupted pool
    for (int sw = 0; run < cnt && sw * kWave < tiles; ++sw) {
      const uint64_t m = pool[pbase + sw];
      if ((m >> lane) & 1ull) {
        const int t = sw * kWave + lane;
        const int ty = t / rw, tx = t - ty * rw;
        out.put(o_off + run + popc64(m & below),
                (uint32_t)(F.tile0 + (y0 + ty) * cp.grid_x + (x0 + tx)), dep, gid);
      }
      run += popc64(m);
    }
  }
  const bool big = count > 0 && w == 0 && r.mask == kNoPool;
  if (ballot(big) == 0) return;  // uniform across the wave
  float px, py, pz;
  const Proj pj = big ? project_gaussian(F, g, px, py, pz) : no_proj();
  const int area = big ? rect_area(pj) : 0;
  OutSink<Out> sink{off, count, g, F.tile0, pj.depth, out};
  uint64_t mask;
  sweep_all(cand_from(pj), area, false, lane, cp.block_x, cp.block_y, cp.grid_x, sink, mask);
}

// Emit pass in Gaussian order: keys go to their canonical slots.
__global__ __launch_bounds__(kPreThreads) void ps_pre_emit_rec(
    FrameDev F, int nb, const int* __restrict__ counts, const int* __restrict__ block_sums,
    const unsigned long long* __restrict__ chunk_sums, const EmitRec* __restrict__ recs,
    const uint64_t* __restrict__ pool, int* __restrict__ state, int* __restrict__ curr_offset,
    int capacity, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  bind_camera(F);
  const int g = blockIdx.x * kPreThreads + threadIdx.x;
  const int count = g < F.P ? counts[g] : 0;
  // The record is loaded before the scan (a lane that emits nothing ignores
  // whatever its slot holds), so its latency overlaps the scan's.
  const EmitRec r = g < F.P ? recs[g] : EmitRec{0u, 0u, 0u, 0u};
  emit_records_block(F, blockIdx.x, nb, g, count, r, block_sums, chunk_sums, pool, state,
                     curr_offset, capacity, Out64{keys, vals});
}

// ---------------------------------------------------- depth-first emission
// Slot s of the depth order holds Gaussian depth_order(s): the depth sort's
// values, or s itself when no digit of the sort ran.
struct DepthOrder {
  const uint32_t* ids;  // NULL: the identity
  __device__ __forceinline__ int at(int s) const { return ids ? (int)ids[s] : s; }
};

__device__ __forceinline__ DepthOrder depth_order(const uint32_t* __restrict__ plan,
                                                  uint32_t* const vb0, uint32_t* const vb1,
                                                  uint32_t* s_k) {
  const int tid = threadIdx.x;
  if (tid < kWave) {
    uint32_t o = plan[tid * kPlanStride], a = plan[(kPlanSlots + tid) * kPlanStride];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      o |= (uint32_t)__shfl_xor((int)o, d);
      a &= (uint32_t)__shfl_xor((int)a, d);
    }
    if (tid == 0) *s_k = (uint32_t)depth_digits_running(o == 0u ? 0u : (o ^ a));
  }
  __syncthreads();
  const int k = (int)*s_k;
  return DepthOrder{k == 0 ? nullptr : ((k - 1) & 1 ? vb1 : vb0)};
}

// Counts in depth order, their block sums and the chunk sums of those.
__global__ __launch_bounds__(kPreThreads) void ps_es_count(
    int P, const uint32_t* __restrict__ plan, uint32_t* vb0, uint32_t* vb1,
    const int* __restrict__ counts, int* __restrict__ counts_sorted, int* __restrict__ block_sums,
    unsigned long long* __restrict__ chunk_sums) {
  __shared__ int scratch[4];
  __shared__ uint32_t s_k;
  const DepthOrder ord = depth_order(plan, vb0, vb1, &s_k);
  const int s = blockIdx.x * kPreThreads + threadIdx.x;
  const int c = s < P ? counts[ord.at(s)] : 0;
  if (s < P) counts_sorted[s] = c;
  int total;
  block256_excl_scan(c, scratch, total);
  if (threadIdx.x == 0) {
    block_sums[blockIdx.x] = total;
    atomicAdd(&chunk_sums[blockIdx.x / kChunkBlocks], (unsigned long long)total);
  }
}

// Emit pass in depth order: (tile id, Gaussian) pairs, each Gaussian's in
// row-major tile order, Gaussians in (depth, id) order.
__global__ __launch_bounds__(kPreThreads) void ps_es_emit(
    FrameDev F, int nb, const uint32_t* __restrict__ plan, uint32_t* vb0, uint32_t* vb1,
    const int* __restrict__ counts_sorted, const int* __restrict__ block_sums,
    const unsigned long long* __restrict__ chunk_sums, const EmitRec* __restrict__ recs,
    const uint64_t* __restrict__ pool, int* __restrict__ state, int* __restrict__ curr_offset,
    int capacity, uint32_t* __restrict__ tiles, uint32_t* __restrict__ vals) {
  bind_camera(F);
  __shared__ uint32_t s_k;
  const DepthOrder ord = depth_order(plan, vb0, vb1, &s_k);
  const int s = blockIdx.x * kPreThreads + threadIdx.x;
  const int count = s < F.P ? counts_sorted[s] : 0;
  const int g = s < F.P ? ord.at(s) : F.P;
  const EmitRec r = count > 0 ? recs[g] : EmitRec{0u, 0u, 0u, 0u};
  emit_records_block(F, blockIdx.x, nb, g, count, r, block_sums, chunk_sums, pool, state,
                     curr_offset, capacity, Out32{tiles, vals});
}

struct PreWs {
  int* state;
  int* counts;
  int* block_sums;
  int* block_offsets;
  EmitRec* recs;                    // per Gaussian, written by the count pass
  unsigned long long* chunk_sums;   // pairs per chunk of kChunkBlocks workgroups
  uint64_t* pool;                   // hit-pool words (see EmitRec, kPoolPerBlock)
  long long pool_cap;               // words in all
  long long pool_overflow;          // words of the shared overflow area (the last ones)
  // depth-first pipeline: depth words, sort plan, depth-sort ping-pong
  // buffers, counts in depth order
  uint32_t* dkey;
  uint32_t* plan;
  uint32_t* kb[2];
  uint32_t* vb[2];
  int* counts_sorted;
};

// Workspace for `nb` preprocess workgroups (a single frame has
// ceil(P / 256); a batch the sum over its frames).
inline size_t pre_layout_blocks(int nb, char* base, PreWs* w) {
  size_t off = 0;
  const size_t b = (size_t)(nb > 0 ? nb : 1);
  if (w) w->state = reinterpret_cast<int*>(base + off);
  off += align_up(kStateWords * sizeof(int), 256);
  if (w) w->counts = reinterpret_cast<int*>(base + off);
  off += align_up(b * kPreThreads * sizeof(int), 256);
  if (w) w->block_sums = reinterpret_cast<int*>(base + off);
  off += align_up(b * sizeof(int), 256);
  if (w) w->block_offsets = reinterpret_cast<int*>(base + off);
  off += align_up(b * sizeof(int), 256);
  if (w) w->recs = reinterpret_cast<EmitRec*>(base + off);
  off += align_up(b * kPreThreads * sizeof(EmitRec), 256);
  if (w) w->chunk_sums = reinterpret_cast<unsigned long long*>(base + off);
  off += align_up((size_t)chunks_for((int)b) * sizeof(unsigned long long), 256);
  if (w) {
    w->pool = reinterpret_cast<uint64_t*>(base + off);
    w->pool_cap = (long long)pool_words_for((int)b);
    w->pool_overflow = (long long)pool_overflow_words((int)b);
  }
  off += align_up(pool_words_for((int)b) * sizeof(uint64_t), 256);
  const size_t g = b * kPreThreads * sizeof(uint32_t);  // one word per Gaussian slot
  if (w) w->dkey = reinterpret_cast<uint32_t*>(base + off);
  off += align_up(g, 256);
  if (w) w->plan = reinterpret_cast<uint32_t*>(base + off);
  off += align_up((size_t)2 * kPlanSlots * kPlanStride * sizeof(uint32_t), 256);
  for (int i = 0; i < 2; ++i) {
    if (w) w->kb[i] = reinterpret_cast<uint32_t*>(base + off);
    off += align_up(g, 256);
    if (w) w->vb[i] = reinterpret_cast<uint32_t*>(base + off);
    off += align_up(g, 256);
  }
  if (w) w->counts_sorted = reinterpret_cast<int*>(base + off);
  off += align_up(g, 256);
  return off;
}

inline int blocks_for(int P) { return (P + kPreThreads - 1) / kPreThreads; }

// Key emission from the count pass's records (two launches and a chunk
// reset) unless POLYSPLAT_EMIT=3pass selects count, scan and a re-projecting
// emit.  Both give the same keys; they differ only in what an overflow leaves
// behind (the three-launch form writes no pair at all).
inline bool record_emission() {
  const char* e = getenv("POLYSPLAT_EMIT");
  return !(e && e[0] == '3');
}

inline size_t pre_layout(int P, char* base, PreWs* w) {
  return pre_layout_blocks(blocks_for(P), base, w);
}

template <typename K1>
int launch_count(int sh_prec, int nb, hipStream_t stream, K1 launch) {
  if (nb <= 0) return PS_OK;
  if (sh_prec == PS_PREC_F32) launch(std::integral_constant<int, kShF32>{});
  else if (sh_prec == PS_PREC_F16) launch(std::integral_constant<int, kShF16>{});
  else launch(std::integral_constant<int, kShBF16>{});
  PS_LAUNCH_CHECK();
  return PS_OK;
}

}  // namespace

FrameDev make_frame_dev(int P, const float* positions, const void* shs, const float* opacities,
                        const float* cov3d, const ps_camera* cam, float* xy, float* rgbd,
                        float* conic, void* rec, int rec_prec, uint8_t* out, const float* bg) {
  FrameDev F;
  memset(&F, 0, sizeof(F));
  F.pos = positions;
  F.opac = opacities;
  F.cov3 = cov3d;
  F.shs = shs;
  F.xy = reinterpret_cast<float2*>(xy);
  F.rgbd = reinterpret_cast<float4*>(rgbd);
  F.conic = reinterpret_cast<float4*>(conic);
  F.rec = reinterpret_cast<FeatRec16*>(rec);
  F.out = out;
  F.P = P;
  F.rec_prec = rec_prec;
  F.cp = make_cam_params(cam);
  F.tiles = F.cp.grid_x * F.cp.grid_y;
  if (bg) {
    F.bg[0] = bg[0];
    F.bg[1] = bg[1];
    F.bg[2] = bg[2];
  }
  return F;
}

static int check_frame_args(int P, const void* positions, const void* shs, const void* opacities,
                            const void* cov3d, const ps_camera* cam) {
  PS_REQUIRE(P >= 0, "P must be >= 0 (got %d)", P);
  PS_REQUIRE(cam != nullptr, "camera is NULL");
  PS_REQUIRE(cam->width > 0 && cam->height > 0, "image size must be positive");
  PS_REQUIRE(cam->block_x > 0 && cam->block_y > 0, "tile size must be positive");
  const long long tiles = (long long)tiles_x(cam) * tiles_y(cam);
  PS_REQUIRE(tiles < (1LL << 31), "too many tiles");
  PS_REQUIRE(tiles_x(cam) < (1 << 16) && tiles_y(cam) < (1 << 16),
             "too many tile columns or rows (emit records hold 16 bits of each)");
  PS_REQUIRE(P == 0 || (positions && shs && opacities && cov3d), "scene pointer is NULL");
  return PS_OK;
}

// Internal entry used by ps_forward as well.
int preprocess_launch(int P, const float* positions, const void* shs, int sh_prec,
                      const float* opacities, const float* cov3d, const ps_camera* cam,
                      float* points_xy, float* rgb_depth, float* conic_opacity, void* feat_records,
                      int feat_prec, uint64_t* keys, uint32_t* values, int capacity,
                      int* curr_offset, void* workspace, size_t workspace_bytes,
                      hipStream_t stream, bool reset_offset, const CamParams* cam_dev) {
  int st = check_frame_args(P, positions, shs, opacities, cov3d, cam);
  if (st != PS_OK) return st;
  PS_REQUIRE(sh_prec == PS_PREC_F32 || sh_prec == PS_PREC_F16 || sh_prec == PS_PREC_BF16,
             "unknown SH precision %d", sh_prec);
  PS_REQUIRE(feat_records == nullptr || feat_prec == PS_PREC_F16 || feat_prec == PS_PREC_BF16,
             "feature records need F16 or BF16 precision");
  PS_REQUIRE(capacity >= 0, "capacity must be >= 0");
  PS_REQUIRE(curr_offset != nullptr, "curr_offset is NULL");
  PS_REQUIRE(workspace != nullptr && workspace_bytes >= ps_preprocess_workspace_size(P),
             "preprocess workspace too small (%zu < %zu)", workspace_bytes,
             ps_preprocess_workspace_size(P));
  PS_REQUIRE(capacity == 0 || (keys && values), "key/value buffers are NULL");

  PreWs w;
  pre_layout(P, static_cast<char*>(workspace), &w);
  FrameDev F = make_frame_dev(P, positions, shs, opacities, cov3d, cam, points_xy, rgb_depth,
                              conic_opacity, feat_records, feat_prec, nullptr, nullptr);
  F.cam_dev = cam_dev;
  const int nb = blocks_for(P);
  if (nb == 0) {
    // No Gaussians: the scan kernel alone publishes the (unchanged) offset.
    hipLaunchKernelGGL(ps_pre_scan, dim3(1), dim3(1024), 0, stream, 0, w.block_sums,
                       w.block_offsets, curr_offset, capacity, reset_offset ? 1 : 0, w.state,
                       (const FrameDev*)nullptr, 0, (int*)nullptr);
    PS_LAUNCH_CHECK();
    return PS_OK;
  }
  if (record_emission()) {
    const int nc = chunks_for(nb);
    hipLaunchKernelGGL(ps_pre_chunk_reset, dim3((nc + 255) / 256), dim3(256), 0, stream,
                       w.chunk_sums, nc,
                       reinterpret_cast<unsigned long long*>(w.state + kStatePool),
                       (uint32_t*)nullptr);
    PS_LAUNCH_CHECK();
    st = launch_count(sh_prec, nb, stream, [&](auto shp) {
      hipLaunchKernelGGL(ps_pre_count<decltype(shp)::value>, dim3(nb), dim3(kPreThreads), 0,
                         stream, F, w.counts, w.block_sums, w.recs, w.chunk_sums, w.state,
                         curr_offset, reset_offset ? 1 : 0, w.pool, w.pool_cap,
                         w.pool_overflow, (uint32_t*)nullptr, (uint32_t*)nullptr);
    });
    if (st != PS_OK) return st;
    hipLaunchKernelGGL(ps_pre_emit_rec, dim3(nb), dim3(kPreThreads), 0, stream, F, nb, w.counts,
                       w.block_sums, (const unsigned long long*)w.chunk_sums,
                       (const EmitRec*)w.recs, (const uint64_t*)w.pool, w.state, curr_offset,
                       capacity, keys, values);
    PS_LAUNCH_CHECK();
    return PS_OK;
  }
  st = launch_count(sh_prec, nb, stream, [&](auto shp) {
    hipLaunchKernelGGL(ps_pre_count<decltype(shp)::value>, dim3(nb), dim3(kPreThreads), 0, stream,
                       F, w.counts, w.block_sums, (EmitRec*)nullptr,
                       (unsigned long long*)nullptr, w.state, curr_offset, 0, (uint64_t*)nullptr,
                       0LL, 0LL, (uint32_t*)nullptr, (uint32_t*)nullptr);
  });
  if (st != PS_OK) return st;
  hipLaunchKernelGGL(ps_pre_scan, dim3(1), dim3(1024), 0, stream, nb, w.block_sums,
                     w.block_offsets, curr_offset, capacity, reset_offset ? 1 : 0, w.state,
                     (const FrameDev*)nullptr, 0, (int*)nullptr);
  PS_LAUNCH_CHECK();
  hipLaunchKernelGGL(ps_pre_emit, dim3(nb), dim3(kPreThreads), 0, stream, F, w.counts,
                     w.block_offsets, w.state, keys, values);
  PS_LAUNCH_CHECK();
  return PS_OK;
}

// Depth-first preprocess: count pass (features, records, hit pool, depth
// words, sort plan), the stable depth sort of the Gaussians, then the
// (tile id, Gaussian) pairs emitted in (depth, id) order into tiles/vals
// from offset 0.  The sort space supplies the depth sort's digit counts.
// The same state words as preprocess_launch are published (count, overflow).
int preprocess_es_launch(int P, const float* positions, const void* shs, int sh_prec,
                         const float* opacities, const float* cov3d, const ps_camera* cam,
                         float* points_xy, float* rgb_depth, float* conic_opacity,
                         void* feat_records, int feat_prec, uint32_t* tiles, uint32_t* vals,
                         int capacity, int* curr_offset, void* workspace, size_t workspace_bytes,
                         void* sort_space, size_t sort_space_bytes, int sort_max_pairs,
                         hipStream_t stream, const CamParams* cam_dev) {
  int st = check_frame_args(P, positions, shs, opacities, cov3d, cam);
  if (st != PS_OK) return st;
  PS_REQUIRE(sh_prec == PS_PREC_F32 || sh_prec == PS_PREC_F16 || sh_prec == PS_PREC_BF16,
             "unknown SH precision %d", sh_prec);
  PS_REQUIRE(capacity >= 0 && curr_offset != nullptr, "bad key buffers");
  PS_REQUIRE(capacity == 0 || (tiles && vals), "key/value buffers are NULL");
  PS_REQUIRE(workspace != nullptr && workspace_bytes >= ps_preprocess_workspace_size(P),
             "preprocess workspace too small");
  PreWs w;
  pre_layout(P, static_cast<char*>(workspace), &w);
  FrameDev F = make_frame_dev(P, positions, shs, opacities, cov3d, cam, points_xy, rgb_depth,
                              conic_opacity, feat_records, feat_prec, nullptr, nullptr);
  F.cam_dev = cam_dev;
  const int nb = blocks_for(P);
  if (nb == 0) {
    hipLaunchKernelGGL(ps_pre_scan, dim3(1), dim3(1024), 0, stream, 0, w.block_sums,
                       w.block_offsets, curr_offset, capacity, 1, w.state,
                       (const FrameDev*)nullptr, 0, (int*)nullptr);
    PS_LAUNCH_CHECK();
    return PS_OK;
  }
  const int nc = chunks_for(nb);
  hipLaunchKernelGGL(ps_pre_chunk_reset, dim3((nc + 255) / 256), dim3(256), 0, stream,
                     w.chunk_sums, nc, reinterpret_cast<unsigned long long*>(w.state + kStatePool),
                     w.plan);
  PS_LAUNCH_CHECK();
  st = launch_count(sh_prec, nb, stream, [&](auto shp) {
    hipLaunchKernelGGL(ps_pre_count<decltype(shp)::value>, dim3(nb), dim3(kPreThreads), 0, stream,
                       F, w.counts, w.block_sums, w.recs, (unsigned long long*)nullptr, w.state,
                       curr_offset, 1, w.pool, w.pool_cap, w.pool_overflow, w.dkey, w.plan);
  });
  if (st != PS_OK) return st;
  DepthSortArgs A;
  A.plan = w.plan;
  A.words = w.dkey;
  A.kb[0] = w.kb[0];
  A.kb[1] = w.kb[1];
  A.vb[0] = w.vb[0];
  A.vb[1] = w.vb[1];
  A.n = P;
  st = depth_sort_launch(A, sort_max_pairs, sort_space, sort_space_bytes, stream);
  if (st != PS_OK) return st;
  hipLaunchKernelGGL(ps_es_count, dim3(nb), dim3(kPreThreads), 0, stream, P,
                     (const uint32_t*)w.plan, w.vb[0], w.vb[1], (const int*)w.counts,
                     w.counts_sorted, w.block_sums, w.chunk_sums);
  PS_LAUNCH_CHECK();
  hipLaunchKernelGGL(ps_es_emit, dim3(nb), dim3(kPreThreads), 0, stream, F, nb,
                     (const uint32_t*)w.plan, w.vb[0], w.vb[1], (const int*)w.counts_sorted,
                     (const int*)w.block_sums, (const unsigned long long*)w.chunk_sums,
                     (const EmitRec*)w.recs, (const uint64_t*)w.pool, w.state, curr_offset,
                     capacity, tiles, vals);
  PS_LAUNCH_CHECK();
  return PS_OK;
}

// Pinned staging ring for batch descriptors, one per host thread.  A copy
// from pageable memory would block the host until the stream reaches it;
// from a pinned slot it is queued like a kernel.  A slot is rewritten only
// after the copy that last read it has executed (its event), so up to
// kDescSlots batches can be queued ahead of the device.
struct DescRing {
  static constexpr int kDescSlots = 16;
  FrameDev* host = nullptr;
  hipEvent_t ev[kDescSlots] = {};
  bool used[kDescSlots] = {};
  int next = 0;
  bool ok = false;
  DescRing() {
    if (hipHostMalloc(reinterpret_cast<void**>(&host),
                      sizeof(FrameDev) * PS_MAX_BATCH_FRAMES * kDescSlots,
                      hipHostMallocDefault) != hipSuccess) {
      host = nullptr;
      return;
    }
    for (int i = 0; i < kDescSlots; ++i)
      if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) return;
    ok = true;
  }
  ~DescRing() {
    for (int i = 0; i < kDescSlots; ++i)
      if (ev[i]) (void)hipEventDestroy(ev[i]);
    if (host) (void)hipHostFree(host);
  }
};

int upload_descriptors(FrameDev* dst, const FrameDev* src, int nf, hipStream_t stream) {
  PS_REQUIRE(nf > 0 && nf <= PS_MAX_BATCH_FRAMES, "batch of %d frames", nf);
  static thread_local DescRing ring;
  if (!ring.ok) {
    PS_HIP_CHECK(hipMemcpy(dst, src, nf * sizeof(FrameDev), hipMemcpyHostToDevice));
    return PS_OK;
  }
  const int i = ring.next;
  ring.next = (i + 1) % DescRing::kDescSlots;
  if (ring.used[i]) PS_HIP_CHECK(hipEventSynchronize(ring.ev[i]));
  FrameDev* slot = ring.host + (size_t)i * PS_MAX_BATCH_FRAMES;
  memcpy(slot, src, nf * sizeof(FrameDev));
  PS_HIP_CHECK(hipMemcpyAsync(dst, slot, nf * sizeof(FrameDev), hipMemcpyHostToDevice, stream));
  PS_HIP_CHECK(hipEventRecord(ring.ev[i], stream));
  ring.used[i] = true;
  return PS_OK;
}

size_t preprocess_batch_workspace_size(int nf, const int* P) {
  int nb = 0;
  for (int f = 0; f < nf; ++f) nb += blocks_for(P[f] > 0 ? P[f] : 0);
  return align_up(pre_layout_blocks(nb, nullptr, nullptr), 256) +
         align_up((size_t)(nf > 0 ? nf : 1) * sizeof(FrameDev), 256) +
         align_up((size_t)(nf > 0 ? nf : 1) * sizeof(int), 256);
}

// Batched preprocess.  Fills in each frame's first workgroup (block0) and
// first global tile id (tile0), consecutive in frame order, copies the
// descriptors into the workspace and emits the keys of all frames into one
// array with global tile ids.  Returns the total tile count in *tiles_out.
int preprocess_batch_launch(int nf, FrameDev* frames_host, int sh_prec, uint64_t* keys,
                            uint32_t* values, int capacity, int* total_out, void* workspace,
                            size_t workspace_bytes, int** frame_counts_dev, const int** state_count,
                            int* tiles_out, hipStream_t stream) {
  PS_REQUIRE(nf > 0, "empty batch");
  PS_REQUIRE(sh_prec == PS_PREC_F32 || sh_prec == PS_PREC_F16 || sh_prec == PS_PREC_BF16,
             "unknown SH precision %d", sh_prec);
  PS_REQUIRE(capacity == 0 || (keys && values), "key/value buffers are NULL");
  PS_REQUIRE(total_out != nullptr, "curr_offset is NULL");
  int nb = 0;
  long long tiles = 0;
  for (int f = 0; f < nf; ++f) {
    frames_host[f].block0 = nb;
    frames_host[f].tile0 = (int)tiles;
    nb += blocks_for(frames_host[f].P);
    tiles += frames_host[f].tiles;
    PS_REQUIRE(tiles < (1LL << 31), "too many tiles in the batch");
  }
  if (tiles_out) *tiles_out = (int)tiles;
  const size_t pre_bytes = align_up(pre_layout_blocks(nb, nullptr, nullptr), 256);
  PS_REQUIRE(workspace != nullptr &&
                 workspace_bytes >= pre_bytes + align_up(nf * sizeof(FrameDev), 256) +
                                        align_up(nf * sizeof(int), 256),
             "batch workspace too small");
  char* ws = static_cast<char*>(workspace);
  PreWs w;
  pre_layout_blocks(nb, ws, &w);
  FrameDev* dframes = reinterpret_cast<FrameDev*>(ws + pre_bytes);
  int* fcounts = reinterpret_cast<int*>(ws + pre_bytes + align_up(nf * sizeof(FrameDev), 256));
  int st = upload_descriptors(dframes, frames_host, nf, stream);
  if (st != PS_OK) return st;
  st = launch_count(sh_prec, nb, stream, [&](auto shp) {
    hipLaunchKernelGGL(ps_pre_count_batch<decltype(shp)::value>, dim3(nb), dim3(kPreThreads), 0,
                       stream, (const FrameDev*)dframes, nf, w.counts, w.block_sums);
  });
  if (st != PS_OK) return st;
  hipLaunchKernelGGL(ps_pre_scan, dim3(1), dim3(1024), 0, stream, nb, w.block_sums,
                     w.block_offsets, total_out, capacity, 1, w.state, (const FrameDev*)dframes, nf,
                     fcounts);
  PS_LAUNCH_CHECK();
  if (nb > 0) {
    hipLaunchKernelGGL(ps_pre_emit_batch, dim3(nb), dim3(kPreThreads), 0, stream,
                       (const FrameDev*)dframes, nf, w.counts, w.block_offsets, w.state, keys,
                       values);
    PS_LAUNCH_CHECK();
  }
  if (frame_counts_dev) *frame_counts_dev = fcounts;
  if (state_count) *state_count = w.state + kStateCount;
  return PS_OK;
}

const FrameDev* batch_frames_dev(int nf, const FrameDev* frames_host, void* workspace) {
  int nb = 0;
  for (int f = 0; f < nf; ++f) nb += blocks_for(frames_host[f].P);
  return reinterpret_cast<const FrameDev*>(static_cast<char*>(workspace) +
                                           align_up(pre_layout_blocks(nb, nullptr, nullptr), 256));
}

int* batch_counts_dev(int nf, const int* P, void* workspace) {
  int nb = 0;
  for (int f = 0; f < nf; ++f) nb += blocks_for(P[f] > 0 ? P[f] : 0);
  return reinterpret_cast<int*>(static_cast<char*>(workspace) +
                                align_up(pre_layout_blocks(nb, nullptr, nullptr), 256) +
                                align_up((size_t)nf * sizeof(FrameDev), 256));
}

int check_frame(const FrameDev& F, const ps_camera* cam) {
  return check_frame_args(F.P, F.pos, F.shs, F.opac, F.cov3, cam);
}

}  // namespace ps

extern "C" size_t ps_preprocess_workspace_size(int P) {
  return ps::pre_layout(P < 0 ? 0 : P, nullptr, nullptr);
}

extern "C" const int* ps_preprocess_count_ptr(const void* workspace) {
  return static_cast<const int*>(workspace) + ps::kStateCount;
}

extern "C" const int* ps_preprocess_status_ptr(const void* workspace) {
  return static_cast<const int*>(workspace) + ps::kStateFlags;
}

extern "C" int ps_preprocess(int P, const float* positions, const void* shs, int sh_prec,
                             const float* opacities, const float* cov3d, const ps_camera* cam,
                             float* points_xy, float* rgb_depth, float* conic_opacity,
                             void* feat_records, int feat_prec, uint64_t* keys_unsorted,
                             uint32_t* values_unsorted, int capacity, int* curr_offset,
                             void* workspace, size_t workspace_bytes, ps_stream_t stream) {
  return ps::preprocess_launch(P, positions, shs, sh_prec, opacities, cov3d, cam, points_xy,
                               rgb_depth, conic_opacity, feat_records, feat_prec, keys_unsorted,
                               values_unsorted, capacity, curr_offset, workspace, workspace_bytes,
                               reinterpret_cast<hipStream_t>(stream), false, nullptr);
}
