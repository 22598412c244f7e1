// Preprocess for the MI355X forward rasterizer: projection, EWA conic, SH
// colour and deterministic tile-key emission.
//
// The reference (preprocess.cu:313-641) emits keys with one global atomic
// counter, so the unsorted order -- and the order of equal keys after the
// stable sort -- changes from run to run.  Here emission is split in three
// launches so that the key array comes out in a canonical order (Gaussian id,
// then tile row-major) and the exact pair count is known on the device before
// anything is written:
//   1. ps_pre_count : project every Gaussian, count the tiles its ellipse
//                     really touches, write its render features, and reduce
//                     the counts per 256-Gaussian block;
//   2. ps_pre_scan  : scan the block sums, publish the pair count, detect
//                     overflow;
//   3. ps_pre_emit  : re-project the Gaussians that emit (bit-identical, so
//                     the same tiles pass), and write keys at their offsets.
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

#include <string.h>

#include <type_traits>

namespace ps {
namespace {

constexpr int kPreThreads = 256;
constexpr int kMassive = 64;

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

// Massive Gaussians: the wave walks each one's rect 64 tiles at a time in
// row-major order.  Returns this lane's own tile count (count mode); in emit
// mode writes keys at off + running index.
template <bool EMIT>
__device__ int sweep_massive(const Cand& me, bool is_massive, int lane, int gid, int off, int cnt,
                             int bx, int by, int grid_x, int tile0, uint64_t* keys,
                             uint32_t* vals) {
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
    const int qoff = EMIT ? bcast(off, i) : 0;
    const int qcnt = EMIT ? bcast(cnt, i) : 0;
    const int qgid = bcast(gid, i);
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
      const int k = run + popc64(m & below);
      // The count pass found qcnt tiles with the same arithmetic; the bound
      // keeps every write inside this Gaussian's slots regardless.
      if (EMIT && ok && k < qcnt) {
        const int slot = qoff + k;
        keys[slot] = make_key(tile0 + (q.ry0 + ty) * grid_x + (q.rx0 + tx), q.depth);
        vals[slot] = (uint32_t)qgid;
      }
      run += popc64(m);
    }
    if (lane == i) mine = run;
  }
  return mine;
}

// Fragmented Gaussians: all lanes' small rects become one stream of
// (Gaussian, tile) tasks, 64 per step; the owner of task t is the first lane
// whose inclusive prefix of areas exceeds t.
template <bool EMIT>
__device__ int sweep_fragmented(const Cand& me, bool is_frag, int lane, int gid, int off, int cnt,
                                int bx, int by, int grid_x, int tile0, uint64_t* keys,
                                uint32_t* vals) {
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
    const int o_gid = permute(gid, o);
    const int o_off = EMIT ? permute(off, o) : 0;
    const int o_done = EMIT ? permute(mine, o) : 0;
    const int o_cnt = EMIT ? permute(cnt, o) : 0;
    bool ok = false;
    int tx = 0, ty = 0;
    if (tv) {
      const int w = q.rx1 - q.rx0;
      const int local = t - o_excl;
      ty = local / w;
      tx = local - ty * w;
      ok = tile_hits(q.rx0 + tx, q.ry0 + ty, proj_from(q), bx, by);
    }
    const uint64_t m = ballot(ok);
    const int seg_lo = imax(o_excl - t0, 0);
    const int k = o_done + popc64(m & below & ~lanes_below(seg_lo));
    if (EMIT && ok && k < o_cnt) {
      const int slot = o_off + k;
      keys[slot] = make_key(tile0 + (q.ry0 + ty) * grid_x + (q.rx0 + tx), q.depth);
      vals[slot] = (uint32_t)o_gid;
    }
    // Each owner adds the passing tasks of its own segment in this step.
    if (is_frag) {
      const int lo_l = imax(excl - t0, 0), hi_l = imin(incl - t0, kWave);
      if (lo_l < hi_l) mine += popc64(m & lane_range(lo_l, hi_l));
    }
  }
  return mine;
}

__device__ __forceinline__ void load_cov(const float* cov3, int g, float c[6]) {
  const float2* p = reinterpret_cast<const float2*>(cov3 + (size_t)g * 6);
  float2 a = p[0], b = p[1], d = p[2];
  c[0] = a.x; c[1] = a.y; c[2] = b.x; c[3] = b.y; c[4] = d.x; c[5] = d.y;
}

template <int SHP>
__device__ __forceinline__ void load_sh(const void* shs, int g, float sh[48]) {
  if (SHP == kShF32) {
    const float4* p = reinterpret_cast<const float4*>(static_cast<const float*>(shs) + (size_t)g * 48);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      float4 v = p[k];
      sh[4 * k + 0] = v.x; sh[4 * k + 1] = v.y; sh[4 * k + 2] = v.z; sh[4 * k + 3] = v.w;
    }
  } else {
    const uint4* p = reinterpret_cast<const uint4*>(static_cast<const uint16_t*>(shs) + (size_t)g * 48);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      uint4 v = p[k];
      uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        uint16_t lo = (uint16_t)(w4[j] & 0xffffu), hi = (uint16_t)(w4[j] >> 16);
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
}

__device__ __forceinline__ uint16_t to_half_bits(float f, int prec) {
  return prec == kFeatBF16 ? f32_to_bf16(f) : __half_as_ushort(__float2half_rn(f));
}

// Count pass for the workgroup `vb` of a launch (vb - F.block0 is the
// frame-local workgroup).  counts/block_sums are indexed by launch position.
template <int SHP>
__device__ __forceinline__ void pre_count_body(const FrameDev& F, int vb, int* __restrict__ counts,
                                               int* __restrict__ block_sums) {
  __shared__ int scratch[4];
  const int tid = threadIdx.x, lane = tid & 63;
  const int g = (vb - F.block0) * kPreThreads + tid;
  const CamParams& cp = F.cp;
  Proj pj;
  pj.valid = false;
  float px = 0.f, py = 0.f, pz = 0.f;
  if (g < F.P) {
    px = F.pos[3 * (size_t)g + 0];
    py = F.pos[3 * (size_t)g + 1];
    pz = F.pos[3 * (size_t)g + 2];
    float c[6];
    load_cov(F.cov3, g, c);
    pj = project(px, py, pz, F.opac[g], c, cp);
  }
  const int area = pj.valid ? (pj.rmax_x - pj.rmin_x) * (pj.rmax_y - pj.rmin_y) : 0;
  int count = 0;
  if (area == 1) count = tile_hits(pj.rmin_x, pj.rmin_y, pj, cp.block_x, cp.block_y) ? 1 : 0;
  const Cand me = cand_from(pj);
  count += sweep_massive<false>(me, area >= kMassive, lane, g, 0, 0, cp.block_x, cp.block_y,
                                cp.grid_x, 0, nullptr, nullptr);
  count += sweep_fragmented<false>(me, area >= 2 && area < kMassive, lane, g, 0, 0, cp.block_x,
                                   cp.block_y, cp.grid_x, 0, nullptr, nullptr);
  counts[(size_t)vb * kPreThreads + tid] = count;
  if (count > 0) {
    const float4 con = make_float4((-0.5f * kLog2e) * pj.con_a, -kLog2e * pj.con_b,
                                   (-0.5f * kLog2e) * pj.con_c, pj.log2_opacity);
    float sh[48];
    load_sh<SHP>(F.shs, g, sh);
    const F3 rgb = sh_color(px, py, pz, cp.campos, sh);
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
  int total;
  block256_excl_scan(count, scratch, total);
  if (tid == 0) block_sums[vb] = total;
}

template <int SHP>
__global__ __launch_bounds__(kPreThreads) void ps_pre_count(FrameDev F, int* __restrict__ counts,
                                                            int* __restrict__ block_sums) {
  pre_count_body<SHP>(F, blockIdx.x, counts, block_sums);
}

__device__ __forceinline__ int frame_of_block(const FrameDev* frames, int nf, int vb) {
  return frame_search(frames, nf, vb, [](const FrameDev& f) { return f.block0; });
}

template <int SHP>
__global__ __launch_bounds__(kPreThreads) void ps_pre_count_batch(const FrameDev* __restrict__ frames,
                                                                  int nf, int* __restrict__ counts,
                                                                  int* __restrict__ block_sums) {
  const int f = frame_of_block(frames, nf, blockIdx.x);
  pre_count_body<SHP>(frames[f], blockIdx.x, counts, block_sums);
}

// state[0] = offset the frame's keys start at, state[1] = pair count for the
// later stages (0 after overflow), state[2] = status bits (bit 0: overflow).
constexpr int kStateBase = 0, kStateCount = 1, kStateFlags = 2, kStateWords = 64;

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
    const bool overflow = end > (long long)capacity || end > 0x7fffffffLL;
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
  Proj pj;
  pj.valid = false;
  if (count > 0) {
    float c[6];
    load_cov(F.cov3, g, c);
    pj = project(F.pos[3 * (size_t)g], F.pos[3 * (size_t)g + 1], F.pos[3 * (size_t)g + 2],
                 F.opac[g], c, cp);
  }
  const int area = (count > 0 && pj.valid) ? (pj.rmax_x - pj.rmin_x) * (pj.rmax_y - pj.rmin_y) : 0;
  if (area == 1) {
    keys[off] = make_key(F.tile0 + pj.rmin_y * cp.grid_x + pj.rmin_x, pj.depth);
    vals[off] = (uint32_t)g;
  }
  const Cand me = cand_from(pj);
  sweep_massive<true>(me, area >= kMassive, lane, g, off, count, cp.block_x, cp.block_y,
                      cp.grid_x, F.tile0, keys, vals);
  sweep_fragmented<true>(me, area >= 2 && area < kMassive, lane, g, off, count, cp.block_x,
                         cp.block_y, cp.grid_x, F.tile0, keys, vals);
}

__global__ __launch_bounds__(kPreThreads) void ps_pre_emit(FrameDev F, const int* __restrict__ counts,
                                                           const int* __restrict__ block_offsets,
                                                           const int* __restrict__ state,
                                                           uint64_t* __restrict__ keys,
                                                           uint32_t* __restrict__ vals) {
  pre_emit_body(F, blockIdx.x, counts, block_offsets, state, keys, vals);
}

__global__ __launch_bounds__(kPreThreads) void ps_pre_emit_batch(
    const FrameDev* __restrict__ frames, int nf, const int* __restrict__ counts,
    const int* __restrict__ block_offsets, const int* __restrict__ state,
    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int f = frame_of_block(frames, nf, blockIdx.x);
  pre_emit_body(frames[f], blockIdx.x, counts, block_offsets, state, keys, vals);
}

struct PreWs {
  int* state;
  int* counts;
  int* block_sums;
  int* block_offsets;
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
  return off;
}

inline int blocks_for(int P) { return (P + kPreThreads - 1) / kPreThreads; }

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
  PS_REQUIRE(P == 0 || (positions && shs && opacities && cov3d), "scene pointer is NULL");
  return PS_OK;
}

// Internal entry used by ps_forward as well.
int preprocess_launch(int P, const float* positions, const void* shs, int sh_prec,
                      const float* opacities, const float* cov3d, const ps_camera* cam,
                      float* points_xy, float* rgb_depth, float* conic_opacity, void* feat_records,
                      int feat_prec, uint64_t* keys, uint32_t* values, int capacity,
                      int* curr_offset, void* workspace, size_t workspace_bytes,
                      hipStream_t stream, bool reset_offset) {
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
  const FrameDev F = make_frame_dev(P, positions, shs, opacities, cov3d, cam, points_xy, rgb_depth,
                                    conic_opacity, feat_records, feat_prec, nullptr, nullptr);
  const int nb = blocks_for(P);
  st = launch_count(sh_prec, nb, stream, [&](auto shp) {
    hipLaunchKernelGGL(ps_pre_count<decltype(shp)::value>, dim3(nb), dim3(kPreThreads), 0, stream,
                       F, w.counts, w.block_sums);
  });
  if (st != PS_OK) return st;
  hipLaunchKernelGGL(ps_pre_scan, dim3(1), dim3(1024), 0, stream, nb, w.block_sums,
                     w.block_offsets, curr_offset, capacity, reset_offset ? 1 : 0, w.state,
                     (const FrameDev*)nullptr, 0, (int*)nullptr);
  PS_LAUNCH_CHECK();
  if (nb > 0) {
    hipLaunchKernelGGL(ps_pre_emit, dim3(nb), dim3(kPreThreads), 0, stream, F, w.counts,
                       w.block_offsets, w.state, keys, values);
    PS_LAUNCH_CHECK();
  }
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
                               reinterpret_cast<hipStream_t>(stream), false);
}
