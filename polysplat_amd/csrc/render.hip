// Tile renderers for gfx950: one 64-lane wave per 16x16 tile, one pixel per
// lane in each 8x8 quadrant, front-to-back alpha blending in FP32.
//
// Per pixel (render.cu:228-271): d = centre - pixel (integer pixel coords),
//   power = log2(o) + c.a*dx^2 + c.b*dx*dy + c.c*dy^2  (conic pre-scaled by
//   -0.5*log2(e), -log2(e), -0.5*log2(e)),  alpha = 2^power,
//   C += rgb * alpha * T,  T -= alpha * T.
// The tile stops once every pixel has T < 1e-4; that test runs every
// kCheck Gaussians.  Output (render.cu:375-394): T < 1e-4 -> 0,
//   u8 = trunc(clamp((C + T*bg) * 255, 0, 255)); an empty tile writes bg.
//
// Features reach the shading loop through LDS: each lane stages one
// Gaussian of the next 64-Gaussian batch (loads issued a batch ahead, so they
// overlap the current batch's shading), converting reduced-precision records
// to FP32 once per Gaussian instead of once per lane; the shading loop then
// reads wave-uniform LDS addresses (broadcast, no bank conflicts).
//
// Scheduling: the plain renderer launches one workgroup per tile
// (render.cu:4365-4385).  The persistent renderer launches as many one-wave
// workgroups as fit on the device and pulls chunks of tiles from atomic
// counters, one per XCD, in a caller-owned queue area (render.cu:1717-1864
// uses one __device__ global counter); the last workgroup to finish zeroes
// the queue, so the next launch needs no reset.

#include "ps_common.h"
#include "ps_host.h"
#include "ps_wave.h"

#include <stdlib.h>

#include <mutex>

// Tile-queue words of ps_render_16x16_persistent_noalloc: 2 KiB, the same
// area a caller hands ps_render_16x16_persistent as `queue`.  Each device
// holds its own copy for the life of the process.  Host code reaches it only
// through the address hipGetSymbolAddress returns for the current device
// (see noalloc_queue below); the array's name is not a device address.
static __device__ __align__(256) unsigned scratch[512];

namespace ps {
namespace {

constexpr int kTile = 16;
constexpr int kBatch = 64;
constexpr float kTMin = 0.0001f;

// Range of tile `t` in the sorted keys, found by a cooperative 64-ary search
// (ceil(log64(n)) dependent loads per bound).
__device__ __forceinline__ int wave_bound(const uint64_t* keys, int n, uint32_t t, bool upper,
                                          int lane) {
  int lo = 0, hi = n;
  while (hi - lo > kWave) {
    const int stride = (hi - lo + kWave - 1) / kWave;
    const int pos = lo + lane * stride;
    bool below = false;
    if (pos < hi) {
      const uint32_t kt = (uint32_t)(keys[pos] >> 32);
      below = upper ? (kt <= t) : (kt < t);
    }
    const int k = popc64(ballot(below));
    if (k == 0) return lo;
    const int nlo = lo + (k - 1) * stride + 1;
    int nhi = lo + k * stride;
    if (nhi > hi) nhi = hi;
    lo = nlo;
    hi = nhi;
  }
  const int pos = lo + lane;
  bool below = false;
  if (pos < hi) {
    const uint32_t kt = (uint32_t)(keys[pos] >> 32);
    below = upper ? (kt <= t) : (kt < t);
  }
  return lo + popc64(ballot(below));
}

// Nine FP32 values per Gaussian as the shading loop wants them.
struct Feat {
  float x, y, w, a, b, c, r, g, bl;
};

// Stands in for the slots past the end of a tile's list.
__device__ __forceinline__ Feat neutral_feat() {
  Feat f;
  f.x = f.y = f.a = f.b = f.c = f.r = f.g = f.bl = 0.0f;
  f.w = -__builtin_inff();
  return f;
}

template <int PREC>
__device__ __forceinline__ Feat load_feat(int idx, const float2* __restrict__ xy,
                                          const float4* __restrict__ rgbd,
                                          const float4* __restrict__ con,
                                          const FeatRec16* __restrict__ rec, const TopKCache& tk) {
  Feat f;
  if (PREC == kFeatF32) {
    if (tk.xy != nullptr && idx < 0) {  // a cached Gaussian (bit 31 set)
      const int slot = idx & 0x7fffffff;
      if (slot >= tk.n) return neutral_feat();  // not a slot of the cache: nothing
      xy = tk.xy;
      rgbd = tk.rgbd;
      con = tk.con;
      idx = slot;
    }
    const float2 p = xy[idx];
    const float4 q = con[idx];
    const float4 c = rgbd[idx];
    f.x = p.x; f.y = p.y;
    f.a = q.x; f.b = q.y; f.c = q.z; f.w = q.w;
    f.r = c.x; f.g = c.y; f.bl = c.z;
  } else {
    const uint2* p = reinterpret_cast<const uint2*>(rec + idx);
    const uint2 u0 = p[0], u1 = p[1], u2 = p[2];
    f.x = __uint_as_float(u0.x);
    f.y = __uint_as_float(u0.y);
    f.w = __uint_as_float(u1.x);
    const uint16_t h[6] = {(uint16_t)(u1.y & 0xffffu), (uint16_t)(u1.y >> 16),
                           (uint16_t)(u2.x & 0xffffu), (uint16_t)(u2.x >> 16),
                           (uint16_t)(u2.y & 0xffffu), (uint16_t)(u2.y >> 16)};
    float v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k)
      v[k] = (PREC == kFeatBF16) ? bf16_to_f32(h[k]) : __half2float(__ushort_as_half(h[k]));
    const float inv = 1.0f / kConicScale;
    f.a = v[0] * inv; f.b = v[1] * inv; f.c = v[2] * inv;
    f.r = v[3]; f.g = v[4]; f.bl = v[5];
  }
  return f;
}

__device__ __forceinline__ uint8_t encode(float x) {
  return (uint8_t)fminf(fmaxf(0.0f, x * 255.0f), 255.0f);
}

// Quadrant bit k (k = qy * 2 + qx) of a Gaussian: set when its alpha can
// reach 2^kQuadSkipPower at some pixel of quadrant k (8x8 pixels) of the tile
// centred on (cx, cy) (its first pixel + 7.5).  The power is w + q(d) with q
// a negative semidefinite quadratic form, and q(d) <= lmax |d|^2 for lmax the
// eigenvalue of q nearest zero (taken with a relative rounding allowance), so
// w + lmax * dist^2(centre, quadrant) bounds the power over the quadrant.
// A Gaussian left out of a quadrant has alpha below 2^-24 at each of its
// pixels: it would have added less than 2^-24 T of colour (T <= 1), under a
// ten-thousandth of the 1/255 output step, and changed T by at most one unit
// in its last place.  The output may differ from blending every Gaussian by
// one level where its scaled value lies that close to an integer (measured:
// 11 of 64M pixels of the 56M-Gaussian city, none at 6M Gaussians
// 3840x2160); against 2^-32 the tighter masks make the sparse path 3-5%
// faster on frames of small splats.  The neutral padding Gaussian (power
// -inf) has no bits; a NaN keeps every bit.
constexpr float kQuadSkipPower = -24.0f;

__device__ __forceinline__ float dist_to_span(float v, float lo, float hi) {
  return v < lo ? lo - v : (v > hi ? v - hi : 0.0f);
}

__device__ __forceinline__ uint32_t quad_mask(const Feat& f, float cx, float cy) {
  const float h = 0.5f * (f.a - f.c);
  const float root = __builtin_sqrtf(h * h + 0.25f * f.b * f.b);
  float lmax = 0.5f * (f.a + f.c) + root;
  lmax += 1e-6f * (fabsf(f.a) + fabsf(f.b) + fabsf(f.c));
  lmax = fminf(lmax, 0.0f);
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float x0 = cx + (8.0f * (float)(k & 1) - 7.5f);
    const float y0 = cy + (8.0f * (float)(k >> 1) - 7.5f);
    const float ddx = dist_to_span(f.x, x0, x0 + 7.0f);
    const float ddy = dist_to_span(f.y, y0, y0 + 7.0f);
    const float bound = f.w + lmax * (ddx * ddx + ddy * ddy);
    if (!(bound < kQuadSkipPower)) m |= 1u << k;
  }
  return m;
}

// The power as a polynomial in the pixel's offset (u, v) from the tile's
// centre (cx, cy) = tile origin + 7.5: with X = x - cx, Y = y - cy,
//   power = F + u (D + a u + b v) + v (E + c v),
//   F = w + X (a X + b Y) + c Y^2,  D = -(2 a X + b Y),  E = -(b X + 2 c Y),
// which is w + a dx^2 + b dx dy + c dy^2 for dx = X - u, dy = Y - v expanded.
// Staged once per (Gaussian, tile), it costs five FMAs per pixel, and for a
// lane's four pixels the row terms (per v) and column terms (per u) are
// shared.  Rounding: F is the power at the tile centre; its magnitude is
// bounded by the conic (at most about 2.4 per square pixel after the
// reference's 0.3 low-pass dilation) times the squared distance, so the
// cancellation error stays near 1e-4 in log2 units, far below 1/255 of
// colour.  The neutral Gaussian stages F = -inf with D = E = 0.
struct Staged {
  float4 A;  // (F, D, E, a)
  float4 B;  // (b, c, r, g)
  float bl;
};

__device__ __forceinline__ Staged stage_feat(const Feat& f, float cx, float cy) {
  const float X = f.x - cx, Y = f.y - cy;
  Staged st;
  st.A.x = fmaf(X, fmaf(f.a, X, f.b * Y), fmaf(f.c * Y, Y, f.w));
  st.A.y = -fmaf(2.0f * f.a, X, f.b * Y);
  st.A.z = -fmaf(f.b, X, 2.0f * f.c * Y);
  st.A.w = f.a;
  st.B = make_float4(f.b, f.c, f.r, f.g);
  st.bl = f.bl;
  return st;
}

// The power at offset (u, v) from the tile centre; the dense and the sparse
// path evaluate exactly this sequence of operations.
__device__ __forceinline__ float staged_power(const float4& A, const float4& B, float u,
                                              float v) {
  return fmaf(u, fmaf(B.x, v, fmaf(A.w, u, A.y)), fmaf(fmaf(B.y, v, A.z), v, A.x));
}

// Relative costs of the two shading paths (per Gaussian for the dense one,
// per live (Gaussian, quadrant) for the sparse one), from their VALU counts
// (about 39 and 12) with the sparse one weighted up: its LDS reads and
// scalar loop control are not free.  14 was measured best among 10..28
// (6M Gaussians at 3840x2160: render 0.573 ms at 12, 0.547-0.552 at 14,
// 0.577 at 17; 30M Gaussians at 3840x2160 2.99 / 2.88-2.96 / 2.96).  A
// dense-only kernel (66 VGPRs) is 5-8% faster on those two scenes but 33%
// slower on the 56M-Gaussian city at 9216x6912, whose small splats leave
// most groups sparse.
constexpr int kDenseCost = 39;
constexpr int kSparseCost = 14;
constexpr int kSparseGroupCost = 2;

// Shared per-tile body.  Returns nothing; writes the tile's pixels.
// GATHERED: features indexed by pair position (no id indirection); CHECK:
// early-exit test cadence; DT: skip pixels whose alpha is below 1/255.
//
// Lane l holds one pixel in each 8x8 quadrant of the tile: (l % 8, l / 8)
// plus (8 qx, 8 qy).  Every CHECK Gaussians of the list the wave picks the
// cheaper of two paths from the group's quadrant masks: dense (each Gaussian
// shades all four pixels of every lane; the per-row and per-column terms of
// the quadratic form are shared) or sparse (for each quadrant, only the
// Gaussians live there, one pixel per lane).  Both evaluate the same
// expression per pixel, in list order, so the path never changes a result.
template <int PREC, bool GATHERED, int CHECK, bool DT>
__device__ __forceinline__ void render_tile(
    int tile, int start, int end, int W, int H, int gx, const uint32_t* __restrict__ order,
    const float2* __restrict__ xy, const float4* __restrict__ rgbd, const float4* __restrict__ con,
    const FeatRec16* __restrict__ rec, const TopKCache& tk, float3 bg,
    uint8_t* __restrict__ out, float4 (*sA)[kBatch], float4 (*sB)[kBatch], float (*sC)[kBatch],
    int lane) {
  static_assert(kBatch % CHECK == 0, "a batch holds whole check groups");
  const int tx = tile % gx, ty = tile / gx;
  // The tile's centre, and this lane's pixel offsets from it.
  const float cx = (float)(tx * kTile) + 7.5f, cy = (float)(ty * kTile) + 7.5f;
  const float uf[2] = {(float)(lane & 7) - 7.5f, (float)(lane & 7) + 0.5f};
  const float vf[2] = {(float)(lane >> 3) - 7.5f, (float)(lane >> 3) + 0.5f};
  float T[4] = {1.f, 1.f, 1.f, 1.f};
  float Cr[4] = {0.f, 0.f, 0.f, 0.f}, Cg[4] = {0.f, 0.f, 0.f, 0.f}, Cb[4] = {0.f, 0.f, 0.f, 0.f};
  const int n = end - start;
  if (n > 0) {
    // Stage batch 0, prefetch batch 1.  Slots past the tile's list hold a
    // neutral Gaussian (power -inf: alpha 0, colour and T unchanged), so the
    // shading loop runs whole groups of CHECK with no per-Gaussian bound.
    Feat pre = neutral_feat();
    if (lane < n) {
      const int idx = GATHERED ? start + lane : (int)order[start + lane];
      pre = load_feat<PREC>(idx, xy, rgbd, con, rec, tk);
    }
    int cur = 0;
    {
      const Staged st = stage_feat(pre, cx, cy);
      sA[0][lane] = st.A;
      sB[0][lane] = st.B;
      sC[0][lane] = st.bl;
    }
    uint64_t live[4];
    {
      const uint32_t m = quad_mask(pre, cx, cy);
#pragma unroll
      for (int q = 0; q < 4; ++q) live[q] = ballot((m >> q) & 1u);
    }
    pre = neutral_feat();
    if (kBatch + lane < n) {
      const int idx = GATHERED ? start + kBatch + lane : (int)order[start + kBatch + lane];
      pre = load_feat<PREC>(idx, xy, rgbd, con, rec, tk);
    }
    __syncthreads();
    for (int b0 = 0;; b0 += kBatch) {
      const int cnt = (n - b0 < kBatch) ? (n - b0) : kBatch;
      const int groups = (cnt + CHECK - 1) / CHECK;
      bool done = false;
      for (int g = 0; g < groups && !done; ++g) {
        const uint64_t gm = (CHECK >= 64 ? ~0ull : ((1ull << CHECK) - 1ull)) << (g * CHECK);
        const uint64_t l0 = live[0] & gm, l1 = live[1] & gm, l2 = live[2] & gm,
                       l3 = live[3] & gm;
        const int nlive = popc64(l0) + popc64(l1) + popc64(l2) + popc64(l3);
        if (kSparseCost * nlive + kSparseGroupCost * CHECK < kDenseCost * CHECK) {
          // Sparse: per quadrant, its live Gaussians, one pixel per lane.
          const uint64_t lq[4] = {l0, l1, l2, l3};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const float u = uf[q & 1], v = vf[q >> 1];
            uint64_t bits = lq[q];
            while (bits) {
              const int k = __builtin_ctzll(bits);
              bits &= bits - 1;
              const float4 A = sA[cur][k];
              const float4 B = sB[cur][k];
              const float bl = sC[cur][k];
              const float pw = staged_power(A, B, u, v);
              if (DT && pw < -7.994353375f) continue;
              const float alpha = __builtin_amdgcn_exp2f(pw);
              const float wgt = alpha * T[q];
              Cr[q] = fmaf(B.z, wgt, Cr[q]);
              Cg[q] = fmaf(B.w, wgt, Cg[q]);
              Cb[q] = fmaf(bl, wgt, Cb[q]);
              T[q] = T[q] - wgt;
            }
          }
        } else {
          // Dense: every Gaussian of the group on all four pixels (not
          // unrolled: the persistent kernel fits 72 VGPRs, 7 waves per SIMD).
#pragma unroll 1
          for (int j = 0; j < CHECK; ++j) {
            const int k = g * CHECK + j;
            const float4 A = sA[cur][k];
            const float4 B = sB[cur][k];
            const float bl = sC[cur][k];
            // staged_power for the lane's four pixels, with the per-row
            // (v) and per-column (u) terms computed once.
            const float rv0 = fmaf(fmaf(B.y, vf[0], A.z), vf[0], A.x);
            const float rv1 = fmaf(fmaf(B.y, vf[1], A.z), vf[1], A.x);
            const float du0 = fmaf(A.w, uf[0], A.y), du1 = fmaf(A.w, uf[1], A.y);
            float pw[4];
            pw[0] = fmaf(uf[0], fmaf(B.x, vf[0], du0), rv0);
            pw[1] = fmaf(uf[1], fmaf(B.x, vf[0], du1), rv0);
            pw[2] = fmaf(uf[0], fmaf(B.x, vf[1], du0), rv1);
            pw[3] = fmaf(uf[1], fmaf(B.x, vf[1], du1), rv1);
#pragma unroll
            for (int p = 0; p < 4; ++p) {
              if (DT && pw[p] < -7.994353375f) continue;
              const float alpha = __builtin_amdgcn_exp2f(pw[p]);
              const float wgt = alpha * T[p];
              Cr[p] = fmaf(B.z, wgt, Cr[p]);
              Cg[p] = fmaf(B.w, wgt, Cg[p]);
              Cb[p] = fmaf(bl, wgt, Cb[p]);
              T[p] = T[p] - wgt;
            }
          }
        }
        // Saturation test after every CHECK Gaussians of the tile (the
        // padded group at the end of the list changes nothing).
        const bool sat = T[0] < kTMin && T[1] < kTMin && T[2] < kTMin && T[3] < kTMin;
        done = ballot(!sat) == 0;
      }
      if (done || b0 + kBatch >= n) break;
      // Publish the prefetched batch, start loading the one after it.
      __syncthreads();
      cur ^= 1;
      {
        const Staged st = stage_feat(pre, cx, cy);
        sA[cur][lane] = st.A;
        sB[cur][lane] = st.B;
        sC[cur][lane] = st.bl;
      }
      {
        const uint32_t m = quad_mask(pre, cx, cy);
#pragma unroll
        for (int q = 0; q < 4; ++q) live[q] = ballot((m >> q) & 1u);
      }
      pre = neutral_feat();
      const int nb = b0 + 2 * kBatch + lane;
      if (nb < n) {
        const int idx = GATHERED ? start + nb : (int)order[start + nb];
        pre = load_feat<PREC>(idx, xy, rgbd, con, rec, tk);
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int px = tx * kTile + 8 * (p & 1) + (lane & 7), py = ty * kTile + 8 * (p >> 1) + (lane >> 3);
    if (px < W && py < H) {
      uint8_t* o = out + ((size_t)py * W + px) * 3;
      if (n > 0) {
        const float t = T[p] < kTMin ? 0.0f : T[p];
        o[0] = encode(Cr[p] + t * bg.x);
        o[1] = encode(Cg[p] + t * bg.y);
        o[2] = encode(Cb[p] + t * bg.z);
      } else {
        o[0] = encode(bg.x);
        o[1] = encode(bg.y);
        o[2] = encode(bg.z);
      }
    }
  }
}

// One 8x8 quadrant q of a tile (q = qy * 2 + qx), one pixel per lane: the
// sparse path of render_tile alone, with its own saturation test.  For
// frames with fewer tiles than waves in flight, so that each tile's work is
// split four ways (a frame's longest tile otherwise sets the render time).
// Per pixel the same operations in the same order as render_tile; the
// Gaussians left out are those whose alpha stays below 2^-24 in this
// quadrant, and the quadrant stops on its own pixels' transmittance, where a
// tile waits for all four (pixels past saturation add < 1e-4 of colour).
template <int PREC, bool GATHERED, int CHECK, bool DT>
__device__ __forceinline__ void render_quad(
    int tile, int q, int start, int end, int W, int H, int gx, const uint32_t* __restrict__ order,
    const float2* __restrict__ xy, const float4* __restrict__ rgbd, const float4* __restrict__ con,
    const FeatRec16* __restrict__ rec, const TopKCache& tk, float3 bg,
    uint8_t* __restrict__ out, float4 (*sA)[kBatch], float4 (*sB)[kBatch], float (*sC)[kBatch],
    int lane) {
  const int tx = tile % gx, ty = tile / gx;
  const float cx = (float)(tx * kTile) + 7.5f, cy = (float)(ty * kTile) + 7.5f;
  const float u = (float)(lane & 7) + 8.0f * (float)(q & 1) - 7.5f;
  const float v = (float)(lane >> 3) + 8.0f * (float)(q >> 1) - 7.5f;
  float T = 1.f, Cr = 0.f, Cg = 0.f, Cb = 0.f;
  const int n = end - start;
  if (n > 0) {
    Feat pre = neutral_feat();
    if (lane < n) {
      const int idx = GATHERED ? start + lane : (int)order[start + lane];
      pre = load_feat<PREC>(idx, xy, rgbd, con, rec, tk);
    }
    int cur = 0;
    {
      const Staged st = stage_feat(pre, cx, cy);
      sA[0][lane] = st.A;
      sB[0][lane] = st.B;
      sC[0][lane] = st.bl;
    }
    uint64_t live = ballot((quad_mask(pre, cx, cy) >> q) & 1u);
    pre = neutral_feat();
    if (kBatch + lane < n) {
      const int idx = GATHERED ? start + kBatch + lane : (int)order[start + kBatch + lane];
      pre = load_feat<PREC>(idx, xy, rgbd, con, rec, tk);
    }
    __syncthreads();
    for (int b0 = 0;; b0 += kBatch) {
      const int cnt = (n - b0 < kBatch) ? (n - b0) : kBatch;
      const int groups = (cnt + CHECK - 1) / CHECK;
      bool done = false;
      for (int g = 0; g < groups && !done; ++g) {
        const uint64_t gm = (CHECK >= 64 ? ~0ull : ((1ull << CHECK) - 1ull)) << (g * CHECK);
        uint64_t bits = live & gm;
        while (bits) {
          const int k = __builtin_ctzll(bits);
          bits &= bits - 1;
          const float4 A = sA[cur][k];
          const float4 B = sB[cur][k];
          const float bl = sC[cur][k];
          const float pw = staged_power(A, B, u, v);
          if (DT && pw < -7.994353375f) continue;
          const float alpha = __builtin_amdgcn_exp2f(pw);
          const float wgt = alpha * T;
          Cr = fmaf(B.z, wgt, Cr);
          Cg = fmaf(B.w, wgt, Cg);
          Cb = fmaf(bl, wgt, Cb);
          T = T - wgt;
        }
        done = ballot(!(T < kTMin)) == 0;
      }
      if (done || b0 + kBatch >= n) break;
      __syncthreads();
      cur ^= 1;
      {
        const Staged st = stage_feat(pre, cx, cy);
        sA[cur][lane] = st.A;
        sB[cur][lane] = st.B;
        sC[cur][lane] = st.bl;
      }
      live = ballot((quad_mask(pre, cx, cy) >> q) & 1u);
      pre = neutral_feat();
      const int nb = b0 + 2 * kBatch + lane;
      if (nb < n) {
        const int idx = GATHERED ? start + nb : (int)order[start + nb];
        pre = load_feat<PREC>(idx, xy, rgbd, con, rec, tk);
      }
      __syncthreads();
    }
  }
  const int px = tx * kTile + 8 * (q & 1) + (lane & 7), py = ty * kTile + 8 * (q >> 1) + (lane >> 3);
  if (px < W && py < H) {
    uint8_t* o = out + ((size_t)py * W + px) * 3;
    if (n > 0) {
      const float t = T < kTMin ? 0.0f : T;
      o[0] = encode(Cr + t * bg.x);
      o[1] = encode(Cg + t * bg.y);
      o[2] = encode(Cb + t * bg.z);
    } else {
      o[0] = encode(bg.x);
      o[1] = encode(bg.y);
      o[2] = encode(bg.z);
    }
  }
}

// One workgroup per tile.
template <int PREC, int CHECK>
__global__ __launch_bounds__(kWave, 6) void ps_render_plain(
    const uint64_t* __restrict__ keys, const int* __restrict__ d_n, int n_host, int W, int H,
    int gx, const uint32_t* __restrict__ order, const float2* __restrict__ xy,
    const float4* __restrict__ rgbd, const float4* __restrict__ con,
    const FeatRec16* __restrict__ rec, float3 bg, uint8_t* __restrict__ out,
    const int2* __restrict__ ranges, const uint32_t* __restrict__ tile_order,
    const int4* __restrict__ tile_desc, const FrameConsts* __restrict__ fc, TopKCache tk) {
  __shared__ float4 sA[2][kBatch];
  __shared__ float4 sB[2][kBatch];
  __shared__ float sC[2][kBatch];
  const int lane = threadIdx.x;
  if (fc) {  // a replayed graph's frame: background and image from device memory
    bg = make_float3(fc->bg[0], fc->bg[1], fc->bg[2]);
    out = fc->out;
  }
  const int pos = blockIdx.y * gx + blockIdx.x;
  const int tiles = gx * (int)gridDim.y;
  // With a tile order, workgroup `pos` renders tile tile_order[pos]
  // (render.cu:4642-4644), or the tile of packed descriptor pos (column,
  // row, range start, range end; render.cu:4959-4962); an entry outside
  // the frame is skipped.
  int tile, start, end;
  if (tile_desc) {
    const int4 d = tile_desc[pos];
    tile = d.y * gx + d.x;
    if ((unsigned)d.x >= (unsigned)gx || (unsigned)tile >= (unsigned)tiles) return;
    start = d.z;
    end = d.w;
  } else {
  tile = tile_order ? (int)tile_order[pos] : pos;
  if (tile_order && (unsigned)tile >= (unsigned)tiles) return;
  if (ranges) {
    const int2 r = ranges[tile];
    start = r.x;
    end = r.y;
  } else {
    const int n = d_n ? *d_n : n_host;
    start = wave_bound(keys, n, (uint32_t)tile, false, lane);
    end = wave_bound(keys, n, (uint32_t)tile, true, lane);
  }
  }
  render_tile<PREC, false, CHECK, false>(tile, start, end, W, H, gx, order, xy, rgbd, con, rec, tk,
                                         bg, out, sA, sB, sC, lane);
}

// Tile queue of the persistent renderer: one claim counter per XCD (128 bytes
// apart), each over every eighth tile row of the frame (queue x: rows x,
// x + 8, x + 16, ..., row-major within a row), plus a count of finished
// workgroups.  Workgroup w starts on queue w % 8 (the XCD the dispatcher
// places it on, so a queue's neighbouring tiles, which share records, are
// rendered at the same time under that XCD's L2) and, once its queue is
// drained, moves on to the others in order.  Dealing rows rather than
// contiguous eighths of the frame gives every XCD a sample of the whole
// frame's work (measured 1-1.6% faster renders at 6M Gaussians 3840x2160,
// the 56M-Gaussian city and 30M Gaussians at 3840x2160 and 7680x4320; strips
// of two or four rows were no better).  A
// frame of many light tiles then spreads its atomics over eight counters
// instead of one (at 7680x4320 with 1M Gaussians the single counter held the
// render at 1.56 ms; the eight take 0.36 ms).  Claims are single tiles:
// chunks of contiguous tiles per claim were measured slower on heavy frames
// (56M-Gaussian city at 9216x6912: 6.0 vs 3.2 ms).  The last workgroup to
// finish zeroes every word, so the next launch needs no reset.
constexpr int kQueueLines = 8;
constexpr int kQuadPairs = 512;  // mean pairs per tile above which quadrants pay
constexpr int kQueueStride = 32;                         // words between counters
constexpr int kQueueDone = kQueueLines * kQueueStride;  // finished workgroups

struct TileQueue {
  unsigned* q;
  int items, lane, home, probe = 0;
  int group;  // items per tile (4 with quadrant items), kept on one queue
  bool inter;
  int strip;  // items per tile row (0: contiguous eighths)
  __device__ TileQueue(unsigned* queue, int nitems, int l, int items_per_tile, bool interleave,
                       int row_items)
      : q(queue), items(nitems), lane(l), home((int)(blockIdx.x % kQueueLines)),
        group(items_per_tile), inter(interleave), strip(row_items) {}
  __device__ int lo(int x) const { return (int)(((long long)x * items) / kQueueLines); }
  // Items of queue x: when the items follow a tile order (interleave), the
  // items of every eighth tile of the order from tile x, so that each XCD
  // starts on its share of the order's head and a tile's quadrants stay on
  // one XCD; else every eighth row of items from row x (the last row may be
  // short), or with no row length a contiguous eighth [lo(x), lo(x + 1)).
  __device__ int len(int x) const {
    if (inter) {
      const int units = items / group;
      return x < units ? group * ((units - 1 - x) / kQueueLines + 1) : 0;
    }
    if (strip <= 0) return lo(x + 1) - lo(x);
    const int full = items / strip, rem = items - full * strip;
    const int rows = full > x ? (full - 1 - x) / kQueueLines + 1 : 0;
    return rows * strip + ((rem > 0 && full % kQueueLines == x) ? rem : 0);
  }
  __device__ int item(int x, int c) const {
    if (inter) return (x + kQueueLines * (c / group)) * group + c % group;
    if (strip <= 0) return lo(x) + c;
    return (x + kQueueLines * (c / strip)) * strip + c % strip;
  }
  // The next item, or -1 when every queue is drained.  The claim is
  // compared unsigned, so a counter that does not start at zero can only
  // end the queue early, never yield an item outside it.
  __device__ int claim() {
    while (probe < kQueueLines) {
      const int x = (home + probe) % kQueueLines;
      const int n = len(x);
      unsigned c = 0;
      if (lane == 0 && n > 0) c = atomicAdd(&q[x * kQueueStride], 1u);
      c = (unsigned)bcast((int)c, 0);
      if (n > 0 && c < (unsigned)n) return item(x, (int)c);
      ++probe;  // drained for good: counters only grow
    }
    return -1;
  }
  __device__ void finish() {
    if (lane == 0) {
      const unsigned done = atomicAdd(&q[kQueueDone], 1u);
      if (done == gridDim.x - 1) {
        for (int x = 0; x < kQueueLines; ++x) atomicExch(&q[x * kQueueStride], 0u);
        atomicExch(&q[kQueueDone], 0u);
      }
    }
  }
};

// Persistent: tiles from the queue.  RANGES: read tile_ranges, else search.
// quads = 1: the queue's items are (tile, quadrant) pairs, item 4 t + q,
// each rendered by render_quad; 2: so when the frame averages at least
// kQuadPairs pairs per tile, else whole tiles; 0: whole tiles.
// Seven waves per SIMD: measured 2-4% faster than six (6M Gaussians at
// 3840x2160: render 0.533 vs 0.545 ms; 30M at 7680x4320: 3.59 vs 3.72 ms)
// and level on the city scene; the few values spilled are reloaded once per
// tile.
template <int PREC, bool RANGES, bool GATHERED, int CHECK, bool DT>
__global__ __launch_bounds__(kWave, 7) void ps_render_persistent(
    const uint64_t* __restrict__ keys, const int* __restrict__ d_n, int n_host,
    const int2* __restrict__ ranges, int W, int H, int gx, int tiles, int quads,
    const uint32_t* __restrict__ order, const float2* __restrict__ xy,
    const float4* __restrict__ rgbd, const float4* __restrict__ con,
    const FeatRec16* __restrict__ rec, float3 bg, uint8_t* __restrict__ out,
    unsigned* __restrict__ queue, const uint32_t* __restrict__ tile_order,
    const int4* __restrict__ tile_desc, int interleave, const FrameConsts* __restrict__ fc,
    TopKCache tk) {
  __shared__ float4 sA[2][kBatch];
  __shared__ float4 sB[2][kBatch];
  __shared__ float sC[2][kBatch];
  const int lane = threadIdx.x;
  if (fc) {  // a replayed graph's frame: background and image from device memory
    bg = make_float3(fc->bg[0], fc->bg[1], fc->bg[2]);
    out = fc->out;
  }
  const int pairs = d_n ? *d_n : n_host;
  const int n = RANGES ? 0 : pairs;
  if (quads == 2) quads = (long long)pairs >= (long long)kQuadPairs * tiles ? 1 : 0;
  const bool ordered = (tile_order != nullptr || tile_desc != nullptr) && interleave;
  TileQueue tq(queue, quads ? 4 * tiles : tiles, lane, quads ? 4 : 1, ordered,
               gx * (quads ? 4 : 1));
  for (int item = tq.claim(); item >= 0; item = tq.claim()) {
    const int pos = quads ? item >> 2 : item;
    int tile, start, end;
    if (tile_desc) {
      // Packed descriptor (column, row, range start, range end).
      const int4 d = tile_desc[pos];
      tile = d.y * gx + d.x;
      start = d.z;
      end = d.w;
      if ((unsigned)d.x >= (unsigned)gx || (unsigned)tile >= (unsigned)tiles) continue;
    } else {
      // A tile order is a permutation of the tiles; an entry outside the
      // frame (a caller's bad order) is skipped rather than rendered.
      tile = tile_order ? (int)tile_order[pos] : pos;
      if (tile_order && (unsigned)tile >= (unsigned)tiles) continue;
      if (RANGES) {
        const int2 r = ranges[tile];
        start = r.x;
        end = r.y;
      } else {
        start = wave_bound(keys, n, (uint32_t)tile, false, lane);
        end = wave_bound(keys, n, (uint32_t)tile, true, lane);
      }
    }
    if (quads)
      render_quad<PREC, GATHERED, CHECK, DT>(tile, item & 3, start, end, W, H, gx, order, xy,
                                             rgbd, con, rec, tk, bg, out, sA, sB, sC, lane);
    else
      render_tile<PREC, GATHERED, CHECK, DT>(tile, start, end, W, H, gx, order, xy, rgbd, con,
                                             rec, tk, bg, out, sA, sB, sC, lane);
  }
  tq.finish();
}

// Persistent renderer over the tiles of several frames (one queue for all):
// global tile t belongs to the last frame whose tile0 <= t.  Frame tile
// offsets are staged in LDS once per workgroup.
constexpr int kMaxBatch = PS_MAX_BATCH_FRAMES;

template <int PREC, int CHECK>
__global__ __launch_bounds__(kWave, 6) void ps_render_persistent_batch(
    const FrameDev* __restrict__ frames, int nf, const int2* __restrict__ ranges, int tiles,
    const uint32_t* __restrict__ order, unsigned* __restrict__ queue) {
  __shared__ float4 sA[2][kBatch];
  __shared__ float4 sB[2][kBatch];
  __shared__ float sC[2][kBatch];
  __shared__ int s_tile0[kMaxBatch];
  const int lane = threadIdx.x;
  if (lane < nf) s_tile0[lane] = frames[lane].tile0;
  __syncthreads();
  int tile = blockIdx.x;
  while (tile < tiles) {
    int lo = 0, hi = nf - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_tile0[mid] <= tile) lo = mid; else hi = mid - 1;
    }
    const FrameDev& F = frames[lo];
    const int2 r = ranges[tile];
    render_tile<PREC, false, CHECK, false>(tile - F.tile0, r.x, r.y, F.cp.W, F.cp.H, F.cp.grid_x,
                                           order, F.xy, F.rgbd, F.conic, F.rec, TopKCache{},
                                           make_float3(F.bg[0], F.bg[1], F.bg[2]), F.out, sA, sB,
                                           sC, lane);
    unsigned next = 0;
    if (lane == 0) next = atomicAdd(&queue[0], 1u) + gridDim.x;
    next = (unsigned)bcast((int)next, 0);
    tile = next < (unsigned)tiles ? (int)next : tiles;  // unsigned: see TileQueue::claim
  }
  if (lane == 0) {
    const unsigned done = atomicAdd(&queue[1], 1u);
    if (done == gridDim.x - 1) {
      atomicExch(&queue[0], 0u);
      atomicExch(&queue[1], 0u);
    }
  }
}

struct GridCache {
  int device = -1;
  int cus = 0;
};

int device_cus(hipStream_t) {
  static thread_local GridCache c;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (c.device != dev) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    c.device = dev;
    c.cus = cus;
  }
  return c.cus;
}

template <typename K>
int persistent_grid(K kernel, int tiles) {
  int cus = device_cus(nullptr);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kWave, 0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 8;
  long long g = (long long)(cus > 0 ? cus : 256) * per_cu;
  if (g > tiles) g = tiles;
  return (int)(g < 1 ? 1 : g);
}

// Work items of the persistent renderer (see ps_render_persistent): quadrants
// when the frame has fewer tiles than waves in flight (every tile then has a
// wave to itself and the longest tile sets the render time), the kernel's
// choice from the pair count when it has fewer than twice as many, whole
// tiles otherwise.  Measured (render ms, whole tiles -> quadrants): 1280x720
// with 3M / 30M Gaussians 0.52 -> 0.32 / 1.27 -> 0.78; 1920x1080 with 10M
// (1532 pairs per tile) 1.14 -> 0.91 but with 1M (153) 0.15 -> 0.20;
// 3840x2160 with 6M 0.59 -> 0.67.  POLYSPLAT_RENDER_QUADS=0 or 1 forces it.
int quadrant_items(int tiles, int waves) {
  const char* e = getenv("POLYSPLAT_RENDER_QUADS");
  if (e && (e[0] == '0' || e[0] == '1')) return e[0] == '1';
  if (tiles < waves) return 1;
  return tiles < 2 * waves ? 2 : 0;
}

// How the tile queue deals out the positions of a tile order: interleaved
// (position p to XCD queue p % 8, so every XCD starts on the heaviest
// tiles) or in contiguous eighths (each XCD a run of the order, which keeps
// a spatial order's neighbourhoods on one XCD's L2).
// POLYSPLAT_ORDER_QUEUES=contiguous selects the second.
int order_queues_interleaved() {
  const char* e = getenv("POLYSPLAT_ORDER_QUEUES");
  return (e && e[0] == 'c') ? 0 : 1;
}

// Arguments of one persistent render launch.
struct PersistArgs {
  const uint64_t* keys;
  const int* d_n;
  int n_host;
  const int* ranges;
  int W, H;
  const uint32_t* order;
  const float* xy;
  const float* rgbd;
  const float* con;
  const void* rec;
  float3 bg;
  uint8_t* out;
  unsigned* queue;
  const uint32_t* tile_order;  // NULL: tiles in row-major order
  const int* tile_desc;        // NULL, or int4 (column, row, start, end) per position
  const FrameConsts* fc;       // NULL, or background and image in device memory
  TopKCache tk;                // top-k feature cache (FP32 features only)
};

template <int PREC, bool RANGES, bool GATHERED, int CHECK, bool DT>
int launch_persistent(const PersistArgs& a, hipStream_t stream) {
  const int gx = (a.W + kTile - 1) / kTile, gy = (a.H + kTile - 1) / kTile;
  const int tiles = gx * gy;
  auto k = ps_render_persistent<PREC, RANGES, GATHERED, CHECK, DT>;
  const int waves = persistent_grid(k, 1 << 30);
  const int quads = quadrant_items(tiles, waves);
  const int items = quads == 1 ? 4 * tiles : tiles;  // quads == 2: tiles >= waves
  const int grid = waves < items ? waves : items;
  hipLaunchKernelGGL(k, dim3(grid), dim3(kWave), 0, stream, a.keys, a.d_n, a.n_host,
                     reinterpret_cast<const int2*>(a.ranges), a.W, a.H, gx, tiles, quads, a.order,
                     (const float2*)a.xy, (const float4*)a.rgbd, (const float4*)a.con,
                     (const FeatRec16*)a.rec, a.bg, a.out, a.queue, a.tile_order,
                     reinterpret_cast<const int4*>(a.tile_desc), order_queues_interleaved(), a.fc,
                     a.tk);
  PS_LAUNCH_CHECK();
  return PS_OK;
}

template <int PREC, bool RANGES, bool GATHERED>
int dispatch_dt(bool dt, const PersistArgs& a, hipStream_t s) {
  // Cadence of the saturation test: every 8 Gaussians with precomputed
  // ranges (render.cu:1814), every 16 with the in-kernel search (render.cu:2124).
  constexpr int kCheck = RANGES ? 8 : 16;
  if (dt) return launch_persistent<PREC, RANGES, GATHERED, kCheck, true>(a, s);
  return launch_persistent<PREC, RANGES, GATHERED, kCheck, false>(a, s);
}

template <int PREC>
int dispatch_persistent(bool gathered, bool dt, const PersistArgs& a, hipStream_t s) {
  if (a.ranges != nullptr || a.tile_desc != nullptr) {
    if (gathered) return dispatch_dt<PREC, true, true>(dt, a, s);
    return dispatch_dt<PREC, true, false>(dt, a, s);
  }
  if (gathered) return dispatch_dt<PREC, false, true>(dt, a, s);
  return dispatch_dt<PREC, false, false>(dt, a, s);
}

}  // namespace

int render_batch_launch(const FrameDev* frames_dev, int nf, int feat_prec, const int* ranges,
                        int tiles, const uint32_t* order, unsigned* queue, hipStream_t stream) {
  PS_REQUIRE(nf > 0 && nf <= kMaxBatch, "batch of %d frames (max %d)", nf, kMaxBatch);
  if (tiles <= 0) return PS_OK;
  auto go = [&](auto kernel) {
    const int grid = persistent_grid(kernel, tiles);
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kWave), 0, stream, frames_dev, nf,
                       reinterpret_cast<const int2*>(ranges), tiles, order, queue);
  };
  switch (feat_prec) {
    case PS_PREC_F32: go(ps_render_persistent_batch<kFeatF32, 8>); break;
    case PS_PREC_F16: go(ps_render_persistent_batch<kFeatF16, 8>); break;
    case PS_PREC_BF16: go(ps_render_persistent_batch<kFeatBF16, 8>); break;
    default:
      set_last_error("unknown feature precision %d", feat_prec);
      return PS_ERR_INVALID;
  }
  PS_LAUNCH_CHECK();
  return PS_OK;
}

int render_plain_launch(int n_host, const int* d_n, int W, int H, const float* xy,
                        const float* rgbd, const float* con, const uint64_t* keys,
                        const uint32_t* order, const float bg[3], uint8_t* out,
                        hipStream_t stream, const int* ranges, const uint32_t* tile_order,
                        const int* tile_desc, const FrameConsts* fc, const TopKCache& tk) {
  PS_REQUIRE(W > 0 && H > 0, "bad image size");
  PS_REQUIRE((out != nullptr || fc != nullptr) && bg != nullptr, "output or background is NULL");
  PS_REQUIRE(d_n != nullptr || n_host >= 0, "negative num_rendered");
  PS_REQUIRE(ranges != nullptr || keys != nullptr || tile_desc != nullptr,
             "need keys_sorted, tile_ranges or tile descriptors");
  const int gx = (W + kTile - 1) / kTile, gy = (H + kTile - 1) / kTile;
  const float3 b = make_float3(bg[0], bg[1], bg[2]);
  // Saturation test every 2 Gaussians, as the reference's plain path does
  // (render.cu:2436-2447).
  hipLaunchKernelGGL((ps_render_plain<kFeatF32, 2>), dim3(gx, gy), dim3(kWave), 0, stream, keys, d_n,
                     n_host, W, H, gx, order, (const float2*)xy, (const float4*)rgbd,
                     (const float4*)con, (const FeatRec16*)nullptr, b, out,
                     reinterpret_cast<const int2*>(ranges), tile_order,
                     reinterpret_cast<const int4*>(tile_desc), fc, tk);
  PS_LAUNCH_CHECK();
  return PS_OK;
}

// Top-k renders (render.cu:6698-6798): the values may name slots of the
// cache (bit 31); ranges from an in-kernel search of the keys; saturation
// test every 2 Gaussians as the reference's top-k kernels do
// (render.cu:2780-2847).  Plain: one workgroup per tile; persistent: the
// tile queue.
int render_topk_launch(int n_host, const int* d_n, int W, int H, const float* xy,
                       const float* rgbd, const float* con, const uint64_t* keys,
                       const uint32_t* values, const TopKCache& tk, bool persistent,
                       const float bg[3], uint8_t* out, unsigned* queue, hipStream_t stream) {
  PS_REQUIRE(W > 0 && H > 0, "bad image size");
  PS_REQUIRE(out != nullptr && bg != nullptr && keys != nullptr && values != nullptr,
             "NULL output, background, keys or values");
  PS_REQUIRE(xy && rgbd && con, "f32 feature arrays are NULL");
  PS_REQUIRE(tk.n >= 0 && (tk.n == 0 || (tk.xy && tk.rgbd && tk.con)),
             "top-k arrays are NULL for %d slots", tk.n);
  PS_REQUIRE(d_n != nullptr || n_host >= 0, "negative num_rendered");
  if (!persistent)
    return render_plain_launch(n_host, d_n, W, H, xy, rgbd, con, keys, values, bg, out, stream,
                               nullptr, nullptr, nullptr, nullptr, tk);
  PS_REQUIRE(queue != nullptr, "queue is NULL");
  PersistArgs a{keys, d_n, n_host, nullptr, W, H, values, xy, rgbd, con, nullptr,
                make_float3(bg[0], bg[1], bg[2]), out, queue, nullptr, nullptr, nullptr, tk};
  return launch_persistent<kFeatF32, false, false, 2, false>(a, stream);
}

int render_persistent_launch(int n_host, const int* d_n, int W, int H, int feat_prec,
                             const float* xy, const float* rgbd, const float* con, const void* rec,
                             int gathered, const uint64_t* keys, const uint32_t* order,
                             const int* ranges, int dt, const float bg[3], uint8_t* out,
                             unsigned* queue, hipStream_t stream, const uint32_t* tile_order,
                             const int* tile_desc, const FrameConsts* fc, const TopKCache& tk) {
  PS_REQUIRE(W > 0 && H > 0, "bad image size");
  PS_REQUIRE((out != nullptr || fc != nullptr) && bg != nullptr && queue != nullptr,
             "NULL output/bg/queue");
  PS_REQUIRE(ranges != nullptr || keys != nullptr || tile_desc != nullptr,
             "need tile_ranges, keys_sorted or tile descriptors");
  PS_REQUIRE(gathered || order != nullptr, "values_sorted is NULL");
  PS_REQUIRE(d_n != nullptr || n_host >= 0, "negative num_rendered");
  PersistArgs a{keys, d_n,  n_host, ranges, W, H, order, xy, rgbd, con, rec,
                make_float3(bg[0], bg[1], bg[2]), out, queue, tile_order, tile_desc, fc, tk};
  const bool g = gathered != 0, d = dt != 0;
  switch (feat_prec) {
    case PS_PREC_F32:
      PS_REQUIRE(xy && rgbd && con, "f32 feature arrays are NULL");
      return dispatch_persistent<kFeatF32>(g, d, a, stream);
    case PS_PREC_F16:
      PS_REQUIRE(rec != nullptr, "records are NULL");
      return dispatch_persistent<kFeatF16>(g, d, a, stream);
    case PS_PREC_BF16:
      PS_REQUIRE(rec != nullptr, "records are NULL");
      return dispatch_persistent<kFeatBF16>(g, d, a, stream);
    default:
      set_last_error("unknown feature precision %d", feat_prec);
      return PS_ERR_INVALID;
  }
}

// ------------------------------------------------ no-allocation entry point
namespace {

constexpr int kMaxDevices = 64;
static_assert(sizeof(scratch) >= 4 * (kQueueDone + 1), "queue words fit the static scratch");

// Clears the queue words before a launch (a kernel rather than a memset, so
// that the call can be recorded into a HIP graph).
__global__ __launch_bounds__(256) void ps_queue_clear(unsigned* __restrict__ q, int words) {
  for (int i = threadIdx.x; i < words; i += 256) q[i] = 0u;
}

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
int noalloc_queue(unsigned** out) {
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  int dev = 0;
  PS_HIP_CHECK(hipGetDevice(&dev));
  PS_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d the library tracks", dev,
             kMaxDevices);
  std::lock_guard<std::mutex> g(mu);
  if (addr[dev] == nullptr) {
    void* p = nullptr;
    PS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PS_REQUIRE(p != nullptr, "no device address for the static tile queue");
    addr[dev] = p;
  }
  *out = static_cast<unsigned*>(addr[dev]);
  return PS_OK;
}

}  // namespace

int render_persistent_noalloc(int n_host, const int* d_n, int W, int H, int feat_prec,
                              const float* xy, const float* rgbd, const float* con,
                              const void* rec, int gathered, const uint64_t* keys,
                              const uint32_t* order, const int* ranges, int dt, const float bg[3],
                              uint8_t* out, hipStream_t stream) {
  unsigned* q = nullptr;
  const int st = noalloc_queue(&q);
  if (st != PS_OK) return st;
  // The queue needs no more words for larger frames: every size the
  // allocating entry point accepts is accepted here.
  const int words = (int)(sizeof(scratch) / sizeof(unsigned));
  hipLaunchKernelGGL(ps_queue_clear, dim3(1), dim3(256), 0, stream, q, words);
  PS_LAUNCH_CHECK();
  return render_persistent_launch(n_host, d_n, W, H, feat_prec, xy, rgbd, con, rec, gathered,
                                  keys, order, ranges, dt, bg, out, q, stream);
}

}  // namespace ps

extern "C" int ps_render_16x16(int num_rendered, const int* d_num_rendered, int width, int height,
                               const float* points_xy, const float* rgb_depth,
                               const float* conic_opacity, const uint64_t* keys_sorted,
                               const uint32_t* values_sorted, const float bg[3],
                               uint8_t* out_color, ps_stream_t stream) {
  return ps::render_plain_launch(num_rendered, d_num_rendered, width, height, points_xy, rgb_depth,
                                 conic_opacity, keys_sorted, values_sorted, bg, out_color,
                                 reinterpret_cast<hipStream_t>(stream));
}

extern "C" int ps_render_16x16_persistent(int num_rendered, const int* d_num_rendered, int width,
                                          int height, int feat_prec, const float* points_xy,
                                          const float* rgb_depth, const float* conic_opacity,
                                          const void* records, int gathered,
                                          const uint64_t* keys_sorted,
                                          const uint32_t* values_sorted, const int* tile_ranges,
                                          int dt, const float bg[3], uint8_t* out_color,
                                          unsigned* queue, ps_stream_t stream) {
  return ps::render_persistent_launch(num_rendered, d_num_rendered, width, height, feat_prec,
                                      points_xy, rgb_depth, conic_opacity, records, gathered,
                                      keys_sorted, values_sorted, tile_ranges, dt, bg, out_color,
                                      queue, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int ps_render_16x16_persistent_noalloc(
    int num_rendered, const int* d_num_rendered, int width, int height, int feat_prec,
    const float* points_xy, const float* rgb_depth, const float* conic_opacity,
    const void* records, int gathered, const uint64_t* keys_sorted, const uint32_t* values_sorted,
    const int* tile_ranges, int dt, const float bg[3], uint8_t* out_color, ps_stream_t stream) {
  return ps::render_persistent_noalloc(num_rendered, d_num_rendered, width, height, feat_prec,
                                       points_xy, rgb_depth, conic_opacity, records, gathered,
                                       keys_sorted, values_sorted, tile_ranges, dt, bg, out_color,
                                       reinterpret_cast<hipStream_t>(stream));
}

extern "C" int ps_render_16x16_reordered(int num_rendered, const int* d_num_rendered, int width,
                                         int height, const float* points_xy,
                                         const float* rgb_depth, const float* conic_opacity,
                                         const uint64_t* keys_sorted,
                                         const uint32_t* values_sorted, const int* tile_ranges,
                                         const uint32_t* tile_order, const float bg[3],
                                         uint8_t* out_color, ps_stream_t stream) {
  PS_REQUIRE(tile_ranges != nullptr && tile_order != nullptr, "tile_ranges or tile_order is NULL");
  return ps::render_plain_launch(num_rendered, d_num_rendered, width, height, points_xy, rgb_depth,
                                 conic_opacity, keys_sorted, values_sorted, bg, out_color,
                                 reinterpret_cast<hipStream_t>(stream), tile_ranges, tile_order);
}

extern "C" int ps_render_16x16_reordered_persistent(
    int num_rendered, const int* d_num_rendered, int width, int height, int feat_prec,
    const float* points_xy, const float* rgb_depth, const float* conic_opacity,
    const void* records, const uint32_t* values_sorted, const int* tile_ranges,
    const uint32_t* tile_order, const int* tile_desc, int dt, const float bg[3],
    uint8_t* out_color, unsigned* queue, ps_stream_t stream) {
  PS_REQUIRE(tile_desc != nullptr || (tile_ranges != nullptr && tile_order != nullptr),
             "need tile_desc, or tile_ranges and tile_order");
  return ps::render_persistent_launch(num_rendered, d_num_rendered, width, height, feat_prec,
                                      points_xy, rgb_depth, conic_opacity, records, 0, nullptr,
                                      values_sorted, tile_desc ? nullptr : tile_ranges, dt, bg,
                                      out_color, queue, reinterpret_cast<hipStream_t>(stream),
                                      tile_desc ? nullptr : tile_order, tile_desc);
}

extern "C" int ps_render_16x16_reordered_desc(int num_rendered, const int* d_num_rendered,
                                              int width, int height, const float* points_xy,
                                              const float* rgb_depth, const float* conic_opacity,
                                              const uint32_t* values_sorted, const int* tile_desc,
                                              const float bg[3], uint8_t* out_color,
                                              ps_stream_t stream) {
  PS_REQUIRE(tile_desc != nullptr, "tile_desc is NULL");
  return ps::render_plain_launch(num_rendered, d_num_rendered, width, height, points_xy, rgb_depth,
                                 conic_opacity, nullptr, values_sorted, bg, out_color,
                                 reinterpret_cast<hipStream_t>(stream), nullptr, nullptr,
                                 tile_desc);
}

extern "C" int ps_render_16x16_topk(int num_rendered, const int* d_num_rendered, int width,
                                    int height, const float* points_xy, const float* rgb_depth,
                                    const float* conic_opacity, const uint64_t* keys_sorted,
                                    const uint32_t* values_sorted, const float* topk_xy,
                                    const float* topk_rgb_depth, const float* topk_conic,
                                    int num_topk, int persistent, const float bg[3],
                                    uint8_t* out_color, unsigned* queue, ps_stream_t stream) {
  ps::TopKCache tk;
  tk.xy = reinterpret_cast<const float2*>(topk_xy);
  tk.rgbd = reinterpret_cast<const float4*>(topk_rgb_depth);
  tk.con = reinterpret_cast<const float4*>(topk_conic);
  tk.n = num_topk;
  if (num_topk == 0) tk.xy = nullptr;  // nothing cached: values are plain ids
  return ps::render_topk_launch(num_rendered, d_num_rendered, width, height, points_xy, rgb_depth,
                                conic_opacity, keys_sorted, values_sorted, tk, persistent != 0,
                                bg, out_color, queue, reinterpret_cast<hipStream_t>(stream));
}
