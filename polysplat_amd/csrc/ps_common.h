// Shared device-side definitions for the polysplat MI355X rasterizer.
//
// Everything that decides a tile key (projection, EWA covariance, conic,
// opacity power, bounding rectangle, ellipse/tile intersection) lives here
// and is compiled with FP contraction off, correctly rounded sqrt and
// division, and a self-written log2, so that the same FP32 decisions can be
// restated bit for bit on the host.  The equations follow the reference's
// forward pass (csrc/cuda_rasterizer/preprocess.cu:78-311, :354-414).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#define PS_HD __host__ __device__ __forceinline__

namespace ps {

// ---------------------------------------------------------------------------
// Constants (preprocess.cu:13-38).  SH constants are the textbook real
// spherical-harmonic normalisations; the others are the reference's cutoffs.
// ---------------------------------------------------------------------------
constexpr float kLog2e = 1.4426950216293334961f;
constexpr float kLn2 = 0.69314718055f;
constexpr float kCutoff = 8.0f;    // opacity dynamic range kept: 2^-8
constexpr float kBMinSq = 0.09f;   // thin-Gaussian cull: min semi-axis^2
constexpr float kOThin = 0.5f;     // thin cull applies below this opacity
constexpr float kNearCull = 0.2f;  // view-space z at or below which a point is dropped

constexpr float kShC0 = 0.28209479177387814f;
constexpr float kShC1 = 0.4886025119029199f;
constexpr float kShC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                            -1.0925484305920792f, 0.5462742152960396f};
constexpr float kShC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                            0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                            -0.5900435899266435f};

// Feature record precisions for the render path.
enum FeatPrec : int { kFeatF32 = 0, kFeatF16 = 1, kFeatBF16 = 2 };
// SH storage precisions.
enum ShPrec : int { kShF32 = 0, kShF16 = 1, kShBF16 = 2 };

// Camera state handed to the kernels by value.  Matrices are column-major
// 4x4 (element [c*4+r]) exactly as the host builds them.
struct CamParams {
  float view[16];
  float proj[16];
  float campos[3];
  float tan_fovx, tan_fovy;
  float focal_x, focal_y;
  int W, H;
  int block_x, block_y;
  int grid_x, grid_y;
};

// ---------------------------------------------------------------------------
// Bit-reproducible scalar helpers.
// ---------------------------------------------------------------------------

// float -> int with truncation toward zero; NaN -> 0 and saturation at the
// int range, so host and device agree on every input.
PS_HD int f2i_trunc(float x) {
  if (!(x == x)) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return (-2147483647 - 1);
  return (int)x;
}

PS_HD float u2f(uint32_t u) {
  union { uint32_t u; float f; } c;
  c.u = u;
  return c.f;
}
PS_HD uint32_t f2u(float f) {
  union { uint32_t u; float f; } c;
  c.f = f;
  return c.u;
}

// log2 for the opacity power term.  x = 2^e * m with m in [sqrt(1/2), sqrt(2)),
// ln(m) = 2*atanh(s), s = (m-1)/(m+1), summed to s^9 (|s| <= 0.1716, so the
// truncation error is below 4e-10).  Only IEEE +,-,*,/ are used, so the result
// is identical on every IEEE machine with contraction disabled.
PS_HD float log2_det(float x) {
  if (!(x > 0.0f)) return (x == 0.0f) ? -__builtin_inff() : __builtin_nanf("");
  if (x == __builtin_inff()) return x;
  int e_adj = 0;
  if (x < 1.17549435e-38f) {  // subnormal: rescale into the normal range
    x = x * 8388608.0f;
    e_adj = -23;
  }
  uint32_t bits = f2u(x);
  int e = (int)((bits >> 23) & 0xffu) - 127 + e_adj;
  float m = u2f((bits & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e = e + 1;
  }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float p = s2 * (0.333333343f + s2 * (0.2f + s2 * (0.142857149f + s2 * 0.111111112f)));
  float ln_m = (s + s) + (s + s) * p;
  return (float)e + ln_m * 1.44269502f;
}

// ---------------------------------------------------------------------------
// Projection (preprocess.cu:78-114).
// ---------------------------------------------------------------------------
PS_HD float ndc2pix(float v, int S) {
  // The reference evaluates this with double literals; keep double here.
  const double r = (((double)v + 1.0) * (double)S - 1.0) * 0.5;
  return (float)r;
}

struct F3 { float x, y, z; };

PS_HD F3 xform4x3(float px, float py, float pz, const float* m) {
  F3 t;
  t.x = m[0] * px + m[4] * py + m[8] * pz + m[12];
  t.y = m[1] * px + m[5] * py + m[9] * pz + m[13];
  t.z = m[2] * px + m[6] * py + m[10] * pz + m[14];
  return t;
}

PS_HD float xform4x4_w(float px, float py, float pz, const float* m) {
  return m[3] * px + m[7] * py + m[11] * pz + m[15];
}

PS_HD float fminf_(float a, float b) { return a < b ? a : b; }
PS_HD float fmaxf_(float a, float b) { return a > b ? a : b; }
PS_HD int imin(int a, int b) { return a < b ? a : b; }
PS_HD int imax(int a, int b) { return a > b ? a : b; }

// EWA screen-space covariance (preprocess.cu:117-157, Zwicker et al. 2002
// eqs. 29/31) with the 3x3 products expanded in the column-major order the
// reference's matrix library uses, then the +0.3 low-pass on the diagonal.
// Returns (c00, c01, c11).
PS_HD F3 cov2d(float px, float py, float pz, float fx, float fy, float tfx, float tfy,
               const float* s /*6*/, const float* v /*view 16*/) {
  F3 t = xform4x3(px, py, pz, v);
  const float limx = 1.3f * tfx;
  const float limy = 1.3f * tfy;
  const float txtz = t.x / t.z;
  const float tytz = t.y / t.z;
  t.x = fminf_(limx, fmaxf_(-limx, txtz)) * t.z;
  t.y = fminf_(limy, fmaxf_(-limy, tytz)) * t.z;

  // J columns (third column is zero).
  float J[3][3] = {{fx / t.z, 0.0f, -(fx * t.x) / (t.z * t.z)},
                   {0.0f, fy / t.z, -(fy * t.y) / (t.z * t.z)},
                   {0.0f, 0.0f, 0.0f}};
  // W[c][r] = view[r*4 + c] (transposed upper-left 3x3 of the view matrix).
  float Wm[3][3] = {{v[0], v[4], v[8]}, {v[1], v[5], v[9]}, {v[2], v[6], v[10]}};
  // T = W * J : T[c][r] = W[0][r]*J[c][0] + W[1][r]*J[c][1] + W[2][r]*J[c][2]
  float T[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r)
      T[c][r] = Wm[0][r] * J[c][0] + Wm[1][r] * J[c][1] + Wm[2][r] * J[c][2];
  // Vrk is symmetric, so transpose(Vrk) == Vrk element for element.
  float Vk[3][3] = {{s[0], s[1], s[2]}, {s[1], s[3], s[4]}, {s[2], s[4], s[5]}};
  // A = transpose(T) * Vrk ; At[c][r] = T[r][c]
  float A[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r)
      A[c][r] = T[r][0] * Vk[c][0] + T[r][1] * Vk[c][1] + T[r][2] * Vk[c][2];
  // cov = A * T, only [0][0], [0][1], [1][1] are needed.
  float c00 = A[0][0] * T[0][0] + A[1][0] * T[0][1] + A[2][0] * T[0][2];
  float c01 = A[0][1] * T[0][0] + A[1][1] * T[0][1] + A[2][1] * T[0][2];
  float c11 = A[0][1] * T[1][0] + A[1][1] * T[1][1] + A[2][1] * T[1][2];
  F3 r;
  r.x = c00 + 0.3f;
  r.y = c01;
  r.z = c11 + 0.3f;
  return r;
}

// Result of projecting one Gaussian: everything the tile test and the
// feature write need.
struct Proj {
  float xy_x, xy_y;       // pixel-space centre
  float con_a, con_b, con_c;  // inverse 2D covariance (conic)
  float power;            // ln(256 * opacity): ellipse level for the tile test
  float log2_opacity;
  float depth;            // view-space z (key low word)
  int rmin_x, rmin_y, rmax_x, rmax_y;  // candidate tile rectangle [min, max)
  bool valid;             // survived all culls and the rectangle is non-empty
};

// q = (x0, y0, x1, y1) / (bx, by, bx, by) for tile sizes other than 16.  Not
// inlined, so that the 16x16 path does not evaluate these divisions
// speculatively.
struct Quot4 {
  float x0, y0, x1, y1;
};
__host__ __device__ __attribute__((noinline)) inline Quot4 divide_by_block(Quot4 q, int bx,
                                                                          int by) {
  return Quot4{q.x0 / (float)bx, q.y0 / (float)by, q.x1 / (float)bx, q.y1 / (float)by};
}

// Cull + project + bounding rectangle (preprocess.cu:354-414).
PS_HD Proj project(float px, float py, float pz, float opacity, const float* cov3,
                   const CamParams& cp) {
  Proj o;
  o.valid = false;
  o.rmin_x = o.rmin_y = o.rmax_x = o.rmax_y = 0;
  F3 pv = xform4x3(px, py, pz, cp.view);
  o.depth = pv.z;
  if (pv.z <= kNearCull) return o;
  F3 ph3 = xform4x3(px, py, pz, cp.proj);
  float phw = xform4x4_w(px, py, pz, cp.proj);
  float pw = 1.0f / (phw + 0.0000001f);
  float ndc_x = ph3.x * pw;
  float ndc_y = ph3.y * pw;
  F3 cov = cov2d(px, py, pz, cp.focal_x, cp.focal_y, cp.tan_fovx, cp.tan_fovy, cov3, cp.view);
  float det = cov.x * cov.z - cov.y * cov.y;
  float det_inv = 1.0f / det;
  o.con_a = cov.z * det_inv;
  o.con_b = -cov.y * det_inv;
  o.con_c = cov.x * det_inv;
  o.log2_opacity = log2_det(opacity);
  o.power = kLn2 * kCutoff + kLn2 * o.log2_opacity;
  float trace = cov.x + cov.z;
  if (2.0f * det * o.power < kBMinSq * trace && opacity < kOThin) return o;
  int width = f2i_trunc(1.414214f * __builtin_sqrtf(cov.x * o.power) + 1.0f);
  int height = f2i_trunc(1.414214f * __builtin_sqrtf(cov.z * o.power) + 1.0f);
  o.xy_x = ndc2pix(ndc_x, cp.W);
  o.xy_y = ndc2pix(ndc_y, cp.H);
  Quot4 q{o.xy_x - (float)width, o.xy_y - (float)height, o.xy_x + (float)width,
          o.xy_y + (float)height};
  if (cp.block_x == 16 && cp.block_y == 16) {
    // v / 16 == v * 0.0625 exactly (a power-of-two divisor): the quotient
    // and the product are the same real number, rounded the same way.
    q = Quot4{q.x0 * 0.0625f, q.y0 * 0.0625f, q.x1 * 0.0625f, q.y1 * 0.0625f};
  } else {
    q = divide_by_block(q, cp.block_x, cp.block_y);
  }
  const float qx0 = q.x0, qy0 = q.y0, qx1 = q.x1, qy1 = q.y1;
  o.rmin_x = imin(cp.grid_x, imax(0, f2i_trunc(qx0)));
  o.rmin_y = imin(cp.grid_y, imax(0, f2i_trunc(qy0)));
  o.rmax_x = imin(cp.grid_x, imax(0, f2i_trunc(qx1) + 1));
  o.rmax_y = imin(cp.grid_y, imax(0, f2i_trunc(qy1) + 1));
  o.valid = (o.rmax_x - o.rmin_x) * (o.rmax_y - o.rmin_y) > 0;
  return o;
}

// ---------------------------------------------------------------------------
// Exact ellipse / tile test (preprocess.cu:257-311).  The ellipse is
// {q(x) <= 2*power} for the quadratic form of the conic.
// ---------------------------------------------------------------------------
PS_HD bool segment_hits_ellipse(float a, float b, float c, float d, float l, float r) {
  float delta = b * b - 4.0f * a * c;
  float t1 = (l - d) * (2.0f * a) + b;
  float t2 = (r - d) * (2.0f * a) + b;
  return delta >= 0.0f && (t1 <= 0.0f || t1 * t1 <= delta) && (t2 >= 0.0f || t2 * t2 <= delta);
}

PS_HD bool tile_hits(int tx, int ty, const Proj& g, int bx, int by) {
  const int pminx = tx * bx, pminy = ty * by;
  const int pmaxx = pminx + bx - 1, pmaxy = pminy + by - 1;
  const float cx = g.xy_x, cy = g.xy_y;
  if (cx >= (float)pminx && cx <= (float)pmaxx && cy >= (float)pminy && cy <= (float)pmaxy)
    return true;
  if (!(g.power > 0.0f)) return false;
  const float w = 2.0f * g.power;
  float dx = (cx * 2.0f < (float)(pminx + pmaxx)) ? cx - (float)pminx : cx - (float)pmaxx;
  float a = g.con_c;
  float b = -2.0f * g.con_b * dx;
  float c = g.con_a * dx * dx - w;
  if (segment_hits_ellipse(a, b, c, cy, (float)pminy, (float)pmaxy)) return true;
  float dy = (cy * 2.0f < (float)(pminy + pmaxy)) ? cy - (float)pminy : cy - (float)pmaxy;
  a = g.con_a;
  b = -2.0f * g.con_b * dy;
  c = g.con_c * dy * dy - w;
  return segment_hits_ellipse(a, b, c, cx, (float)pminx, (float)pmaxx);
}

// 64-bit sort key: tile id in the high word, depth bits in the low word
// (preprocess.cu:428-433).
PS_HD uint64_t make_key(int tile, float depth) {
  return ((uint64_t)(uint32_t)tile << 32) | (uint64_t)f2u(depth);
}

}  // namespace ps

namespace ps {

// ---------------------------------------------------------------------------
// Degree-3 SH colour (preprocess.cu:159-255).  `sh` holds 48 floats laid out
// as 16 coefficients x RGB.  Evaluation order follows the reference so that
// the FP32 result is reproducible on the host.  Each coefficient is
// accumulated by one fused multiply-add, in the reference's term order (the
// contraction its compiler applies to `r - b * sh + ...`); with 16-bit SH the
// widening of each coefficient folds into the same instruction
// (v_fma_mix_f32), so a channel costs 16 instructions.
// ---------------------------------------------------------------------------
PS_HD F3 sh_color(float px, float py, float pz, const float* campos, const float* sh) {
  float dx = px - campos[0], dy = py - campos[1], dz = pz - campos[2];
  float l2 = dx * dx + dy * dy + dz * dz;
  float rs = 1.0f / __builtin_sqrtf(l2);
  float x = dx * rs, y = dy * rs, z = dz * rs;
  float xx = x * x, yy = y * y, zz = z * z;
  float xy = x * y, yz = y * z, xz = x * z;
  float c1y = kShC1 * y, c1z = kShC1 * z, c1x = kShC1 * x;
  float b4 = kShC2[0] * xy, b5 = kShC2[1] * yz, b6 = kShC2[2] * (2.0f * zz - xx - yy);
  float b7 = kShC2[3] * xz, b8 = kShC2[4] * (xx - yy);
  float b9 = kShC3[0] * y * (3.0f * xx - yy);
  float b10 = kShC3[1] * xy * z;
  float b11 = kShC3[2] * y * (4.0f * zz - xx - yy);
  float b12 = kShC3[3] * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
  float b13 = kShC3[4] * x * (4.0f * zz - xx - yy);
  float b14 = kShC3[5] * z * (xx - yy);
  float b15 = kShC3[6] * x * (xx - 3.0f * yy);
  // Basis weights in term order; the degree-1 terms the reference subtracts
  // enter negated (r - b * s == fma(-b, s, r) exactly).
  const float basis[16] = {kShC0, -c1y, c1z, -c1x, b4,  b5,  b6,  b7,
                           b8,    b9,   b10, b11,  b12, b13, b14, b15};
  float out[3];
  for (int c = 0; c < 3; ++c) {
    float r = fmaf(basis[0], sh[c], 0.5f);
#pragma unroll
    for (int k = 1; k < 16; ++k) r = fmaf(basis[k], sh[3 * k + c], r);
    out[c] = fmaxf_(r, 0.0f);
  }
  F3 o;
  o.x = out[0];
  o.y = out[1];
  o.z = out[2];
  return o;
}

// bf16 <-> f32 (round to nearest even on the way down).
PS_HD float bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }
PS_HD uint16_t f32_to_bf16(float f) {
  uint32_t u = f2u(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Packed render record for reduced-precision features (24 bytes):
//   x, y, log2(opacity) in FP32 (projection and opacity stay exact),
//   conic scaled by 2^12 and RGB in FP16 or BF16.
// The conic scale keeps the smallest inverse variances of very wide
// Gaussians out of the FP16 subnormal range; the renderer multiplies pixel
// offsets by 2^-6 so the quadratic form is unchanged.
constexpr float kConicScale = 4096.0f;
constexpr float kDxScale = 0.015625f;
struct alignas(8) FeatRec16 {
  float x, y, w;
  uint16_t ca, cb, cc, r, g, b;
};
static_assert(sizeof(FeatRec16) == 24, "record layout");

}  // namespace ps

namespace ps {

// Depth-sort plan of the depth-first pipeline: the count pass ORs and ANDs
// the depth words of emitting Gaussians into kPlanSlots slot pairs (slot
// words kPlanStride apart, so that the atomics of different slots fall on
// different cache lines); the sort kernels combine the slots.
constexpr int kPlanSlots = 64;
constexpr int kPlanStride = 32;

// Buffers of the depth sort (sort.hip) and of the emission that reads its
// order (preprocess.hip).
struct DepthSortArgs {
  const uint32_t* plan;   // kPlanSlots OR words, then kPlanSlots AND words (kPlanStride apart)
  const uint32_t* words;  // depth word per Gaussian, 0 when it emits nothing
  uint32_t* kb[2];        // keys between digits
  uint32_t* vb[2];        // values (Gaussian ids) between digits
  int n;                  // Gaussians
};

// Plan of the depth sort: varying = OR ^ AND of the emitting depth words (0
// when nothing emits); digit q (8 bits) runs when its byte of varying is
// non-zero.  The sorted ids end in vb[(running digits - 1) % 2], or are the
// identity when no digit runs.
__host__ __device__ inline int depth_digits_running(uint32_t varying) {
  int k = 0;
  for (int q = 0; q < 4; ++q) k += ((varying >> (8 * q)) & 0xffu) != 0u ? 1 : 0;
  return k;
}

// Everything the preprocess and render kernels need about one frame.  A
// single-frame launch passes it by value; a batched launch reads an array of
// them from device memory and maps each workgroup / tile to its frame.
struct FrameDev {
  const float* pos;
  const float* opac;
  const float* cov3;
  const void* shs;
  float2* xy;        // FP32 features (may be null)
  float4* rgbd;
  float4* conic;
  FeatRec16* rec;    // reduced-precision records (may be null)
  uint8_t* out;      // u8[H][W][3]
  int P;
  int rec_prec;
  int block0;        // first preprocess workgroup of this frame in the launch
  int tile0;         // first global tile id of this frame
  int tiles;         // number of 16x16 tiles of this frame
  float bg[3];
  CamParams cp;
  // When set, the camera is read from device memory instead of `cp` (frames
  // replayed from a captured graph, whose kernel arguments are fixed).
  const CamParams* cam_dev;
};

// The frame's camera: `cp`, or the device copy a replayed graph points at.
__device__ __forceinline__ void bind_camera(FrameDev& F) {
  if (F.cam_dev) F.cp = *F.cam_dev;
}

// Top-k cache (render.cu:396-513, :600-623, example.py:168-197): the FP32
// features of the most-referenced Gaussians in compact arrays; a sorted
// value with bit 31 set names slot (value & 0x7fffffff) of them instead of a
// Gaussian id.  The compact arrays (40 bytes per slot) stay resident in L2.
struct TopKCache {
  const float2* xy = nullptr;
  const float4* rgbd = nullptr;
  const float4* con = nullptr;
  int n = 0;
};

// Per-frame constants of a frame replayed from a captured graph, in device
// memory: the kernels read them instead of their (captured) arguments.
struct FrameConsts {
  CamParams cp;
  float bg[4];
  uint8_t* out;  // u8[H][W][3]
};

// Largest f with key >= field(frames[f]) (frames sorted by that field).
template <typename Field>
__device__ __forceinline__ int frame_search(const FrameDev* frames, int nf, int key, Field field) {
  int lo = 0, hi = nf - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (field(frames[mid]) <= key) lo = mid; else hi = mid - 1;
  }
  return lo;
}

}  // namespace ps
