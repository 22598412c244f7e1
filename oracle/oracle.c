/*
 * CPU oracle for the forward 3DGS rasterizer -- TEST INFRASTRUCTURE ONLY
 * (see oracle.h).  Every function cites the reference lines it restates.
 *
 * FP32 parts follow the reference's operation order literally; the build
 * disables contraction, so each `a*b + c` rounds twice, exactly as written.
 * The reference's approximate hardware instructions (lg2/rsqrt/sqrt/ex2
 * .approx, preprocess.cu:40-66, render.cu:129-134) are replaced by
 * correctly rounded sqrt/division and a series log2 defined below; the GPU
 * build makes the same substitution, so tile decisions can match bit for bit.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static float bits2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* ------------------------------------------------------------------ halves */
float orc_f16_to_f32(uint16_t h) {
  uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 0x1f, m = h & 0x3ff;
  if (e == 0) {
    if (m == 0) return bits2f(s);
    /* subnormal: m * 2^-24 */
    float v = (float)m * 5.9604644775390625e-8f;
    return s ? -v : v;
  }
  if (e == 31) return bits2f(s | 0x7f800000u | (m << 13));
  return bits2f(s | ((e + 112) << 23) | (m << 13));
}

uint16_t orc_f32_to_f16(float f) {
  uint32_t u = f2bits(f);
  uint32_t s = (u >> 16) & 0x8000u;
  uint32_t a = u & 0x7fffffffu;
  if (a > 0x7f800000u) return (uint16_t)(s | 0x7e00u);  /* NaN */
  if (a >= 0x477ff000u) return (uint16_t)(s | 0x7c00u); /* rounds to inf */
  if (a < 0x38800000u) {                                  /* below 2^-14: subnormal half */
    /* value / 2^-24, rounded to nearest even integer */
    double v = (double)bits2f(a) * 16777216.0;
    double r = nearbyint(v); /* default rounding mode: nearest even */
    return (uint16_t)(s | (uint32_t)r);
  }
  uint32_t e = ((a >> 23) - 112) << 10;
  uint32_t m = (a >> 13) & 0x3ffu;
  uint32_t rest = a & 0x1fffu;
  uint32_t h = e | m;
  if (rest > 0x1000u || (rest == 0x1000u && (m & 1u))) h += 1;
  return (uint16_t)(s | h);
}

float orc_bf16_to_f32(uint16_t h) { return bits2f((uint32_t)h << 16); }

uint16_t orc_f32_to_bf16(float f) {
  uint32_t u = f2bits(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

/* ------------------------------------------------------------------ camera */
/* getViewMatrix / getProjectionMatrix, preprocess.cu:643-671, with the
 * rotation tensor read as the columns of a column-major matrix
 * (pybind.cpp:140-143).  Matrices are column-major: m[col*4 + row]. */
void orc_camera_matrices(const orc_camera* c, float view[16], float proj[16], float tan_fov[2]) {
  float R[3][3]; /* R[col][row] */
  int i, j, k;
  for (i = 0; i < 3; ++i)
    for (j = 0; j < 3; ++j) R[i][j] = c->rotation[i * 3 + j];
  float np[3] = {-c->position[0], -c->position[1], -c->position[2]};
  float t[3];
  for (j = 0; j < 3; ++j) t[j] = R[0][j] * np[0] + R[1][j] * np[1] + R[2][j] * np[2];
  for (i = 0; i < 3; ++i) {
    for (j = 0; j < 3; ++j) view[i * 4 + j] = R[i][j];
    view[i * 4 + 3] = 0.0f;
  }
  view[12] = t[0]; view[13] = t[1]; view[14] = t[2]; view[15] = 1.0f;

  float n = c->z_near, f = c->z_far;
  float top = (float)c->height / (2.0f * c->focal_y) * n;
  float bottom = -top;
  float right = (float)c->width / (2.0f * c->focal_x) * n;
  float left = -right;
  float P[4][4];
  memset(P, 0, sizeof(P));
  P[0][0] = 2.0f * n / (right - left);
  P[1][1] = 2.0f * n / (top - bottom);
  P[0][2] = (right + left) / (right - left);
  P[1][2] = (top + bottom) / (top - bottom);
  P[3][2] = 1.0f;
  P[2][2] = 1.0f * f / (f - n);
  P[2][3] = -(f * n) / (f - n);
  /* proj = transpose(P) * view, column by column: out[j] = sum_k A[k]*V[j][k] */
  for (j = 0; j < 4; ++j) {
    for (i = 0; i < 4; ++i) {
      float acc = P[i][0] * view[j * 4 + 0];
      acc = acc + P[i][1] * view[j * 4 + 1];
      acc = acc + P[i][2] * view[j * 4 + 2];
      acc = acc + P[i][3] * view[j * 4 + 3];
      proj[j * 4 + i] = acc;
    }
  }
  (void)k;
  tan_fov[0] = (float)c->width / (2.0f * c->focal_x);
  tan_fov[1] = (float)c->height / (2.0f * c->focal_y);
}

/* ------------------------------------------------------------------ scalar math */
/* log2 by the atanh series: x = 2^e m, m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 (s + s^3/3 + s^5/5 + s^7/7 + s^9/9), s = (m - 1)/(m + 1). */
float orc_log2(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
  if (isinf(x)) return x;
  int e_adj = 0;
  if (x < 1.17549435e-38f) {
    x = x * 8388608.0f;
    e_adj = -23;
  }
  uint32_t b = f2bits(x);
  int e = (int)((b >> 23) & 0xffu) - 127 + e_adj;
  float m = bits2f((b & 0x007fffffu) | 0x3f800000u);
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

/* (int) cast with NaN -> 0 and saturation (the GPU conversion semantics). */
static int trunc_i(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return -2147483647 - 1;
  return (int)x;
}

static float minf(float a, float b) { return a < b ? a : b; }
static float maxf(float a, float b) { return a > b ? a : b; }
static int mini(int a, int b) { return a < b ? a : b; }
static int maxi(int a, int b) { return a > b ? a : b; }

/* transformPoint4x3, preprocess.cu:84-92 */
static void xf43(const float* m, float x, float y, float z, float out[3]) {
  out[0] = m[0] * x + m[4] * y + m[8] * z + m[12];
  out[1] = m[1] * x + m[5] * y + m[9] * z + m[13];
  out[2] = m[2] * x + m[6] * y + m[10] * z + m[14];
}

/* computeCov2D, preprocess.cu:117-157: cov = T^T Vrk^T T with T = W J,
 * matrix products in column-major order (Result[c][r] = sum_k A[k][r] B[c][k],
 * summed left to right). */
static void cov2d_f32(float px, float py, float pz, float fx, float fy, float tfx, float tfy,
                      const float* s, const float* v, float out[3]) {
  float t[3];
  xf43(v, px, py, pz, t);
  float limx = 1.3f * tfx, limy = 1.3f * tfy;
  float txtz = t[0] / t[2], tytz = t[1] / t[2];
  t[0] = minf(limx, maxf(-limx, txtz)) * t[2];
  t[1] = minf(limy, maxf(-limy, tytz)) * t[2];
  float J[3][3] = {{fx / t[2], 0.0f, -(fx * t[0]) / (t[2] * t[2])},
                   {0.0f, fy / t[2], -(fy * t[1]) / (t[2] * t[2])},
                   {0.0f, 0.0f, 0.0f}};
  float W[3][3] = {{v[0], v[4], v[8]}, {v[1], v[5], v[9]}, {v[2], v[6], v[10]}};
  float T[3][3], Vk[3][3] = {{s[0], s[1], s[2]}, {s[1], s[3], s[4]}, {s[2], s[4], s[5]}};
  float Tt[3][3], VkT[3][3], A[3][3], C[3][3];
  int c, r;
  for (c = 0; c < 3; ++c)
    for (r = 0; r < 3; ++r) T[c][r] = W[0][r] * J[c][0] + W[1][r] * J[c][1] + W[2][r] * J[c][2];
  for (c = 0; c < 3; ++c)
    for (r = 0; r < 3; ++r) {
      Tt[c][r] = T[r][c];
      VkT[c][r] = Vk[r][c];
    }
  for (c = 0; c < 3; ++c)
    for (r = 0; r < 3; ++r)
      A[c][r] = Tt[0][r] * VkT[c][0] + Tt[1][r] * VkT[c][1] + Tt[2][r] * VkT[c][2];
  for (c = 0; c < 3; ++c)
    for (r = 0; r < 3; ++r) C[c][r] = A[0][r] * T[c][0] + A[1][r] * T[c][1] + A[2][r] * T[c][2];
  C[0][0] += 0.3f;
  C[1][1] += 0.3f;
  out[0] = C[0][0];
  out[1] = C[0][1];
  out[2] = C[1][1];
}

void orc_cov2d(const float p[3], const float cov3[6], const orc_camera* c, float out[3]) {
  float view[16], proj[16], tf[2];
  orc_camera_matrices(c, view, proj, tf);
  cov2d_f32(p[0], p[1], p[2], c->focal_x, c->focal_y, tf[0], tf[1], cov3, view, out);
}

/* computeColorFromSH, preprocess.cu:159-199 (real SH to degree 3). */
static const float SH_C0 = 0.28209479177387814f;
static const float SH_C1 = 0.4886025119029199f;
static const float SH_C2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                               -1.0925484305920792f, 0.5462742152960396f};
static const float SH_C3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                               0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                               -0.5900435899266435f};

void orc_sh_color(const float p[3], const float cp[3], const float sh[48], float out[3]) {
  float d0 = p[0] - cp[0], d1 = p[1] - cp[1], d2 = p[2] - cp[2];
  float l2 = d0 * d0 + d1 * d1 + d2 * d2;
  float rs = 1.0f / sqrtf(l2);
  float x = d0 * rs, y = d1 * rs, z = d2 * rs;
  float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
  int ch;
  for (ch = 0; ch < 3; ++ch) {
    const float* v = sh + ch; /* coefficient k at v[3k] */
    /* basis value of coefficient k, each product in the reference's
     * left-to-right order; the degree-1 terms it subtracts are negated */
    const float b[16] = {SH_C0, -(SH_C1 * y), SH_C1 * z, -(SH_C1 * x),
                         SH_C2[0] * xy, SH_C2[1] * yz, SH_C2[2] * (2.0f * zz - xx - yy),
                         SH_C2[3] * xz, SH_C2[4] * (xx - yy),
                         SH_C3[0] * y * (3.0f * xx - yy), SH_C3[1] * xy * z,
                         SH_C3[2] * y * (4.0f * zz - xx - yy),
                         SH_C3[3] * z * (2.0f * zz - 3.0f * xx - 3.0f * yy),
                         SH_C3[4] * x * (4.0f * zz - xx - yy), SH_C3[5] * z * (xx - yy),
                         SH_C3[6] * x * (xx - 3.0f * yy)};
    /* the sum r = 0.5 + sum_k b[k] * v[3k], accumulated in term order with one
     * correctly rounded fused multiply-add per term (the contraction the
     * reference's compiler applies to `r - b * sh + ...`) */
    float r = fmaf(b[0], v[0], 0.5f);
    int k;
    for (k = 1; k < 16; ++k) r = fmaf(b[k], v[3 * k], r);
    out[ch] = r > 0.0f ? r : 0.0f;
  }
}

/* segment_intersect_ellipse / block_intersect_ellipse / block_contains_center,
 * preprocess.cu:257-311. */
static int seg_hit(float a, float b, float c, float d, float l, float r) {
  float delta = b * b - 4.0f * a * c;
  float t1 = (l - d) * (2.0f * a) + b;
  float t2 = (r - d) * (2.0f * a) + b;
  return delta >= 0.0f && (t1 <= 0.0f || t1 * t1 <= delta) && (t2 >= 0.0f || t2 * t2 <= delta);
}

int orc_tile_hits(int tx, int ty, float cx, float cy, float ca, float cb, float cc, float power,
                  int bx, int by) {
  int x0 = tx * bx, y0 = ty * by, x1 = x0 + bx - 1, y1 = y0 + by - 1;
  if (cx >= (float)x0 && cx <= (float)x1 && cy >= (float)y0 && cy <= (float)y1) return 1;
  if (!(power > 0.0f)) return 0;
  float w = 2.0f * power;
  float dx = (cx * 2.0f < (float)(x0 + x1)) ? cx - (float)x0 : cx - (float)x1;
  if (seg_hit(cc, -2.0f * cb * dx, ca * dx * dx - w, cy, (float)y0, (float)y1)) return 1;
  float dy = (cy * 2.0f < (float)(y0 + y1)) ? cy - (float)y0 : cy - (float)y1;
  return seg_hit(ca, -2.0f * cb * dy, cc * dy * dy - w, cx, (float)x0, (float)x1);
}

/* ------------------------------------------------------------------ preprocess */
typedef struct {
  int valid;
  float x, y, ca, cb, cc, power, l2o, depth;
  int rx0, ry0, rx1, ry1;
} proj_t;

static void load_sh(const void* shs, int type, long long g, float out[48]) {
  int k;
  if (type == 0) {
    memcpy(out, (const float*)shs + g * 48, 48 * sizeof(float));
  } else {
    const uint16_t* h = (const uint16_t*)shs + g * 48;
    for (k = 0; k < 48; ++k) out[k] = type == 1 ? orc_f16_to_f32(h[k]) : orc_bf16_to_f32(h[k]);
  }
}

/* preprocess.cu:354-414 */
static proj_t project(const float* pos, const float* opac, const float* cov3, long long g,
                      const orc_camera* c, const float* view, const float* proj, const float* tf,
                      int gx, int gy) {
  proj_t o;
  memset(&o, 0, sizeof(o));
  float px = pos[3 * g], py = pos[3 * g + 1], pz = pos[3 * g + 2];
  float pv[3];
  xf43(view, px, py, pz, pv);
  o.depth = pv[2];
  if (pv[2] <= 0.2f) return o;
  float op = opac[g];
  float ph[3];
  xf43(proj, px, py, pz, ph);
  float phw = proj[3] * px + proj[7] * py + proj[11] * pz + proj[15];
  float pw = 1.0f / (phw + 0.0000001f);
  float nx = ph[0] * pw, ny = ph[1] * pw;
  float cov[3];
  cov2d_f32(px, py, pz, c->focal_x, c->focal_y, tf[0], tf[1], cov3 + 6 * g, view, cov);
  float det = cov[0] * cov[2] - cov[1] * cov[1];
  float det_inv = 1.0f / det;
  o.ca = cov[2] * det_inv;
  o.cb = -cov[1] * det_inv;
  o.cc = cov[0] * det_inv;
  o.l2o = orc_log2(op);
  o.power = 0.69314718055f * 8.0f + 0.69314718055f * o.l2o;
  float trace = cov[0] + cov[2];
  if (2.0f * det * o.power < 0.09f * trace && op < 0.5f) return o;
  int wdt = trunc_i(1.414214f * sqrtf(cov[0] * o.power) + 1.0f);
  int hgt = trunc_i(1.414214f * sqrtf(cov[2] * o.power) + 1.0f);
  /* ndc2Pix in double, preprocess.cu:78-81 */
  o.x = (float)((((double)nx + 1.0) * (double)c->width - 1.0) * 0.5);
  o.y = (float)((((double)ny + 1.0) * (double)c->height - 1.0) * 0.5);
  float bxf = (float)c->block_x, byf = (float)c->block_y;
  o.rx0 = mini(gx, maxi(0, trunc_i((o.x - (float)wdt) / bxf)));
  o.ry0 = mini(gy, maxi(0, trunc_i((o.y - (float)hgt) / byf)));
  o.rx1 = mini(gx, maxi(0, trunc_i((o.x + (float)wdt) / bxf) + 1));
  o.ry1 = mini(gy, maxi(0, trunc_i((o.y + (float)hgt) / byf) + 1));
  o.valid = (o.rx1 - o.rx0) * (o.ry1 - o.ry0) > 0;
  return o;
}

long long orc_preprocess(int P, const float* pos, const void* shs, int sh_type, const float* opac,
                         const float* cov3, const orc_camera* c, float* xy, float* rgbd,
                         float* conic, int* counts, uint64_t* keys, uint32_t* vals,
                         long long cap) {
  float view[16], proj[16], tf[2];
  orc_camera_matrices(c, view, proj, tf);
  int gx = (c->width + c->block_x - 1) / c->block_x;
  int gy = (c->height + c->block_y - 1) / c->block_y;
  long long g;
  int* cnt = counts ? counts : (int*)malloc(sizeof(int) * (size_t)(P > 0 ? P : 1));
  /* pass 1: count exact tiles, write features */
#pragma omp parallel for schedule(dynamic, 4096)
  for (g = 0; g < P; ++g) {
    proj_t o = project(pos, opac, cov3, g, c, view, proj, tf, gx, gy);
    int n = 0, x, y;
    if (o.valid)
      for (y = o.ry0; y < o.ry1; ++y)
        for (x = o.rx0; x < o.rx1; ++x)
          n += orc_tile_hits(x, y, o.x, o.y, o.ca, o.cb, o.cc, o.power, c->block_x, c->block_y);
    cnt[g] = n;
    if (n > 0) {
      float sh[48], rgb[3], p[3] = {pos[3 * g], pos[3 * g + 1], pos[3 * g + 2]};
      load_sh(shs, sh_type, g, sh);
      orc_sh_color(p, c->position, sh, rgb);
      if (xy) { xy[2 * g] = o.x; xy[2 * g + 1] = o.y; }
      if (conic) {
        conic[4 * g + 0] = (-0.5f * 1.4426950216293334961f) * o.ca;
        conic[4 * g + 1] = (-1.4426950216293334961f) * o.cb;
        conic[4 * g + 2] = (-0.5f * 1.4426950216293334961f) * o.cc;
        conic[4 * g + 3] = o.l2o;
      }
      if (rgbd) {
        rgbd[4 * g + 0] = rgb[0]; rgbd[4 * g + 1] = rgb[1];
        rgbd[4 * g + 2] = rgb[2]; rgbd[4 * g + 3] = o.depth;
      }
    }
  }
  /* exclusive scan of counts -> canonical offsets */
  long long total = 0;
  long long* off = (long long*)malloc(sizeof(long long) * (size_t)(P > 0 ? P : 1));
  for (g = 0; g < P; ++g) {
    off[g] = total;
    total += cnt[g];
  }
  if (keys && vals && total <= cap) {
#pragma omp parallel for schedule(dynamic, 4096)
    for (g = 0; g < P; ++g) {
      if (cnt[g] == 0) continue;
      proj_t o = project(pos, opac, cov3, g, c, view, proj, tf, gx, gy);
      long long k = off[g];
      int x, y;
      for (y = o.ry0; y < o.ry1; ++y)
        for (x = o.rx0; x < o.rx1; ++x)
          if (orc_tile_hits(x, y, o.x, o.y, o.ca, o.cb, o.cc, o.power, c->block_x, c->block_y)) {
            keys[k] = ((uint64_t)(uint32_t)(y * gx + x) << 32) | (uint64_t)f2bits(o.depth);
            vals[k] = (uint32_t)g;
            ++k;
          }
    }
  }
  free(off);
  if (!counts) free(cnt);
  return total;
}

/* ------------------------------------------------------------------ sort */
void orc_sort_pairs(long long n, uint64_t* keys, uint32_t* vals) {
  if (n < 2) return;
  uint64_t* tk = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
  uint32_t* tv = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
  uint64_t *sk = keys, *dk = tk;
  uint32_t *sv = vals, *dv = tv;
  long long w;
  for (w = 1; w < n; w *= 2) {
    long long lo;
#pragma omp parallel for schedule(static)
    for (lo = 0; lo < n; lo += 2 * w) {
      long long mid = lo + w < n ? lo + w : n;
      long long hi = lo + 2 * w < n ? lo + 2 * w : n;
      long long i = lo, j = mid, k = lo;
      while (i < mid && j < hi) {
        if (sk[j] < sk[i]) { dk[k] = sk[j]; dv[k++] = sv[j++]; }
        else { dk[k] = sk[i]; dv[k++] = sv[i++]; }  /* ties: left run first (stable) */
      }
      while (i < mid) { dk[k] = sk[i]; dv[k++] = sv[i++]; }
      while (j < hi) { dk[k] = sk[j]; dv[k++] = sv[j++]; }
    }
    uint64_t* t1 = sk; sk = dk; dk = t1;
    uint32_t* t2 = sv; sv = dv; dv = t2;
  }
  if (sk != keys) {
    memcpy(keys, sk, sizeof(uint64_t) * (size_t)n);
    memcpy(vals, sv, sizeof(uint32_t) * (size_t)n);
  }
  free(tk);
  free(tv);
}

/* ------------------------------------------------------------------ ranges */
void orc_tile_ranges(long long n, const uint64_t* keys, int tiles, int* ranges, int mode) {
  int t;
  long long i;
  if (mode == 1) {
    memset(ranges, 0, sizeof(int) * 2 * (size_t)tiles);
    for (i = 0; i < n; ++i) {
      uint32_t cur = (uint32_t)(keys[i] >> 32);
      if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != cur) ranges[2 * cur] = (int)i;
      if (i == n - 1 || (uint32_t)(keys[i + 1] >> 32) != cur) ranges[2 * cur + 1] = (int)(i + 1);
    }
    return;
  }
  /* linear sweep giving [lower_bound, upper_bound) for every tile */
  i = 0;
  for (t = 0; t < tiles; ++t) {
    while (i < n && (uint32_t)(keys[i] >> 32) < (uint32_t)t) ++i;
    long long j = i;
    while (j < n && (uint32_t)(keys[j] >> 32) == (uint32_t)t) ++j;
    ranges[2 * t] = (int)i;
    ranges[2 * t + 1] = (int)j;
    i = j;
  }
}

/* ------------------------------------------------------------- tile order */
/* Morton code: bit i of x to bit 2i, bit i of y to bit 2i+1 (render.cu:4430-4441),
 * one bit at a time. */
uint32_t orc_morton(uint32_t x, uint32_t y) {
  uint32_t z = 0;
  int b;
  for (b = 0; b < 16; ++b) {
    z |= ((x >> b) & 1u) << (2 * b);
    z |= ((y >> b) & 1u) << (2 * b + 1);
  }
  return z;
}

/* extractTileCounts / extractTileCountsMorton (render.cu:4414-4427, :4496-4518),
 * SortPairsDescending (render.cu:6974-6978: stable, so ties keep ascending
 * tile ids), gatherTileDesc / gatherTileDescZigzag (render.cu:4812-4866). */
void orc_tile_order(const int* ranges, int tiles, int gx, int zigzag, int bucket,
                    uint32_t* order, uint32_t* keys_sorted, int* desc) {
  uint32_t* key = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(tiles > 0 ? tiles : 1));
  uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(tiles > 0 ? tiles : 1));
  int t, p, w;
  for (t = 0; t < tiles; ++t) {
    uint32_t count = (uint32_t)(ranges[2 * t + 1] - ranges[2 * t]);
    if (bucket > 0) {
      uint32_t b = count / (uint32_t)bucket;
      if (b > 0xffffu) b = 0xffffu;
      uint32_t m = orc_morton((uint32_t)(t % gx), (uint32_t)(t / gx)) & 0xffffu;
      key[t] = (b << 16) | (0xffffu - m);
    } else {
      key[t] = count;
    }
    order[t] = (uint32_t)t;
  }
  /* bottom-up merge sort by key, descending; ties: left run first (stable) */
  for (w = 1; w < tiles; w *= 2) {
    int lo;
    for (lo = 0; lo < tiles; lo += 2 * w) {
      int mid = lo + w < tiles ? lo + w : tiles;
      int hi = lo + 2 * w < tiles ? lo + 2 * w : tiles;
      int i = lo, j = mid, k = lo;
      while (i < mid && j < hi) {
        if (key[order[j]] > key[order[i]]) tmp[k++] = order[j++];
        else tmp[k++] = order[i++];
      }
      while (i < mid) tmp[k++] = order[i++];
      while (j < hi) tmp[k++] = order[j++];
    }
    memcpy(order, tmp, sizeof(uint32_t) * (size_t)tiles);
  }
  for (p = 0; p < tiles; ++p) keys_sorted[p] = key[order[p]];
  if (desc) {
    for (p = 0; p < tiles; ++p) {
      int src = p;
      if (zigzag > 1) {
        int group = p / zigzag, slot = p % zigzag, base = group * zigzag;
        int end = base + zigzag < tiles ? base + zigzag : tiles;
        src = (group & 1) ? base + (end - base - 1 - slot) : base + slot;
        if (src >= tiles) src = tiles - 1;
      }
      t = (int)order[src];
      desc[4 * p + 0] = t % gx;
      desc[4 * p + 1] = t / gx;
      desc[4 * p + 2] = ranges[2 * t];
      desc[4 * p + 3] = ranges[2 * t + 1];
    }
  }
  free(key);
  free(tmp);
}

/* ------------------------------------------------------------------ render */
static uint8_t enc(float x) {
  float v = x * 255.0f;
  v = v > 0.0f ? v : 0.0f;
  v = v < 255.0f ? v : 255.0f;
  return (uint8_t)v;
}

/* pixel_shader_axis_shared_16x16 (render.cu:228-271), write_color
 * (render.cu:375-394), saturation cadence (render.cu:1814-1827). */
void orc_render_f32(int W, int H, const int* ranges, const uint32_t* vals, const float* xy,
                    const float* rgbd, const float* conic, const float bg[3], int check_every,
                    int dt, uint8_t* out, float* out_f, int* stop, int* need) {
  int gx = (W + 15) / 16, gy = (H + 15) / 16, tiles = gx * gy, t;
#pragma omp parallel for schedule(dynamic, 1)
  for (t = 0; t < tiles; ++t) {
    int tx = t % gx, ty = t / gx;
    float T[256], C[256][3];
    int last[256];
    int p, k;
    for (p = 0; p < 256; ++p) { T[p] = 1.0f; C[p][0] = C[p][1] = C[p][2] = 0.0f; last[p] = -1; }
    int s = ranges[2 * t], e = ranges[2 * t + 1];
    int n = e - s, done = 0;
    for (k = 0; k < n && !done; ++k) {
      uint32_t g = vals[s + k];
      float gxx = xy[2 * g], gyy = xy[2 * g + 1];
      const float* q = conic + 4 * (size_t)g;
      const float* c = rgbd + 4 * (size_t)g;
      for (p = 0; p < 256; ++p) {
        int px = tx * 16 + (p & 15), py = ty * 16 + (p >> 4);
        float dx = gxx - (float)px, dy = gyy - (float)py;
        float xs = q[0] * dx * dx, xc = q[1] * dx, ys = q[2] * dy * dy;
        float pw = fmaf(xc, dy, q[3] + xs + ys);
        if (dt && pw < -7.994353375f) continue;
        float a = exp2f(pw);
        if (T[p] >= 0.0001f) last[p] = k;
        float wgt = a * T[p];
        C[p][0] += c[0] * wgt;
        C[p][1] += c[1] * wgt;
        C[p][2] += c[2] * wgt;
        T[p] -= wgt;
      }
      if ((k % check_every) == check_every - 1) {
        int sat = 1;
        for (p = 0; p < 256 && sat; ++p) sat = T[p] < 0.0001f;
        if (sat) done = 1;
      }
    }
    if (stop) stop[t] = k;
    if (need) {
      int m = -1;
      for (p = 0; p < 256; ++p) m = last[p] > m ? last[p] : m;
      need[t] = m + 1;
    }
    for (p = 0; p < 256; ++p) {
      int px = tx * 16 + (p & 15), py = ty * 16 + (p >> 4);
      if (px >= W || py >= H) continue;
      size_t o = ((size_t)py * W + px) * 3;
      int ch;
      if (n > 0) {
        float tt = T[p] < 0.0001f ? 0.0f : T[p];
        for (ch = 0; ch < 3; ++ch) {
          float v = C[p][ch] + tt * bg[ch];
          out[o + ch] = enc(v);
          if (out_f) out_f[o + ch] = v;
        }
      } else {
        for (ch = 0; ch < 3; ++ch) {
          out[o + ch] = enc(bg[ch]);
          if (out_f) out_f[o + ch] = bg[ch];
        }
      }
    }
  }
  (void)gy;
}

/* ------------------------------------------------------------------ FP64 */
static void xf43d(const double* m, double x, double y, double z, double out[3]) {
  out[0] = m[0] * x + m[4] * y + m[8] * z + m[12];
  out[1] = m[1] * x + m[5] * y + m[9] * z + m[13];
  out[2] = m[2] * x + m[6] * y + m[10] * z + m[14];
}

void orc_features_f64(int n_ids, const uint32_t* ids, const float* pos, const void* shs,
                      int sh_type, const float* opac, const float* cov3, const orc_camera* c,
                      double* xy, double* conic, double* rgb) {
  /* Camera matrices in double: view = [R | -R c], projection reduced to the
   * pinhole map the reference's matrices encode (x_ndc = 2 fx x / (W z)). */
  double R[3][3];
  int i, j;
  for (i = 0; i < 3; ++i)
    for (j = 0; j < 3; ++j) R[i][j] = c->rotation[i * 3 + j]; /* R[col][row] */
  double v[16];
  memset(v, 0, sizeof(v));
  for (i = 0; i < 3; ++i)
    for (j = 0; j < 3; ++j) v[i * 4 + j] = R[i][j];
  for (j = 0; j < 3; ++j)
    v[12 + j] = -(R[0][j] * c->position[0] + R[1][j] * c->position[1] + R[2][j] * c->position[2]);
  v[15] = 1.0;
  double fx = c->focal_x, fy = c->focal_y, W = c->width, H = c->height;
  double tfx = W / (2.0 * fx), tfy = H / (2.0 * fy);
  int k;
#pragma omp parallel for schedule(static)
  for (k = 0; k < n_ids; ++k) {
    long long g = ids[k];
    double p[3] = {pos[3 * g], pos[3 * g + 1], pos[3 * g + 2]};
    double t[3];
    xf43d(v, p[0], p[1], p[2], t);
    double z = t[2];
    /* pixel centre: ((x_ndc + 1) W - 1) / 2 with x_ndc = (2 fx / W) x / z */
    double pw = 1.0 / (z + 1e-7);
    xy[2 * k] = ((2.0 * fx / W) * t[0] * pw + 1.0) * W * 0.5 - 0.5;
    xy[2 * k + 1] = ((2.0 * fy / H) * t[1] * pw + 1.0) * H * 0.5 - 0.5;
    double limx = 1.3 * tfx, limy = 1.3 * tfy;
    double tx = fmin(limx, fmax(-limx, t[0] / z)) * z;
    double ty = fmin(limy, fmax(-limy, t[1] / z)) * z;
    double J[2][3] = {{fx / z, 0.0, -fx * tx / (z * z)}, {0.0, fy / z, -fy * ty / (z * z)}};
    /* Wr = rows of the view rotation: T = J * Wr (2x3), cov = T Sigma T^T */
    double Wr[3][3];
    for (i = 0; i < 3; ++i)
      for (j = 0; j < 3; ++j) Wr[i][j] = v[j * 4 + i];
    double Tm[2][3];
    int a, b;
    for (a = 0; a < 2; ++a)
      for (b = 0; b < 3; ++b) Tm[a][b] = J[a][0] * Wr[0][b] + J[a][1] * Wr[1][b] + J[a][2] * Wr[2][b];
    const float* s = cov3 + 6 * g;
    double S[3][3] = {{s[0], s[1], s[2]}, {s[1], s[3], s[4]}, {s[2], s[4], s[5]}};
    double TS[2][3], cv[2][2];
    for (a = 0; a < 2; ++a)
      for (b = 0; b < 3; ++b) TS[a][b] = Tm[a][0] * S[0][b] + Tm[a][1] * S[1][b] + Tm[a][2] * S[2][b];
    for (a = 0; a < 2; ++a)
      for (b = 0; b < 2; ++b) cv[a][b] = TS[a][0] * Tm[b][0] + TS[a][1] * Tm[b][1] + TS[a][2] * Tm[b][2];
    cv[0][0] += 0.3;
    cv[1][1] += 0.3;
    double det = cv[0][0] * cv[1][1] - cv[0][1] * cv[0][1];
    double ca = cv[1][1] / det, cb = -cv[0][1] / det, cc = cv[0][0] / det;
    double l2e = 1.0 / log(2.0);
    conic[4 * k + 0] = -0.5 * l2e * ca;
    conic[4 * k + 1] = -l2e * cb;
    conic[4 * k + 2] = -0.5 * l2e * cc;
    conic[4 * k + 3] = log2((double)opac[g]);
    /* SH in double */
    float shf[48];
    load_sh(shs, sh_type, g, shf);
    double d0 = p[0] - c->position[0], d1 = p[1] - c->position[1], d2 = p[2] - c->position[2];
    double nrm = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    double x = d0 / nrm, y = d1 / nrm, zz0 = d2 / nrm;
    double xx = x * x, yy = y * y, zz = zz0 * zz0;
    double basis[16];
    basis[0] = 0.28209479177387814;
    basis[1] = -0.4886025119029199 * y;
    basis[2] = 0.4886025119029199 * zz0;
    basis[3] = -0.4886025119029199 * x;
    basis[4] = 1.0925484305920792 * x * y;
    basis[5] = -1.0925484305920792 * y * zz0;
    basis[6] = 0.31539156525252005 * (2.0 * zz - xx - yy);
    basis[7] = -1.0925484305920792 * x * zz0;
    basis[8] = 0.5462742152960396 * (xx - yy);
    basis[9] = -0.5900435899266435 * y * (3.0 * xx - yy);
    basis[10] = 2.890611442640554 * x * y * zz0;
    basis[11] = -0.4570457994644658 * y * (4.0 * zz - xx - yy);
    basis[12] = 0.3731763325901154 * zz0 * (2.0 * zz - 3.0 * xx - 3.0 * yy);
    basis[13] = -0.4570457994644658 * x * (4.0 * zz - xx - yy);
    basis[14] = 1.445305721320277 * zz0 * (xx - yy);
    basis[15] = -0.5900435899266435 * x * (xx - 3.0 * yy);
    int ch, m;
    for (ch = 0; ch < 3; ++ch) {
      double acc = 0.5;
      for (m = 0; m < 16; ++m) acc += basis[m] * (double)shf[3 * m + ch];
      rgb[3 * k + ch] = acc > 0.0 ? acc : 0.0;
    }
  }
}

void orc_render_f64(int W, int H, const int* ranges, const uint32_t* vals, const int* stop,
                    const int* slot, const double* xy, const double* conic, const double* rgb,
                    const float bg[3], uint8_t* out, double* out_f) {
  int gx = (W + 15) / 16, gy = (H + 15) / 16, tiles = gx * gy, t;
#pragma omp parallel for schedule(dynamic, 1)
  for (t = 0; t < tiles; ++t) {
    int tx = t % gx, ty = t / gx, p, k, ch;
    double T[256], C[256][3];
    for (p = 0; p < 256; ++p) { T[p] = 1.0; C[p][0] = C[p][1] = C[p][2] = 0.0; }
    int s = ranges[2 * t], n = ranges[2 * t + 1] - s;
    int m = stop ? stop[t] : n;
    for (k = 0; k < m; ++k) {
      int sl = slot[vals[s + k]];
      const double* q = conic + 4 * (size_t)sl;
      for (p = 0; p < 256; ++p) {
        double dx = xy[2 * (size_t)sl] - (double)(tx * 16 + (p & 15));
        double dy = xy[2 * (size_t)sl + 1] - (double)(ty * 16 + (p >> 4));
        double pw = q[3] + q[0] * dx * dx + q[1] * dx * dy + q[2] * dy * dy;
        double a = exp2(pw);
        double wgt = a * T[p];
        for (ch = 0; ch < 3; ++ch) C[p][ch] += rgb[3 * (size_t)sl + ch] * wgt;
        T[p] -= wgt;
      }
    }
    for (p = 0; p < 256; ++p) {
      int px = tx * 16 + (p & 15), py = ty * 16 + (p >> 4);
      if (px >= W || py >= H) continue;
      size_t o = ((size_t)py * W + px) * 3;
      double tt = (n > 0 && T[p] < 0.0001) ? 0.0 : T[p];
      for (ch = 0; ch < 3; ++ch) {
        double v = (n > 0) ? C[p][ch] + tt * bg[ch] : bg[ch];
        double u = v * 255.0;
        u = u > 0.0 ? u : 0.0;
        u = u < 255.0 ? u : 255.0;
        out[o + ch] = (uint8_t)u;
        if (out_f) out_f[o + ch] = v;
      }
    }
  }
  (void)gy;
}

void orc_quantize_features(int P, int prec, const float* rgbd, const float* conic, float* rgbd_q,
                           float* conic_q) {
  long long g;
#pragma omp parallel for schedule(static)
  for (g = 0; g < P; ++g) {
    int k;
    for (k = 0; k < 3; ++k) {
      float cs = conic[4 * g + k] * 4096.0f;
      float cq = prec == 2 ? orc_bf16_to_f32(orc_f32_to_bf16(cs)) : orc_f16_to_f32(orc_f32_to_f16(cs));
      conic_q[4 * g + k] = cq * (1.0f / 4096.0f);
      float r = rgbd[4 * g + k];
      rgbd_q[4 * g + k] =
          prec == 2 ? orc_bf16_to_f32(orc_f32_to_bf16(r)) : orc_f16_to_f32(orc_f32_to_f16(r));
    }
    conic_q[4 * g + 3] = conic[4 * g + 3];
    rgbd_q[4 * g + 3] = rgbd[4 * g + 3];
  }
}
