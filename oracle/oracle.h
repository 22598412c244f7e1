/*
 * CPU oracle for the forward 3DGS rasterizer -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of LuoLongZan/PolySplat's forward equations, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  Nothing in the product links or calls it.
 *
 * Build: oracle/Makefile (gcc, -ffp-contract=off, IEEE sqrt and division).
 */
#ifndef PS_ORACLE_H
#define PS_ORACLE_H

#include <stdint.h>

typedef struct orc_camera {
  int width, height, block_x, block_y;
  float position[3];
  float rotation[9];
  float focal_x, focal_y, z_far, z_near;
} orc_camera;

/* preprocess.cu:643-671, :1263-1266 */
void orc_camera_matrices(const orc_camera* c, float view[16], float proj[16], float tan_fov[2]);

/* Stand-alone scalar pieces exposed for closed-form tests. */
float orc_log2(float x);
void orc_cov2d(const float p[3], const float cov3[6], const orc_camera* c, float out[3]);
void orc_sh_color(const float p[3], const float campos[3], const float sh[48], float out[3]);
int orc_tile_hits(int tx, int ty, float cx, float cy, float ca, float cb, float cc, float power,
                  int bx, int by);

/* FP32 preprocess (preprocess.cu:313-641).  sh_type 0: f32, 1: f16 bits,
 * 2: bf16 bits.  Writes per-Gaussian features for Gaussians emitting >= 1
 * key (others untouched), per-Gaussian tile counts, and the keys/values in
 * canonical order (Gaussian id, then tile row-major) when they fit in cap.
 * Returns the total number of pairs. */
long long orc_preprocess(int P, const float* pos, const void* shs, int sh_type, const float* opac,
                         const float* cov3, const orc_camera* c, float* xy, float* rgbd,
                         float* conic, int* counts, uint64_t* keys, uint32_t* vals,
                         long long cap);

/* Stable sort of (key, value) pairs by key (bottom-up merge sort). */
void orc_sort_pairs(long long n, uint64_t* keys, uint32_t* vals);

/* mode 0: [lower_bound, upper_bound) per tile (empty -> (lo, lo));
 * mode 1: boundary scan (empty -> (0, 0)).  render.cu:3509-3557 */
void orc_tile_ranges(long long n, const uint64_t* keys, int tiles, int* ranges, int mode);

/* FP32 render (render.cu:228-271, :375-394).  Features are indexed through
 * vals.  check_every: saturation-test cadence.  Optional outputs: out_f
 * (H*W*3 float, C + T*bg before encoding), stop (per tile: Gaussians blended)
 * and need (per tile: 1 + index of the last Gaussian any pixel blends while
 * its T >= 1e-4, i.e. the cadence-free work bound). */
void orc_render_f32(int W, int H, const int* ranges, const uint32_t* vals, const float* xy,
                    const float* rgbd, const float* conic, const float bg[3], int check_every,
                    int dt, uint8_t* out, float* out_f, int* stop, int* need);

/* Per-Gaussian continuous quantities in double for the given ids:
 * xy[2], conic[4] (scaled like conic_opacity, .w = log2 opacity), rgb[3]. */
void orc_features_f64(int n_ids, const uint32_t* ids, const float* pos, const void* shs,
                      int sh_type, const float* opac, const float* cov3, const orc_camera* c,
                      double* xy, double* conic, double* rgb);

/* FP64 blend of the first stop[t] Gaussians of every tile, features indexed
 * by gaussian id through vals and a dense id -> slot map (slot[id] < 0 means
 * unused). */
void orc_render_f64(int W, int H, const int* ranges, const uint32_t* vals, const int* stop,
                    const int* slot, const double* xy, const double* conic, const double* rgb,
                    const float bg[3], uint8_t* out, double* out_f);

/* float -> half / bf16 bits (round to nearest even) and back. */
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);
uint16_t orc_f32_to_bf16(float f);
float orc_bf16_to_f32(uint16_t h);

/* Values the reduced-precision records hold once expanded to FP32:
 * conic (x, y, z scaled by 4096 before rounding) and RGB rounded to
 * f16 (prec 1) or bf16 (prec 2); xy and log2 opacity unchanged. */
void orc_quantize_features(int P, int prec, const float* rgbd, const float* conic,
                           float* rgbd_q, float* conic_q);

/* Tile order of the reordered renderers (render.cu:4411-4520, :4790-4866,
 * :6959-7087).  key[t] = count of tile t (ranges[2t+1] - ranges[2t]) or,
 * with bucket > 0, (min(count / bucket, 0xffff) << 16) |
 * (0xffff - (morton(t % gx, t / gx) & 0xffff)); order = tiles by key,
 * descending, ties by ascending tile id; keys_sorted[p] = key of order[p];
 * desc (may be NULL): 4 ints per position p = (column, row, range start,
 * range end) of order[zigzag(p)], zigzag reversing every odd group of
 * `zigzag` positions (no regrouping for zigzag <= 1). */
uint32_t orc_morton(uint32_t x, uint32_t y);
void orc_tile_order(const int* ranges, int tiles, int gx, int zigzag, int bucket,
                    uint32_t* order, uint32_t* keys_sorted, int* desc);

int orc_num_threads(void);

#endif
