/* CPU oracle for the splat frame path: TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Built with -ffp-contract=off so every float32 expression rounds where WGSL's does when evaluated left to
 * right without fusion. dot(a, b) is taken as ((a.x b.x + a.y b.y) + a.z b.z) + a.w b.w and m * v as
 * m[0] v.x + m[1] v.y + m[2] v.z (+ m[3] v.w), the order the HIP kernels use too.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define INVALID_KEY 0xffffffffu
#define NUM_BUCKETS 65536u
#define MAX_PARTS 32u
#define PART_HIDDEN 4u
#define EMITTER_MATERIAL 254u

/* RenderFrame word offsets (shaders/common.wgsl:44-77; src/main.js:526-556) */
enum {
  FR_VIEW = 0, FR_PALETTE = 32, FR_PART_MATERIAL = 48, FR_LAMP = 64, FR_LAMP_COLOR = 68, FR_VIEWPORT = 72,
  FR_FOCAL = 74, FR_NEAR = 76, FR_FAR = 77, FR_COUNT = 78, FR_VIEW_MODE = 79, FR_ALBEDO_SCALE = 80,
  FR_LOWPASS = 83, FR_PALETTE_EXT = 84, FR_PART_MATERIAL_EXT = 164, FR_PART_FADE = 264
};

static float fw(const void *frame, int k) { float v; memcpy(&v, (const char *)frame + 4 * k, 4); return v; }
static uint32_t uw(const void *frame, int k) { uint32_t v; memcpy(&v, (const char *)frame + 4 * k, 4); return v; }
static float bits_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* palette_slot (shaders/project.wgsl:27-31) */
static void palette_slot(const void *frame, uint32_t m, float out[4]) {
  int base = -1;
  if (m < 3u) base = FR_PALETTE + 4 * (int)m;
  else if (m >= 4u && m < 8u) base = FR_PALETTE_EXT + 4 * (int)(m - 4u);
  if (base < 0) { out[0] = out[1] = out[2] = 1.0f; out[3] = 0.0f; return; }
  for (int k = 0; k < 4; k++) out[k] = fw(frame, base + k);
}

static uint32_t part_material(const void *frame, uint32_t part) {
  if (part < 16u) return uw(frame, FR_PART_MATERIAL + (int)part);
  return uw(frame, FR_PART_MATERIAL_EXT + (int)(part - 16u));
}

uint32_t or_depth_key(float depth, float near_, float far_) {
  /* shaders/project.wgsl:366-367. log is the correctly rounded float log (via double). */
  float num = (float)log((double)(depth / near_));
  float den = (float)log((double)(far_ / near_));
  float t = num / den;
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  float s = t * 65535.0f;
  s = s + 0.5f;
  uint32_t k = 65535u - (uint32_t)s;
  return k < 0xfffeu ? k : 0xfffeu;
}

/* ---- project, float and double from one body ------------------------------------------------------------- */

/* float log as the correctly rounded value (double log rounded once), the definition the HIP kernels share */
static float logf_cr(float x) { return (float)log((double)x); }
static float expf_cr_(float x) { return (float)exp((double)x); }

#define REAL float
#define SUFFIX f32
#define SQRTR sqrtf
#define RLOG logf_cr
#define REXP expf_cr_
#include "project_body.inc"
#undef REAL
#undef SUFFIX
#undef SQRTR
#undef RLOG
#undef REXP

#define REAL double
#define SUFFIX f64
#define SQRTR sqrt
#define RLOG log
#define REXP exp
#include "project_body.inc"
#undef REXP
#undef REAL
#undef SUFFIX
#undef SQRTR
#undef RLOG

uint32_t or_project_f32(const void *frame, const void *gaussians, const void *parts, const float *colour,
                        const float *lo_b, const void *emitters, uint32_t n_emit, uint32_t *splats,
                        uint32_t *keys) {
  return or_project_lod_f32(frame, gaussians, parts, colour, lo_b, emitters, n_emit, NULL, splats, keys);
}

uint32_t or_project_lod_f32(const void *frame, const void *gaussians, const void *parts, const float *colour,
                            const float *lo_b, const void *emitters, uint32_t n_emit, const or_lod_view *lod,
                            uint32_t *splats, uint32_t *keys) {
  uint32_t count = uw(frame, FR_COUNT);
  uint32_t nslots = count + 1u + n_emit + (lod ? lod->n_nodes : 0u);
  uint32_t visible = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : visible)
  for (uint32_t i = 0; i < nslots; i++) {
    float s[12];
    float depth;
    uint32_t sflags = 0;
    keys[i] = INVALID_KEY;
    if (!project_one_f32(frame, gaussians, parts, colour, lo_b, emitters, n_emit, lod, i, s, &depth, &sflags))
      continue;
    for (int k = 0; k < 11; k++) splats[12 * i + k] = f_bits(s[k]);
    splats[12 * i + 11] = (f_bits(depth) & ~3u) | sflags;
    keys[i] = or_depth_key(depth, fw(frame, FR_NEAR), fw(frame, FR_FAR));
    visible++;
  }
  return visible;
}

void or_project_f64(const void *frame, const void *gaussians, const void *parts, const float *colour,
                    const float *lo_b, const void *emitters, uint32_t n_emit, double *out, uint8_t *valid) {
  uint32_t count = uw(frame, FR_COUNT);
  uint32_t nslots = count + 1u + n_emit;
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint32_t i = 0; i < nslots; i++) {
    double s[12];
    double depth = 0.0;
    uint32_t sflags = 0;
    valid[i] = (uint8_t)project_one_f64(frame, gaussians, parts, colour, lo_b, emitters, n_emit, NULL, i, s, &depth,
                                        &sflags);
    for (int k = 0; k < 11; k++) out[12 * i + k] = valid[i] ? s[k] : 0.0;
    out[12 * i + 11] = valid[i] ? depth : 0.0;
  }
}

/* ---- sort artefacts ---------------------------------------------------------------------------------------- */

uint32_t or_sort_keys(const uint32_t *keys, uint32_t n, uint32_t *hist, uint32_t *offsets, uint32_t *order) {
  /* shaders/project.wgsl:369 (histogram), shaders/scan.wgsl:16-43 (offsets, visible), shaders/radix.wgsl:96-182
   * (two stable 8-bit LSD passes on bits 0-15 == one stable counting sort on the low 16 bits). */
  uint32_t *h = (uint32_t *)calloc(NUM_BUCKETS, 4);
  uint32_t *o = (uint32_t *)malloc(NUM_BUCKETS * 4);
  uint32_t visible = 0;
  for (uint32_t i = 0; i < n; i++) if (keys[i] != INVALID_KEY) { h[keys[i]]++; visible++; }
  uint32_t run = 0;
  for (uint32_t b = 0; b < NUM_BUCKETS; b++) { o[b] = run; run += h[b]; }
  if (hist) memcpy(hist, h, NUM_BUCKETS * 4);
  if (offsets) memcpy(offsets, o, NUM_BUCKETS * 4);
  if (order) {
    /* stable placement on the low 16 bits: valid keys first, INVALID (low bits 0xFFFF) after them in index order */
    uint32_t *next = o;
    uint32_t inv = visible;
    for (uint32_t i = 0; i < n; i++) {
      if (keys[i] == INVALID_KEY) order[inv++] = i;
      else order[next[keys[i]]++] = i;
    }
  }
  free(h);
  free(o);
  return visible;
}

/* ---- tiles ------------------------------------------------------------------------------------------------- */

uint32_t or_tile_box(const uint32_t *s, uint32_t W, uint32_t H, uint32_t tile_px, uint32_t box[4]) {
  /* The splat's extent box (shaders/project.wgsl:314: centre +- (|axis0| + |axis1|), NDC) in pixel centres:
   * pixel (px, py) has its centre at NDC ((px + 0.5) 2 / W - 1, 1 - (py + 0.5) 2 / H). */
  float cx = bits_f(s[0]), cy = bits_f(s[1]);
  float ex = fabsf(bits_f(s[2])) + fabsf(bits_f(s[4]));
  float ey = fabsf(bits_f(s[3])) + fabsf(bits_f(s[5]));
  float hw = 0.5f * (float)W, hh = 0.5f * (float)H;
  float x0 = (cx - ex + 1.0f) * hw - 0.5f;
  float x1 = (cx + ex + 1.0f) * hw - 0.5f;
  float y0 = (1.0f - (cy + ey)) * hh - 0.5f;
  float y1 = (1.0f - (cy - ey)) * hh - 0.5f;
  float fx0 = ceilf(x0), fx1 = floorf(x1), fy0 = ceilf(y0), fy1 = floorf(y1);
  if (fx0 < 0.0f) fx0 = 0.0f;
  if (fy0 < 0.0f) fy0 = 0.0f;
  if (fx1 > (float)(W - 1)) fx1 = (float)(W - 1);
  if (fy1 > (float)(H - 1)) fy1 = (float)(H - 1);
  if (!(fx0 <= fx1) || !(fy0 <= fy1)) { box[0] = box[1] = 1; box[2] = box[3] = 0; return 0; }
  box[0] = (uint32_t)fx0 / tile_px;
  box[1] = (uint32_t)fy0 / tile_px;
  box[2] = (uint32_t)fx1 / tile_px;
  box[3] = (uint32_t)fy1 / tile_px;
  return (box[2] - box[0] + 1) * (box[3] - box[1] + 1);
}

uint64_t or_tile_lists(const uint32_t *splats, const uint32_t *order, uint32_t visible, uint32_t W, uint32_t H,
                       uint32_t tile_px, uint64_t *tile_start, uint32_t *vals) {
  uint32_t tw = (W + tile_px - 1) / tile_px, th = (H + tile_px - 1) / tile_px;
  uint64_t nt = (uint64_t)tw * th;
  uint64_t *cnt = (uint64_t *)calloc(nt + 1, 8);
  uint64_t P = 0;
  uint32_t box[4];
  for (uint32_t r = 0; r < visible; r++) {
    const uint32_t *s = splats + 12 * (size_t)order[r];
    if (!or_tile_box(s, W, H, tile_px, box)) continue;
    for (uint32_t ty = box[1]; ty <= box[3]; ty++)
      for (uint32_t tx = box[0]; tx <= box[2]; tx++) { cnt[(uint64_t)ty * tw + tx]++; P++; }
  }
  if (tile_start) {
    uint64_t run = 0;
    for (uint64_t t = 0; t < nt; t++) { tile_start[t] = run; run += cnt[t]; }
    tile_start[nt] = run;
    if (vals) {
      uint64_t *next = (uint64_t *)malloc(nt * 8);
      memcpy(next, tile_start, nt * 8);
      for (uint32_t r = 0; r < visible; r++) {
        const uint32_t *s = splats + 12 * (size_t)order[r];
        if (!or_tile_box(s, W, H, tile_px, box)) continue;
        for (uint32_t ty = box[1]; ty <= box[3]; ty++)
          for (uint32_t tx = box[0]; tx <= box[2]; tx++) vals[next[(uint64_t)ty * tw + tx]++] = order[r];
      }
      free(next);
    }
  }
  free(cnt);
  return P;
}

/* ---- composite --------------------------------------------------------------------------------------------- */

/* exp as the correctly rounded float value (double exp rounded once): the coverage test (alpha >= 1/255) is
 * then decided identically by every implementation that uses this definition. */
static float expf_cr(float x) { return (float)exp((double)x); }

static float to_half_rtne(float x) {
  /* IEEE binary16 round-to-nearest-even, returned widened (the rgba16float target's storage) */
  uint32_t u = f_bits(x);
  uint32_t sign = u & 0x80000000u;
  uint32_t a = u & 0x7fffffffu;
  if (a >= 0x7f800000u) return x;                       /* inf / nan */
  if (a >= 0x477ff000u) return bits_f(sign | 0x7f800000u); /* rounds to >= 65520 -> inf */
  if (a < 0x33000000u) return bits_f(sign);               /* below half the smallest subnormal */
  int e = (int)(a >> 23) - 127;
  int shift = e < -14 ? 13 + (-14 - e) : 13;             /* mantissa bits dropped */
  uint32_t m = (a & 0x7fffffu) | 0x800000u;
  uint32_t keep = m >> shift;
  uint32_t rem = m & ((1u << shift) - 1u);
  uint32_t halfway = 1u << (shift - 1);
  if (rem > halfway || (rem == halfway && (keep & 1u))) keep++;
  double v = ldexp((double)keep, e - 23 + shift);
  return (sign ? -1.0f : 1.0f) * (float)v;
}

void or_composite(const void *frame, const uint32_t *splats, const uint32_t *order, uint32_t visible,
                  uint32_t W, uint32_t H, int mode, float *out) {
  or_composite_rows(frame, splats, order, visible, W, H, mode, 1u, out);
}

/* The same blend on the rows py % row_step == 0 only (a bounded sample for timing); other rows stay 0. Threads take
 * bands of 16 rows; within a band every splat is visited in the back-to-front order, so each pixel's sequence of
 * blends (and its float result) is the serial one. */
void or_composite_rows(const void *frame, const uint32_t *splats, const uint32_t *order, uint32_t visible,
                       uint32_t W, uint32_t H, int mode, uint32_t row_step, float *out) {
  const float ALPHA_MIN = 1.0f / 255.0f;
  const uint32_t BAND = 16u;
  uint32_t view_mode = uw(frame, FR_VIEW_MODE);
  float near_ = fw(frame, FR_NEAR), far_ = fw(frame, FR_FAR);
  memset(out, 0, sizeof(float) * 4 * (size_t)W * H);
  /* pixel range of each quad's box, in pixels (not tiles); rows past H when it draws nothing */
  uint32_t *boxes = (uint32_t *)malloc(sizeof(uint32_t) * 4 * (size_t)(visible ? visible : 1));
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint32_t r = 0; r < visible; r++) {
    uint32_t *b = boxes + 4 * (size_t)r;
    const uint32_t *s = splats + 12 * (size_t)order[r];
    if (!or_tile_box(s, W, H, 1, b)) b[1] = b[3] = 0xFFFFFFFFu;
  }
  const uint32_t bands = (H + BAND - 1u) / BAND;
#pragma omp parallel for schedule(dynamic, 1)
  for (uint32_t band = 0; band < bands; band++) {
    const uint32_t y0 = band * BAND, y1 = (y0 + BAND < H ? y0 + BAND : H) - 1u;
    for (uint32_t r = 0; r < visible; r++) {        /* back to front: ascending key (shaders/splat.wgsl:52-55) */
      const uint32_t *box = boxes + 4 * (size_t)r;
      if (box[1] == 0xFFFFFFFFu || box[1] > y1 || box[3] < y0) continue;
      const uint32_t *s = splats + 12 * (size_t)order[r];
      float cx = bits_f(s[0]), cy = bits_f(s[1]);
      float a0x = bits_f(s[2]), a0y = bits_f(s[3]), a1x = bits_f(s[4]), a1y = bits_f(s[5]);
      float rr = bits_f(s[6]), tau0 = bits_f(s[7]);
      float rgb[3] = {bits_f(s[8]), bits_f(s[9]), bits_f(s[10])};
      float depth = bits_f(s[11]);
      if (view_mode == 1u) {
        /* depth view: rgb = 1 - clamp(t) (shaders/splat.wgsl:28,41,76) */
        float t = logf(depth / near_) / logf(far_ / near_);
        t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
        rgb[0] = rgb[1] = rgb[2] = 1.0f - t;
      }
      float det = a0x * a1y - a1x * a0y;
      if (det == 0.0f) continue;
      const uint32_t lo = box[1] > y0 ? box[1] : y0, hi = box[3] < y1 ? box[3] : y1;
      for (uint32_t py = (lo + row_step - 1u) / row_step * row_step; py <= hi; py += row_step) {
        float ny = 1.0f - ((float)py + 0.5f) * 2.0f / (float)H;
        for (uint32_t px = box[0]; px <= box[2]; px++) {
          float nx = ((float)px + 0.5f) * 2.0f / (float)W - 1.0f;
          float dx = nx - cx, dy = ny - cy;
          /* the pixel's corner coordinates on the quad c + u a0 + v a1 (the interpolated `corner`) */
          float u = (dx * a1y - a1x * dy) / det;
          float v = (a0x * dy - dx * a0y) / det;
          if (fabsf(u) > 1.0f || fabsf(v) > 1.0f) continue;
          float lx = u * rr, ly = v * rr;
          float e = expf_cr(-0.5f * (lx * lx + ly * ly));
          float alpha = 1.0f - expf_cr(-tau0 * e);
          if (alpha < ALPHA_MIN) continue;            /* shaders/splat.wgsl:73 (tau_c = 0: cov = 0) */
          float *d = out + 4 * ((size_t)py * W + px);
          if (view_mode == 2u) {                      /* overdraw: additive (1, 0, 0, 0) (src/renderer.js:141-143) */
            d[0] = d[0] + 1.0f;
            if (mode == 2) d[0] = to_half_rtne(d[0]);
            continue;
          }
          float src[4] = {rgb[0] * alpha, rgb[1] * alpha, rgb[2] * alpha, alpha};
          float keep = 1.0f - alpha;
          for (int k = 0; k < 4; k++) {
            d[k] = src[k] + d[k] * keep;               /* over: src * 1 + dst * (1 - src.a) */
            if (mode == 2) d[k] = to_half_rtne(d[k]);
          }
        }
      }
    }
  }
  free(boxes);
}

void or_composite_f64(const void *frame, const double *sp, const uint32_t *order, uint32_t visible,
                      uint32_t W, uint32_t H, double *out) {
  const double ALPHA_MIN = 1.0 / 255.0;
  uint32_t view_mode = uw(frame, FR_VIEW_MODE);
  double near_ = fw(frame, FR_NEAR), far_ = fw(frame, FR_FAR);
  memset(out, 0, sizeof(double) * 4 * (size_t)W * H);
  for (uint32_t r = 0; r < visible; r++) {
    const double *s = sp + 12 * (size_t)order[r];
    double cx = s[0], cy = s[1], a0x = s[2], a0y = s[3], a1x = s[4], a1y = s[5], rr = s[6], tau0 = s[7];
    double rgb[3] = {s[8], s[9], s[10]};
    if (view_mode == 1u) {
      double t = log(s[11] / near_) / log(far_ / near_);
      t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
      rgb[0] = rgb[1] = rgb[2] = 1.0 - t;
    }
    double ex = fabs(a0x) + fabs(a1x), ey = fabs(a0y) + fabs(a1y);
    double x0 = ceil((cx - ex + 1.0) * 0.5 * W - 0.5), x1 = floor((cx + ex + 1.0) * 0.5 * W - 0.5);
    double y0 = ceil((1.0 - (cy + ey)) * 0.5 * H - 0.5), y1 = floor((1.0 - (cy - ey)) * 0.5 * H - 0.5);
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > W - 1) x1 = W - 1;
    if (y1 > H - 1) y1 = H - 1;
    double det = a0x * a1y - a1x * a0y;
    if (det == 0.0) continue;
    for (double py = y0; py <= y1; py += 1.0) {
      double ny = 1.0 - (py + 0.5) * 2.0 / H;
      for (double px = x0; px <= x1; px += 1.0) {
        double nx = (px + 0.5) * 2.0 / W - 1.0;
        double dx = nx - cx, dy = ny - cy;
        double u = (dx * a1y - a1x * dy) / det, v = (a0x * dy - dx * a0y) / det;
        if (fabs(u) > 1.0 || fabs(v) > 1.0) continue;
        double lx = u * rr, ly = v * rr;
        double alpha = 1.0 - exp(-tau0 * exp(-0.5 * (lx * lx + ly * ly)));
        if (alpha < ALPHA_MIN) continue;
        double *d = out + 4 * ((size_t)py * W + (size_t)px);
        if (view_mode == 2u) { d[0] += 1.0; continue; }
        double keep = 1.0 - alpha;
        d[0] = rgb[0] * alpha + d[0] * keep;
        d[1] = rgb[1] * alpha + d[1] * keep;
        d[2] = rgb[2] * alpha + d[2] * keep;
        d[3] = alpha + d[3] * keep;
      }
    }
  }
}

/* The threads the oracle's parallel loops use (OMP_NUM_THREADS, else the host's cores). */
uint32_t or_threads(void) {
#ifdef _OPENMP
  return (uint32_t)omp_get_max_threads();
#else
  return 1u;
#endif
}
