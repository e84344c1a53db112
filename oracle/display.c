/* CPU oracle, display pass (TEST INFRASTRUCTURE ONLY): the resolve of the linear HDR target to the display, after
 * shaders/display.wgsl:22-189. Float32 in the shader's expression order, no contraction (-ffp-contract=off); the
 * shader's constant expressions (abstract floats in WGSL) are folded in double and rounded once to float, and its
 * exp, pow, log2, cos and sin (whose precision WGSL leaves to the implementation) are the correctly rounded ones.
 *
 * Pinned by tests/test_display_oracle.py: the sky recurrence against the 16 explicit basis functions of
 * sky_radiance and against scipy's complex spherical harmonics for bands 4-8, the lobes' peaks and integrals, the
 * heat ramp's named colours, sRGB's published 8-bit table points, and the tone curves' fixed points, limits and
 * invariants (grey stays grey, Reinhard's x / (1 + x), Hable's white point, PBR Neutral's identity band).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

/* exp, pow and log2 (and the horizon lobes' cos and sin) correctly rounded: the double function rounded once */
#define EXF(x) ((float)exp((double)(x)))
#define PWF(x, y) ((float)pow((double)(x), (double)(y)))
#define LG2F(x) ((float)log2((double)(x)))

typedef struct { float x, y, z; } v3;

static v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static float clampf_(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

/* display.wgsl:22-32: the lighting sky, 16 real SH coefficients (Env.sh) */
static v3 sky_radiance(const float *env, v3 w) {
  const float x = w.x, y = w.y, z = w.z;
  const float b[16] = {0.282095f, 0.488603f * y, 0.488603f * z, 0.488603f * x,
                       1.092548f * x * y, 1.092548f * y * z, 0.315392f * (3.0f * z * z - 1.0f), 1.092548f * x * z,
                       0.546274f * (x * x - y * y),
                       0.590044f * y * (3.0f * x * x - y * y), 2.890611f * x * y * z,
                       0.457046f * y * (5.0f * z * z - 1.0f), 0.373176f * z * (5.0f * z * z - 3.0f),
                       0.457046f * x * (5.0f * z * z - 1.0f), 1.445306f * z * (x * x - y * y),
                       0.590044f * x * (x * x - 3.0f * y * y)};
  v3 c = mk(0.0f, 0.0f, 0.0f);
  for (int k = 0; k < 16; k++) {
    c.x = c.x + b[k] * env[4 * k + 0];
    c.y = c.y + b[k] * env[4 * k + 1];
    c.z = c.z + b[k] * env[4 * k + 2];
  }
  return mk(fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f));
}

/* display.wgsl:36-80: the background sky (SkyBg): real SH by recurrence up to band 8, the glow SG, the horizon's
 * ASG ring and the ground below the horizon; the lighting sky when SkyBg holds no coefficients */
static v3 sky_background(const float *env, const float *bg, v3 w) {
  const uint32_t *bgu = (const uint32_t *)bg;
  uint32_t n = bgu[4 * 81];
  if (n > 81u) n = 81u;
  if (n == 0u) return sky_radiance(env, w);
  const uint32_t lmax = (uint32_t)roundf(sqrtf((float)n)) - 1u;
  v3 c = mk(0.0f, 0.0f, 0.0f);
  float cm = 1.0f, sm = 0.0f, pmm = 1.0f;
  for (uint32_t m = 0; m <= lmax; m++) {
    if (m > 0u) {
      const float c2 = cm * w.x - sm * w.y;
      sm = cm * w.y + sm * w.x;
      cm = c2;
      pmm *= (float)(2u * m - 1u);
    }
    float p_prev = 0.0f, p = pmm;
    float kf = (float)(1.0 / (4.0 * 3.14159265));
    for (uint32_t j = 1; j <= 2u * m; j++) kf /= (float)j;
    for (uint32_t l = m; l <= lmax; l++) {
      if (l == m + 1u) {
        p_prev = p;
        p = w.z * (float)(2u * m + 1u) * pmm;
      } else if (l > m + 1u) {
        const float pn = ((float)(2u * l - 1u) * w.z * p - (float)(l + m - 1u) * p_prev) / (float)(l - m);
        p_prev = p;
        p = pn;
        kf *= (float)(l - m) / (float)(l + m);
      }
      if (l == m + 1u) kf *= 1.0f / (float)(2u * m + 1u);
      const float k = sqrtf((float)(2u * l + 1u) * kf);
      const uint32_t i = l * l + l;
      if (m == 0u) {
        c.x = c.x + k * p * bg[4 * i + 0];
        c.y = c.y + k * p * bg[4 * i + 1];
        c.z = c.z + k * p * bg[4 * i + 2];
      } else {
        const float s = 1.41421356f * k * p;
        c.x = c.x + s * (cm * bg[4 * (i + m) + 0] + sm * bg[4 * (i - m) + 0]);
        c.y = c.y + s * (cm * bg[4 * (i + m) + 1] + sm * bg[4 * (i - m) + 1]);
        c.z = c.z + s * (cm * bg[4 * (i + m) + 2] + sm * bg[4 * (i - m) + 2]);
      }
    }
  }
  c = mk(fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f));
  const float *glow = bg + 4 * 93, *gdir = bg + 4 * 94;
  if (glow[3] > 0.0f) {
    const float s2 = glow[3] * glow[3];
    const float e = EXF((w.x * gdir[0] + w.y * gdir[1] + w.z * gdir[2] - 1.0f) / s2);
    const float d = (float)(2.0 * 3.14159265) * s2;
    c.x = c.x + glow[0] / d * e;
    c.y = c.y + glow[1] / d * e;
    c.z = c.z + glow[2] / d * e;
  }
  uint32_t nr = bgu[4 * 81 + 1];
  if (nr > 8u) nr = 8u;
  if (nr == 0u) return c;
  const float *ring = bg + 4 * 83, *ground = bg + 4 * 91, *par = bg + 4 * 92;
  for (uint32_t k = 0; k < nr; k++) {
    const float a = 6.28318531f * (float)k / (float)nr;
    const float zx = (float)cos((double)a), zy = (float)sin((double)a);
    const float along = w.x * -zy + w.y * zx;
    const float e = EXF(-par[0] * along * along - par[1] * w.z * w.z);
    const float f = fmaxf(w.x * zx + w.y * zy, 0.0f);
    c.x = c.x + ring[4 * k + 0] * f * e;
    c.y = c.y + ring[4 * k + 1] * f * e;
    c.z = c.z + ring[4 * k + 2] * f * e;
  }
  const float t = clampf_(w.z / ground[3] * 0.5f + 0.5f, 0.0f, 1.0f);
  return mk(ground[0] * (1.0f - t) + c.x * t, ground[1] * (1.0f - t) + c.y * t, ground[2] * (1.0f - t) + c.z * t);
}

/* display.wgsl:84-93: the distant lights' discs, an SG of each light's irradiance over its angular std (or the
 * visible disc's, SkyBg.disc_std, when that is positive) */
static v3 lights_radiance(const float *env, const float *bg, v3 w) {
  const uint32_t *envu = (const uint32_t *)env;
  v3 c = mk(0.0f, 0.0f, 0.0f);
  const uint32_t n = envu[96] < 4u ? envu[96] : 4u;
  for (uint32_t k = 0; k < n; k++) {
    const float *ds = env + 64 + 8 * k, *irr = env + 68 + 8 * k;
    const float dstd = bg[4 * 82 + k];
    const float sd = dstd > 0.0f ? dstd : ds[3];
    const float s2 = fmaxf(sd * sd, 1e-8f);
    const float e = EXF((w.x * ds[0] + w.y * ds[1] + w.z * ds[2] - 1.0f) / s2);
    const float d = (float)(2.0 * 3.14159265) * s2;
    c.x = c.x + irr[0] / d * e;
    c.y = c.y + irr[1] / d * e;
    c.z = c.z + irr[2] / d * e;
  }
  return c;
}

static v3 normalize3(v3 v) {
  const float l = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  return mk(v.x / l, v.y / l, v.z / l);
}

/* display.wgsl:95-99: the world direction through the fragment at (px, py) (the pixel centre: + 0.5) */
static v3 view_dir(const float *fr, float px, float py) {
  const v3 v = normalize3(mk((px - 0.5f * fr[72]) / fr[74], -(py - 0.5f * fr[73]) / fr[75], -1.0f));
  /* transpose(R) v with R's columns view[0..2].xyz: component i is column i . v */
  return normalize3(mk(fr[0] * v.x + fr[1] * v.y + fr[2] * v.z, fr[4] * v.x + fr[5] * v.y + fr[6] * v.z,
                       fr[8] * v.x + fr[9] * v.y + fr[10] * v.z));
}

static float srgb1(float c) {
  if (c <= 0.0031308f) return c * 12.92f;
  return 1.055f * PWF(c, (float)(1.0 / 2.4)) - 0.055f;
}

/* display.wgsl:113-117: black -> blue -> green -> yellow -> red */
static v3 heat(float t) {
  const float x = clampf_(t, 0.0f, 1.0f) * 4.0f;
  return mk(clampf_(x - 2.0f, 0.0f, 1.0f), clampf_(fminf(x - 1.0f, 4.0f - x), 0.0f, 1.0f),
            clampf_(fminf(x, 2.0f - x), 0.0f, 1.0f));
}

static float hable1(float x) {
  return (x * (0.15f * x + 0.05f) + 0.004f) / (x * (0.15f * x + 0.5f) + 0.06f) - (float)(0.02 / 0.3);
}

/* display.wgsl:161-172 */
static v3 pbr_neutral(v3 c) {
  const float start = (float)(0.8 - 0.04), desat = 0.15f;
  const float lo = fminf(c.x, fminf(c.y, c.z));
  const float off = lo < 0.08f ? lo - 6.25f * lo * lo : 0.04f;
  v3 x = mk(c.x - off, c.y - off, c.z - off);
  const float peak = fmaxf(x.x, fmaxf(x.y, x.z));
  if (peak < start) return mk(clampf_(x.x, 0.0f, 1.0f), clampf_(x.y, 0.0f, 1.0f), clampf_(x.z, 0.0f, 1.0f));
  const float d = 1.0f - start;
  const float new_peak = 1.0f - d * d / (peak + d - start);
  const float s = new_peak / peak;
  x = mk(x.x * s, x.y * s, x.z * s);
  const float g = 1.0f - 1.0f / (desat * (peak - new_peak) + 1.0f);
  return mk(clampf_(x.x * (1.0f - g) + new_peak * g, 0.0f, 1.0f), clampf_(x.y * (1.0f - g) + new_peak * g, 0.0f, 1.0f),
            clampf_(x.z * (1.0f - g) + new_peak * g, 0.0f, 1.0f));
}

/* display.wgsl:176-189: inset (columns), log2 encode over [-12.47393, 4.026069], the sigmoid polynomial, outset,
 * 2.2 power */
static const float AGX_M[3][3] = {{0.842479062253094f, 0.0423282422610123f, 0.0423756549057051f},
                                  {0.0784335999999992f, 0.878468636469772f, 0.0784336f},
                                  {0.0792237451477643f, 0.0791661274605434f, 0.879142973793104f}};
static const float AGX_MI[3][3] = {{1.19687900512017f, -0.0528968517574562f, -0.0529716355144438f},
                                   {-0.0980208811401368f, 1.15190312990417f, -0.0980434501171241f},
                                   {-0.0990297440797205f, -0.0989611768448433f, 1.15107367264116f}};

static v3 mat_col(const float m[3][3], v3 v) {
  return mk(m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z, m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z,
            m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z);
}

static float agx_curve(float xin) {
  const float lo = -12.47393f, hi = 4.026069f;
  float x = clampf_(LG2F(fmaxf(xin, 1e-10f)), lo, hi);
  x = (x - lo) / (hi - lo);
  const float x2 = x * x, x4 = x2 * x2;
  return 15.5f * x4 * x2 - 40.14f * x4 * x + 31.96f * x4 - 6.868f * x2 * x + 0.4298f * x2 + 0.1191f * x - 0.00232f;
}

static v3 agx(v3 v) {
  const v3 a = mat_col(AGX_M, v);
  const v3 y = mat_col(AGX_MI, mk(agx_curve(a.x), agx_curve(a.y), agx_curve(a.z)));
  return mk(clampf_(PWF(fmaxf(y.x, 0.0f), 2.2f), 0.0f, 1.0f), clampf_(PWF(fmaxf(y.y, 0.0f), 2.2f), 0.0f, 1.0f),
            clampf_(PWF(fmaxf(y.z, 0.0f), 2.2f), 0.0f, 1.0f));
}

static float aces1(float x) {
  return clampf_(x * (2.51f * x + 0.03f) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f, 1.0f);
}

/* display.wgsl:140-152 */
static v3 tone_map(v3 x, uint32_t mode) {
  if (mode == 1u) return agx(x);
  if (mode == 2u) return mk(aces1(x.x), aces1(x.y), aces1(x.z));
  if (mode == 3u) {
    const float l = x.x * 0.2126f + x.y * 0.7152f + x.z * 0.0722f;
    return mk(clampf_(x.x / (1.0f + l), 0.0f, 1.0f), clampf_(x.y / (1.0f + l), 0.0f, 1.0f),
              clampf_(x.z / (1.0f + l), 0.0f, 1.0f));
  }
  if (mode == 4u) return pbr_neutral(x);
  if (mode == 5u) {
    const float wt = hable1(11.2f);
    return mk(clampf_(hable1(2.0f * x.x) / wt, 0.0f, 1.0f), clampf_(hable1(2.0f * x.y) / wt, 0.0f, 1.0f),
              clampf_(hable1(2.0f * x.z) / wt, 0.0f, 1.0f));
  }
  return mk(clampf_(x.x, 0.0f, 1.0f), clampf_(x.y, 0.0f, 1.0f), clampf_(x.z, 0.0f, 1.0f));
}

static uint8_t unorm8(float v) { return (uint8_t)floorf(clampf_(v, 0.0f, 1.0f) * 255.0f + 0.5f); }

static const float ZERO_ENV[104], ZERO_BG[380];

void or_display(const void *frame, const float *hdr, uint32_t W, uint32_t H, const void *env_, const void *bg_,
                float *out, uint8_t *bytes) {
  const float *fr = (const float *)frame;
  const uint32_t *fu = (const uint32_t *)frame;
  const float *env = env_ ? (const float *)env_ : ZERO_ENV;
  const float *bg = bg_ ? (const float *)bg_ : ZERO_BG;
  const uint32_t *envu = (const uint32_t *)env;
  const uint32_t view_mode = fu[79], tone = fu[260];
  const float exposure = fr[82];
  const size_t npx = (size_t)W * H;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < npx; i++) {
    const float *c = hdr + 4 * i;
    v3 o;
    if (view_mode == 2u) {
      o = heat(LG2F(1.0f + c[0]) / 7.0f);
    } else {
      v3 rgb = mk(c[0], c[1], c[2]);
      if (view_mode == 3u && (envu[96] > 0u || envu[97] != 0u)) {
        const float bare = 1.0f - clampf_(c[3], 0.0f, 1.0f);
        if (bare > 1.0f / 1024.0f) {
          const v3 w = view_dir(fr, (float)(i % W) + 0.5f, (float)(i / W) + 0.5f);
          const v3 s = envu[97] != 0u ? sky_background(env, bg, w) : mk(0.0f, 0.0f, 0.0f);
          const v3 l = lights_radiance(env, bg, w);
          rgb = mk(rgb.x + bare * (s.x + l.x), rgb.y + bare * (s.y + l.y), rgb.z + bare * (s.z + l.z));
        }
      }
      const v3 t = tone_map(mk(fmaxf(rgb.x * exposure, 0.0f), fmaxf(rgb.y * exposure, 0.0f),
                               fmaxf(rgb.z * exposure, 0.0f)), tone);
      o = mk(srgb1(t.x), srgb1(t.y), srgb1(t.z));
    }
    if (out) {
      out[4 * i + 0] = o.x;
      out[4 * i + 1] = o.y;
      out[4 * i + 2] = o.z;
      out[4 * i + 3] = 1.0f;
    }
    if (bytes) {
      bytes[4 * i + 0] = unorm8(o.x);
      bytes[4 * i + 1] = unorm8(o.y);
      bytes[4 * i + 2] = unorm8(o.z);
      bytes[4 * i + 3] = 255u;
    }
  }
}

/* pieces for the pins: n directions / values at a time */
void or_tone_map_n(const float *x, uint32_t n, uint32_t mode, float *y) {
  for (uint32_t i = 0; i < n; i++) {
    const v3 t = tone_map(mk(x[3 * i], x[3 * i + 1], x[3 * i + 2]), mode);
    y[3 * i] = t.x;
    y[3 * i + 1] = t.y;
    y[3 * i + 2] = t.z;
  }
}

void or_srgb_encode_n(const float *x, uint32_t n, float *y) {
  for (uint32_t i = 0; i < n; i++) y[i] = srgb1(x[i]);
}

void or_heat_n(const float *t, uint32_t n, float *y) {
  for (uint32_t i = 0; i < n; i++) {
    const v3 h = heat(t[i]);
    y[3 * i] = h.x;
    y[3 * i + 1] = h.y;
    y[3 * i + 2] = h.z;
  }
}

/* which: 0 sky_radiance (Env), 1 sky_background (SkyBg, else Env), 2 lights_radiance */
void or_sky_n(const void *env_, const void *bg_, const float *w, uint32_t n, uint32_t which, float *y) {
  const float *env = env_ ? (const float *)env_ : ZERO_ENV;
  const float *bg = bg_ ? (const float *)bg_ : ZERO_BG;
  for (uint32_t i = 0; i < n; i++) {
    const v3 d = mk(w[3 * i], w[3 * i + 1], w[3 * i + 2]);
    const v3 c = which == 0u ? sky_radiance(env, d) : which == 1u ? sky_background(env, bg, d)
                                                                   : lights_radiance(env, bg, d);
    y[3 * i] = c.x;
    y[3 * i + 1] = c.y;
    y[3 * i + 2] = c.z;
  }
}

void or_view_dir_n(const void *frame, const float *pxy, uint32_t n, float *w) {
  for (uint32_t i = 0; i < n; i++) {
    const v3 d = view_dir((const float *)frame, pxy[2 * i], pxy[2 * i + 1]);
    w[3 * i] = d.x;
    w[3 * i + 1] = d.y;
    w[3 * i + 2] = d.z;
  }
}
