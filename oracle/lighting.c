/* CPU oracle for the relighting of the splat frame path: TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Float32 restatement of runtime/wgsl/{common,iso_init,emissive,light_shadow,sun_map,environment}.wgsl for the
 * default static M39 matte path, evaluated in WGSL expression order without contraction. The opacity-map splats
 * use the correctly rounded float exp (double exp rounded once) so their fixed-point accumulators are integers
 * that any implementation using the same definition reproduces exactly.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z; } v3;
typedef struct { float x, y; } v2;
typedef struct { v2 c0, c1; } m2;          /* column-major 2x2 */
typedef struct { v3 c0, c1, c2; } m3;      /* column-major 3x3 */

static const float PI_F = 3.14159265358979f;
static const uint32_t NO_CHILD = 0xFFFFFFFFu;

static v3 V3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static v2 V2(float x, float y) { v2 r = {x, y}; return r; }
static v3 add3(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub3(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 mul3s(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static float dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static float dot2(v2 a, v2 b) { return a.x * b.x + a.y * b.y; }
static float len3(v3 a) { return sqrtf(dot3(a, a)); }
static v3 norm3(v3 a) { float l = len3(a); return V3(a.x / l, a.y / l, a.z / l); }
static v3 cross3(v3 a, v3 b) { return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static float clampf_(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
static float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
/* e^x for the opacity-map splats (as the device defines it): Cody-Waite reduction by ln 2 split in two, fused,
   and the degree-7 Taylor polynomial on [-ln2/2, ln2/2] in Horner form with explicit fmaf; x clamped to
   [-86, 88]. Every operation is a correctly rounded IEEE one, so host and device agree bit for bit. */
static float exp_map(float x) {
  if (!(x == x)) return x;
  x = fminf(fmaxf(x, -86.0f), 88.0f);
  const float k = rintf(x * 1.44269504f);
  float r = fmaf(k, -0.693145751953125f, x);
  r = fmaf(k, -1.42860677e-6f, r);
  float p = 1.98412698e-4f;
  p = fmaf(p, r, 1.38888889e-3f);
  p = fmaf(p, r, 8.33333333e-3f);
  p = fmaf(p, r, 4.16666667e-2f);
  p = fmaf(p, r, 1.66666667e-1f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}
void or_exp_map_n(const float *x, float *y, uint32_t n) {
  for (uint32_t i = 0; i < n; i++) y[i] = exp_map(x[i]);
}
/* picks of a cell, mip level or cull decision: double, rounded once (as the device does) */
static float log2_cr(float x) { return (float)log2((double)x); }
static float asin_cr(float x) { return (float)asin((double)x); }
static float acos_cr(float x) { return (float)acos((double)x); }
static uint32_t u32_sat(float f) { /* WGSL u32(f32): truncate, saturate */
  if (!(f > 0.0f)) return 0u;
  if (f >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)f;
}
static float rdf(const void *p, size_t word) { float v; memcpy(&v, (const char *)p + 4 * word, 4); return v; }
static uint32_t rdu(const void *p, size_t word) { uint32_t v; memcpy(&v, (const char *)p + 4 * word, 4); return v; }
static v3 rd3(const void *p, size_t word) { return V3(rdf(p, word), rdf(p, word + 1), rdf(p, word + 2)); }

/* m * v, column-major: m[0] v.x + m[1] v.y + m[2] v.z */
static v3 m3v(m3 m, v3 v) { return add3(add3(mul3s(m.c0, v.x), mul3s(m.c1, v.y)), mul3s(m.c2, v.z)); }
static m3 m3mul(m3 a, m3 b) { m3 r = {m3v(a, b.c0), m3v(a, b.c1), m3v(a, b.c2)}; return r; }
static m3 m3t(m3 m) {
  m3 r = {V3(m.c0.x, m.c1.x, m.c2.x), V3(m.c0.y, m.c1.y, m.c2.y), V3(m.c0.z, m.c1.z, m.c2.z)};
  return r;
}
static v2 m2v(m2 m, v2 v) { return V2(m.c0.x * v.x + m.c1.x * v.y, m.c0.y * v.x + m.c1.y * v.y); }
static m2 m2mul(m2 a, m2 b) { m2 r = {m2v(a, b.c0), m2v(a, b.c1)}; return r; }
static m2 m2t(m2 m) { m2 r = {V2(m.c0.x, m.c1.x), V2(m.c0.y, m.c1.y)}; return r; }
static m2 m2s(m2 m, float s) { m2 r = {V2(m.c0.x * s, m.c0.y * s), V2(m.c1.x * s, m.c1.y * s)}; return r; }
static float m2det(m2 m) { return m.c0.x * m.c1.y - m.c1.x * m.c0.y; }
static m2 inv2(m2 m) { /* light_shadow.wgsl:112-115 */
  float d = m2det(m);
  m2 r = {V2(m.c1.y, -m.c0.y), V2(-m.c1.x, m.c0.x)};
  return m2s(r, 1.0f / d);
}

/* quat_matrix (runtime/wgsl/common.wgsl:218-223) */
static m3 quat_matrix(float x, float y, float z, float w) {
  m3 r = {V3(1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + z * w), 2.0f * (x * z - y * w)),
          V3(2.0f * (x * y - z * w), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + x * w)),
          V3(2.0f * (x * z + y * w), 2.0f * (y * z - x * w), 1.0f - 2.0f * (x * x + y * y))};
  return r;
}

/* part transforms (runtime/wgsl/common.wgsl:79-94); xf: 12 words [R row | t] */
static v3 xf_point(const void *xf, v3 p) {
  return V3(dot3(rd3(xf, 0), p) + rdf(xf, 3), dot3(rd3(xf, 4), p) + rdf(xf, 7), dot3(rd3(xf, 8), p) + rdf(xf, 11));
}
static v3 xf_dir_inv(const void *xf, v3 d) {
  return add3(add3(mul3s(rd3(xf, 0), d.x), mul3s(rd3(xf, 4), d.y)), mul3s(rd3(xf, 8), d.z));
}
static m3 part_rot(const void *xf) { m3 r = {rd3(xf, 0), rd3(xf, 4), rd3(xf, 8)}; return m3t(r); }

/* erfc for z >= 0 (runtime/wgsl/common.wgsl:113-118, Numerical Recipes erfcc) */
float or_erfc_pos(float z) {
  float t = 1.0f / (1.0f + 0.5f * z);
  float p = -z * z - 1.26551223f + t * (1.00002368f + t * (0.37409196f + t * (0.09678418f + t * (-0.18628806f
            + t * (0.27886807f + t * (-1.13520398f + t * (1.48851587f + t * (-0.82215223f + t * 0.17087277f))))))));
  return t * expf(p);
}

static v3 sym_mul(v3 dg, v3 off, v3 v) {
  return V3((dg.x * v.x + off.x * v.y) + off.y * v.z, (off.x * v.x + dg.y * v.y) + off.z * v.z,
            (off.y * v.x + off.z * v.y) + dg.z * v.z);
}

/* Node words: mean 0, radius 3, cinv_diag 4, kappa 7, cinv_off 8, smax 11, sn_diag 12, part 15, sn_off 16, pad0 22 */
static float dirm(const void *nd, v3 w) {   /* runtime/wgsl/common.wgsl:127-129 */
  float q = dot3(w, sym_mul(rd3(nd, 12), rd3(nd, 16), w));
  return sqrtf(q > 0.0f ? q : 0.0f);
}

/* Ein(x) (runtime/wgsl/common.wgsl:226-239): series below 1, else Abramowitz-Stegun 5.1.56 */
float or_ein(float x) {
  if (x < 1.0f) {
    float term = x, s = x;
    for (int k = 2; k <= 12; k++) {
      term *= -x / (float)k;
      s += term / (float)k;
    }
    return s;
  }
  float p = x * x * x * x + 8.5733287401f * x * x * x + 18.0590169730f * x * x + 8.6347608925f * x + 0.2677737343f;
  float q = x * x * x * x + 9.5733223454f * x * x * x + 25.6329561486f * x * x + 21.0996530827f * x + 3.9584969228f;
  return logf(x) + 0.5772156649f + p / q / (x * expf(x));
}

/* footprint A(w) (runtime/wgsl/common.wgsl:242-246) */
float or_footprint(const void *nd, float detC, float wx, float wy, float wz) {
  v3 w = V3(wx, wy, wz);
  float a = dot3(w, sym_mul(rd3(nd, 4), rd3(nd, 8), w));
  float K = rdf(nd, 7) * dirm(nd, w) * sqrtf(2.0f * PI_F / a);
  return 2.0f * PI_F * sqrtf(detC * a) * or_ein(K);
}

static float vgrid_cell(const uint8_t *v, uint32_t i, uint32_t j) { return (float)v[i * 16u + j] * (1.0f / 255.0f); }

/* bilinear on the 8x16 cell centres (runtime/wgsl/common.wgsl:162-177) */
float or_vgrid_lookup(const uint8_t *v, float wx, float wy, float wz) {
  float el = asin_cr(clampf_(wz, -1.0f, 1.0f));
  float fi = clampf_((0.5f * PI_F - el) / (PI_F / 8.0f) - 0.5f, 0.0f, 7.0f);
  uint32_t i0 = (uint32_t)fi < 6u ? (uint32_t)fi : 6u;
  float ti = fi - (float)i0;
  float az = (fabsf(wx) + fabsf(wy) < 1e-30f) ? 0.0f : atan2f(wy, wx);
  float fj = (az + PI_F) / (2.0f * PI_F) * 16.0f - 0.5f;
  float j0f = floorf(fj);
  float tj = fj - j0f;
  uint32_t j0 = (uint32_t)((int32_t)j0f + 16) % 16u;
  uint32_t j1 = (j0 + 1u) % 16u;
  float a = mixf(vgrid_cell(v, i0, j0), vgrid_cell(v, i0, j1), tj);
  float b = mixf(vgrid_cell(v, i0 + 1u, j0), vgrid_cell(v, i0 + 1u, j1), tj);
  return mixf(a, b, ti);
}

static v3 basis_k(v3 w) {   /* runtime/wgsl/common.wgsl:181-185 */
  v3 a = V3(0.0f, 0.0f, 1.0f);
  if (fabsf(w.z) >= 0.999f) a = V3(1.0f, 0.0f, 0.0f);
  return norm3(cross3(a, w));
}

/* ---- iso_init (runtime/wgsl/iso_init.wgsl:29-56) ------------------------------------------------------------- */

void or_iso_init(const void *gaussians, const void *nodes, const uint8_t *vgrid, uint32_t n, float *iso_norm) {
#pragma omp parallel for schedule(dynamic, 256)
  for (uint32_t g = 0; g < n; g++) {
    const char *G = (const char *)gaussians + 64 * (size_t)g;
    const char *nd = (const char *)nodes + 96 * (size_t)g;
    float s3 = rdf(G, 8) * rdf(G, 9) * rdf(G, 10);
    float S = 0.0f, S0 = 0.0f;
    for (uint32_t c = 0; c < 128u; c++) {
      uint32_t i = c / 16u, j = c % 16u;
      float el = 0.5f * PI_F - ((float)i + 0.5f) * PI_F / 8.0f;
      float az = -PI_F + ((float)j + 0.5f) * 2.0f * PI_F / 16.0f;
      float w[3] = {cosf(el) * cosf(az), cosf(el) * sinf(az), sinf(el)};
      float top = 0.5f * PI_F - (float)i * PI_F / 8.0f;
      float om = (2.0f * PI_F / 16.0f) * (sinf(top) - sinf(top - PI_F / 8.0f));
      float v = (float)vgrid[128 * (size_t)g + c] * (1.0f / 255.0f);
      float a = or_footprint(nd, s3 * s3, w[0], w[1], w[2]) * om;
      S += a * v;
      S0 += a;
    }
    float m = S > 0.02f * S0 ? S : 0.02f * S0;
    iso_norm[g] = 1.0f / (m > 1e-30f ? m : 1e-30f);
  }
}

/* ---- shade_emit (runtime/wgsl/emissive.wgsl:123-150, 853-980), M39 matte ------------------------------------- */

/* EmitParams words: count 0, root 1, scale 2, theta 3, mode 4, node_count 5, skip 6, one_sided 7, normal 8,
 * n5 11, edge_soft 12. EmitNode words (24): mean 0, radius 3, m_diag 4, tag 7, m_off 8, child_lo 11, Le 12,
 * child_hi 15. FrameUniforms: gaussian_count 1, eye 8, material_count 13, iso 15. */
static float smoothstep_(float e0, float e1, float x) {
  float t = clampf_((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

static v3 intensity(const void *ep, const void *e, v3 w) {
  float s = 1.0f;
  if (rdu(ep, 7) != 0u || (rdu(e, 7) & 2u) != 0u) {
    float c = dot3(w, rd3(ep, 8));
    float es = rdf(ep, 12);
    s = es > 0.0f ? smoothstep_(-es, es, c) : (c > 0.0f ? 1.0f : 0.0f);
  }
  float q = dot3(w, sym_mul(rd3(e, 4), rd3(e, 8), w));
  float k = s * rdf(ep, 2) * sqrtf(q > 0.0f ? q : 0.0f);
  return mul3s(rd3(e, 12), k);
}

static float unpack_half(uint32_t h) {
  uint32_t s = (h >> 15) & 1u, e = (h >> 10) & 31u, m = h & 1023u;
  float v = e == 0 ? ldexpf((float)m, -24) : (e == 31 ? INFINITY : ldexpf((float)(m | 1024u), (int)e - 25));
  return s ? -v : v;
}

void or_shade_emit(const void *U, const void *EP, const void *emit_nodes, const void *gaussians, const void *nodes,
                   const uint8_t *vgrid, const void *parts, const float *lamp_T, const void *materials,
                   const float *iso_norm, float *colour) {
  uint32_t n = rdu(U, 1), mcount = rdu(U, 13);
  v3 eye = rd3(U, 8);
  float theta = rdf(EP, 3);
  uint32_t mode = rdu(EP, 4);
#pragma omp parallel for schedule(dynamic, 256)
  for (uint32_t g = 0; g < n; g++) {
    const char *G = (const char *)gaussians + 64 * (size_t)g;
    const char *nd = (const char *)nodes + 96 * (size_t)g;
    uint32_t mi = rdu(G, 15) & 0xFFFFu;
    uint32_t mmax = (mcount > 1u ? mcount : 1u) - 1u;
    const char *mat = (const char *)materials + 32 * (size_t)(mi < mmax ? mi : mmax);
    uint32_t mflags = mcount ? rdu(mat, 5) : 0u;
    if (rdu(EP, 0) == 0u && (mflags & 1u) == 0u) {
      colour[4 * (size_t)g + 0] = colour[4 * (size_t)g + 1] = colour[4 * (size_t)g + 2] = 0.0f;
      colour[4 * (size_t)g + 3] = 1.0f;
      continue;
    }
    const char *xf = (const char *)parts + 128 * (size_t)(rdu(G, 11) & 0xFFFFu);
    v3 P = xf_point(xf, rd3(G, 0));
    m3 Rq = quat_matrix(rdf(G, 4), rdf(G, 5), rdf(G, 6), rdf(G, 7));
    float s3 = rdf(G, 8) * rdf(G, 9) * rdf(G, 10);
    float detC = s3 * s3;
    v3 wo = norm3(sub3(eye, P));
    const uint8_t *vg = vgrid + 128 * (size_t)g;
    v3 acc = V3(0.0f, 0.0f, 0.0f);
    uint32_t stack[48];
    stack[0] = rdu(EP, 1);
    uint32_t sp = 1;
    while (sp > 0) {
      sp -= 1;
      uint32_t ni = stack[sp];
      const char *e = (const char *)emit_nodes + 96 * (size_t)ni;
      uint32_t clo = rdu(e, 11);
      int use;
      if (clo == NO_CHILD || mode == 0u) use = clo == NO_CHILD;
      else {
        v3 d = sub3(P, rd3(e, 0));
        float r = rdf(e, 3);
        use = dot3(d, d) > theta * theta * r * r;
      }
      if (!use) {
        if (sp + 2u <= 48u) { stack[sp] = clo; stack[sp + 1] = rdu(e, 15); sp += 2; }
        continue;
      }
      v3 d = sub3(rd3(e, 0), P);
      float d2 = dot3(d, d);
      if (d2 < 1e-8f) continue;
      v3 ww = V3(d.x / sqrtf(d2), d.y / sqrtf(d2), d.z / sqrtf(d2));
      v3 wi = xf_dir_inv(xf, ww);
      float vi = or_vgrid_lookup(vg, wi.x, wi.y, wi.z);
      if (vi <= 1e-7f) continue;
      float fp = or_footprint(nd, detC, wi.x, wi.y, wi.z);
      v3 I = intensity(EP, e, V3(-ww.x, -ww.y, -ww.z));
      float k = fp / d2 * 1.0f;
      v3 light = V3(I.x * vi * k, I.y * vi * k, I.z * vi * k);
      acc = add3(acc, light);
    }
    float T = lamp_T ? lamp_T[g] : 1.0f;
    float inorm = iso_norm[g];
    v3 alb = rd3(G, 12);
    float tk = T * inorm;
    v3 c = V3(alb.x * acc.x * tk, alb.y * acc.y * tk, alb.z * acc.z * tk);
    if ((mflags & 1u) != 0u && ((mflags & 4u) == 0u || dot3(xf_dir_inv(xf, wo), Rq.c0) > 0.0f)) {
      uint32_t p0 = rdu(mat, 6), p1 = rdu(mat, 7);
      c = add3(c, V3(unpack_half(p0 & 0xFFFFu), unpack_half(p0 >> 16), unpack_half(p1 & 0xFFFFu)));
    }
    colour[4 * (size_t)g + 0] = c.x;
    colour[4 * (size_t)g + 1] = c.y;
    colour[4 * (size_t)g + 2] = c.z;
    colour[4 * (size_t)g + 3] = 1.0f;
  }
}

/* ---- opacity maps: shared pieces (light_shadow.wgsl:272-392, sun_map.wgsl:228-324) ------------------------------ */

static float phi_(float x) {
  float a = fabsf(x) * 0.70710678f;
  float h = 0.5f * or_erfc_pos(a);
  return x >= 0.0f ? 1.0f - h : h;
}

static void resolve_layer(const uint32_t *acc, float fix, uint32_t n, float *mip) {
  for (uint32_t i = 0; i < n; i++) {
    float s0 = (float)acc[3 * (size_t)i] / fix;
    float o[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (s0 > 1e-5f) {
      float a = (float)(-expm1(-(double)s0));   /* 1 - e^-s0 rounded once (the f32 form cancels for small s0) */
      o[0] = a;
      o[1] = a * (float)acc[3 * (size_t)i + 1] / (fix * s0);
      o[2] = a * (float)acc[3 * (size_t)i + 2] / (fix * s0);
    }
    memcpy(mip + 4 * (size_t)i, o, 16);
  }
}

static void down_layer(float *mip, const uint32_t *level_off, uint32_t levels, uint32_t res) {
  for (uint32_t L = 1; L < levels; L++) {
    uint32_t rk = res >> L, pr = rk * 2u;
    for (uint32_t y = 0; y < rk; y++)
      for (uint32_t x = 0; x < rk; x++) {
        size_t src = level_off[L - 1] + (2u * y) * pr + 2u * x;
        float *d = mip + 4 * ((size_t)level_off[L] + y * rk + x);
        for (int c = 0; c < 4; c++)
          d[c] = 0.25f * (((mip[4 * src + c] + mip[4 * (src + 1) + c]) + mip[4 * (src + pr) + c]) + mip[4 * (src + pr + 1) + c]);
      }
  }
}

static void mip_at(const float *mip, const uint32_t *level_off, uint32_t res, uint32_t k, v2 tc, float out[4]) {
  uint32_t rk = (res >> k) > 1u ? (res >> k) : 1u;
  float sc = (float)(1u << k);
  float cx = clampf_(tc.x / sc - 0.5f, 0.0f, (float)(rk - 1u)), cy = clampf_(tc.y / sc - 0.5f, 0.0f, (float)(rk - 1u));
  uint32_t i0x = (uint32_t)floorf(cx), i0y = (uint32_t)floorf(cy);
  uint32_t i1x = i0x + 1u < rk - 1u ? i0x + 1u : rk - 1u, i1y = i0y + 1u < rk - 1u ? i0y + 1u : rk - 1u;
  float fx = cx - (float)i0x, fy = cy - (float)i0y;
  const float *b = mip + 4 * (size_t)level_off[k];
  for (int c = 0; c < 4; c++) {
    float a = mixf(b[4 * (i0y * rk + i0x) + c], b[4 * (i0y * rk + i1x) + c], fx);
    float bb = mixf(b[4 * (i1y * rk + i0x) + c], b[4 * (i1y * rk + i1x) + c], fx);
    out[c] = mixf(a, bb, fy);
  }
}

static void mip_tri(const float *mip, const uint32_t *level_off, uint32_t res, uint32_t levels, float l, v2 tc,
                    float out[4]) {
  float lc = clampf_(l, 0.0f, (float)(levels - 1u));
  uint32_t k0 = (uint32_t)floorf(lc);
  uint32_t k1 = k0 + 1u < levels - 1u ? k0 + 1u : levels - 1u;
  float a[4], b[4];
  mip_at(mip, level_off, res, k0, tc, a);
  mip_at(mip, level_off, res, k1, tc, b);
  float t = lc - (float)k0;
  for (int c = 0; c < 4; c++) out[c] = mixf(a[c], b[c], t);
}

static m3 leaf_cov(const char *G, m3 R) {
  m3 Rq = quat_matrix(rdf(G, 4), rdf(G, 5), rdf(G, 6), rdf(G, 7));
  float s2x = rdf(G, 8) * rdf(G, 8), s2y = rdf(G, 9) * rdf(G, 9), s2z = rdf(G, 10) * rdf(G, 10);
  m3 S = {V3(s2x, 0.0f, 0.0f), V3(0.0f, s2y, 0.0f), V3(0.0f, 0.0f, s2z)};
  m3 C = m3mul(m3mul(Rq, S), m3t(Rq));
  return m3mul(m3mul(R, C), m3t(R));
}

/* ---- A: emitter light opacity map (light_shadow.wgsl:64-107, 191-392) ------------------------------------------- */

/* LSParams words: center 0, res 3, normal 4, levels 7, t1 8, layer_texels 11, t2 12, dirty 15, sigma 16,
 * skip 17, hemi 18, part_count 19, level_off 36.. */
static v2 oct_uv(const void *LS, v3 w) {
  v3 t1 = rd3(LS, 8), t2 = rd3(LS, 12), nn = rd3(LS, 4);
  v3 l = V3(dot3(w, t1), dot3(w, t2), dot3(w, nn));
  if (rdu(LS, 18) != 0u) {
    float z = l.z > 0.0f ? l.z : 0.0f;
    float den = (fabsf(l.x) + fabsf(l.y)) + z;
    den = den > 1e-12f ? den : 1e-12f;
    v2 p = V2(l.x / den, l.y / den);
    return V2(p.x + p.y, p.x - p.y);
  }
  float den = (fabsf(l.x) + fabsf(l.y)) + fabsf(l.z);
  den = den > 1e-12f ? den : 1e-12f;
  v2 p = V2(l.x / den, l.y / den);
  if (l.z < 0.0f) {
    v2 q = V2((1.0f - fabsf(p.y)) * (p.x >= 0.0f ? 1.0f : -1.0f), (1.0f - fabsf(p.x)) * (p.y >= 0.0f ? 1.0f : -1.0f));
    p = q;
  }
  return p;
}

static m2 oct_jac(const void *LS, v3 w, v3 b0, v3 b1) {
  const float e = 1e-3f;
  v2 u0 = oct_uv(LS, w);
  v2 ua = oct_uv(LS, norm3(add3(w, mul3s(b0, e))));
  v2 ub = oct_uv(LS, norm3(add3(w, mul3s(b1, e))));
  m2 r = {V2((ua.x - u0.x) / e, (ua.y - u0.y) / e), V2((ub.x - u0.x) / e, (ub.y - u0.y) / e)};
  return r;
}

void or_lsm_splat(const void *U, const void *LS, const void *gaussians, const void *nodes, const void *parts,
                  uint32_t *acc) {
  const float FIX = 65536.0f, LOWPASS = 0.3f;
  const int MAX_HALF = 48;
  uint32_t n = rdu(U, 1), res = rdu(LS, 3), pc = rdu(LS, 19), skip = rdu(LS, 17), dirty = rdu(LS, 15);
  v3 center = rd3(LS, 0), nrm = rd3(LS, 4);
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint32_t g = 0; g < n; g++) {
    const char *G = (const char *)gaussians + 64 * (size_t)g;
    uint32_t p = rdu(G, 11) & 0xFFFFu;
    if (p >= pc || p == skip || (dirty & (1u << p)) == 0u) continue;
    const char *nd = (const char *)nodes + 96 * (size_t)g;
    const char *xf = (const char *)parts + 128 * (size_t)p;
    m3 R = part_rot(xf);
    v3 m = xf_point(xf, rd3(G, 0));
    v3 dv = sub3(m, center);
    float r = len3(dv);
    if (r < 1e-4f) continue;
    v3 w = V3(dv.x / r, dv.y / r, dv.z / r);
    m3 C = leaf_cov(G, R);
    if (rdu(LS, 18) != 0u && dot3(dv, nrm) < -3.0f * sqrtf(dot3(nrm, m3v(C, nrm)))) continue;
    v3 b0 = basis_k(w), b1 = cross3(w, b0);
    v3 c0 = m3v(C, b0), c1 = m3v(C, b1);
    v3 cw = m3v(C, w);
    m2 Cbb = {V2(dot3(b0, c0), dot3(b0, c1)), V2(dot3(b0, c1), dot3(b1, c1))};
    v2 cbz = V2(dot3(b0, cw), dot3(b1, cw));
    float czz = dot3(w, cw);
    m2 Cbb_i = inv2(Cbb);
    float sz2 = czz - dot2(cbz, m2v(Cbb_i, cbz));
    sz2 = sz2 > 1e-12f ? sz2 : 1e-12f;
    float tau0 = rdf(nd, 7) * dirm(nd, xf_dir_inv(xf, w)) * sqrtf(2.0f * PI_F * sz2);
    float half_res = 0.5f * (float)res;
    m2 J = m2s(oct_jac(LS, w, b0, b1), half_res);
    m2 Sang = m2s(Cbb, 1.0f / (r * r));
    m2 S = m2mul(m2mul(J, Sang), m2t(J));
    float det0 = m2det(S);
    S.c0.x += LOWPASS;
    S.c1.y += LOWPASS;
    float det1 = m2det(S);
    tau0 *= sqrtf((det0 > 0.0f ? det0 : 0.0f) / det1);
    m2 Si = inv2(S);
    v2 cz = m2v(Cbb_i, cbz);
    v2 gz = m2v(m2t(inv2(J)), V2(cz.x * r, cz.y * r));
    v2 uv = oct_uv(LS, w);
    v2 ct = V2((uv.x * 0.5f + 0.5f) * (float)res, (uv.y * 0.5f + 0.5f) * (float)res);
    int hx = (int)ceilf(3.0f * sqrtf(S.c0.x)), hy = (int)ceilf(3.0f * sqrtf(S.c1.y));
    hx = hx < MAX_HALF ? hx : MAX_HALF;
    hy = hy < MAX_HALF ? hy : MAX_HALF;
    int cx = (int)floorf(ct.x), cy = (int)floorf(ct.y);
    size_t base = (size_t)p * res * res;
    int ires = (int)res;
    for (int y = (cy - hy > 0 ? cy - hy : 0); y <= (cy + hy < ires - 1 ? cy + hy : ires - 1); y++)
      for (int x = (cx - hx > 0 ? cx - hx : 0); x <= (cx + hx < ires - 1 ? cx + hx : ires - 1); x++) {
        v2 dt = V2(((float)x + 0.5f) - ct.x, ((float)y + 0.5f) - ct.y);
        float q = dot2(dt, m2v(Si, dt));
        if (!(q <= 9.0f)) continue;   /* outside 3 sigma, or a NaN footprint: adds nothing */
        float t = tau0 * exp_map(-0.5f * q);
        t = t < 64.0f ? t : 64.0f;
        float t_fix = t * FIX;
        if (t_fix < 0.5f) continue;
        float z = r + dot2(gz, dt);
        z = z > 0.0f ? z : 0.0f;
        size_t k = (base + (size_t)y * res + (size_t)x) * 3;
        const uint32_t a0 = u32_sat(t_fix + 0.5f), a1 = u32_sat(t_fix * z + 0.5f);
        const uint32_t a2 = u32_sat(t_fix * (z * z + sz2) + 0.5f);
#pragma omp atomic
        acc[k] += a0;
#pragma omp atomic
        acc[k + 1] += a1;
#pragma omp atomic
        acc[k + 2] += a2;
      }
  }
}

void or_lsm_resolve(const void *LS, const uint32_t *acc, float *mip) {
  uint32_t res = rdu(LS, 3), levels = rdu(LS, 7), lt = rdu(LS, 11), pc = rdu(LS, 19), dirty = rdu(LS, 15);
  uint32_t level_off[16];
  for (int k = 0; k < 16; k++) level_off[k] = rdu(LS, 36 + k);
  for (uint32_t p = 0; p < pc; p++) {
    if ((dirty & (1u << p)) == 0u) continue;
    resolve_layer(acc + 3 * (size_t)p * res * res, 65536.0f, res * res, mip + 4 * (size_t)p * lt);
    down_layer(mip + 4 * (size_t)p * lt, level_off, levels, res);
  }
}

void or_lsm_lookup(const void *U, const void *LS, const void *gaussians, const void *parts, const float *mip,
                   float *T_out) {
  uint32_t n = rdu(U, 1), res = rdu(LS, 3), levels = rdu(LS, 7), lt = rdu(LS, 11), pc = rdu(LS, 19);
  uint32_t skip = rdu(LS, 17);
  float sigma = rdf(LS, 16);
  v3 center = rd3(LS, 0);
  uint32_t level_off[16];
  for (int k = 0; k < 16; k++) level_off[k] = rdu(LS, 36 + k);
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint32_t g = 0; g < n; g++) {
    const char *G = (const char *)gaussians + 64 * (size_t)g;
    uint32_t own = rdu(G, 11) & 0xFFFFu;
    v3 P = xf_point((const char *)parts + 128 * (size_t)own, rd3(G, 0));
    v3 dv = sub3(P, center);
    float z = len3(dv);
    if (z < 1e-4f) { T_out[g] = 1.0f; continue; }
    v3 w = V3(dv.x / z, dv.y / z, dv.z / z);
    v3 b0 = basis_k(w), b1 = cross3(w, b0);
    m2 J = oct_jac(LS, w, b0, b1);
    float dj = fabsf(m2det(J));
    float tex_ang = (2.0f / (float)res) / sqrtf(dj > 1e-12f ? dj : 1e-12f);
    v2 uv = oct_uv(LS, w);
    v2 tc = V2((uv.x * 0.5f + 0.5f) * (float)res, (uv.y * 0.5f + 0.5f) * (float)res);
    float T = 1.0f;
    for (uint32_t p = 0; p < pc; p++) {
      if (p == own || p == skip) continue;
      const char *pp = (const char *)parts + 128 * (size_t)p;
      v3 bsc = rd3(pp, 24);
      float bw = rdf(pp, 27);
      v3 bc = sub3(xf_point(pp, bsc), center);
      float dc = len3(bc);
      float zn = dc - bw > 0.02f ? dc - bw : 0.02f;
      if (z <= zn) continue;
      float search = 2.0f * sigma * (z - zn) / (zn * z);
      if (dc > bw) {
        float ang = acos_cr(clampf_(dot3(V3(bc.x / dc, bc.y / dc, bc.z / dc), w), -1.0f, 1.0f));
        if (ang > asin_cr(bw / dc) + search + 2.0f * tex_ang) continue;
      }
      float x = 2.0f * search / tex_ang;
      float ls = log2_cr(x > 1.0f ? x : 1.0f);
      float a0[4], s[4];
      const float *mp = mip + 4 * (size_t)p * lt;
      mip_tri(mp, level_off, res, levels, ceilf(ls), tc, a0);
      if (a0[0] < 1e-4f) continue;
      float zb = a0[1] / a0[0];
      zb = zb > 0.02f ? zb : 0.02f;
      float pen = sigma * (z - zb > 0.0f ? z - zb : 0.0f) / (zb * z);
      x = 2.0f * pen / tex_ang;
      mip_tri(mp, level_off, res, levels, log2_cr(x > 1.0f ? x : 1.0f), tc, s);
      if (s[0] < 1e-5f) continue;
      float mu = s[1] / s[0];
      float var = s[2] / s[0] - mu * mu;
      float sd = sqrtf(var > 1e-8f ? var : 1e-8f);
      T *= 1.0f - s[0] * phi_((z - mu) / sd);
    }
    T_out[g] = T;
  }
}

/* ---- sun maps (sun_map.wgsl:91-324) ----------------------------------------------------------------------------- */

/* SunParams words: n_layers 0, res 1, levels 2, layer_texels 3, n_lights 4, n_movable 5, margin 8, lights 12..27,
 * part_layers 28.., level_off 92... LayerDesc words: part 0, light 1, kind 2, dirty 3, c 4, half 6.
 * SunLayer (here 12 floats): c 0-1, half 2, zref 3, t1 4-6, part 7, t2 8-10, light 11. */
void or_sun_frame(const void *SM, const void *desc, const void *parts, float *layers) {
  uint32_t nl = rdu(SM, 0);
  float margin = rdf(SM, 8);
  for (uint32_t l = 0; l < nl; l++) {
    const char *d = (const char *)desc + 32 * (size_t)l;
    uint32_t p = rdu(d, 0), k = rdu(d, 1);
    v3 L = rd3(SM, 12 + 4 * k);
    v3 t1 = basis_k(L), t2 = cross3(L, t1);
    const char *pp = (const char *)parts + 128 * (size_t)p;
    v3 c = xf_point(pp, rd3(pp, 24));
    float bw = rdf(pp, 27);
    float ccx = dot3(c, t1), ccy = dot3(c, t2), half = bw + margin;
    if (rdu(d, 2) == 1u) { ccx = rdf(d, 4); ccy = rdf(d, 5); half = rdf(d, 6); }
    float *o = layers + 12 * (size_t)l;
    o[0] = ccx; o[1] = ccy; o[2] = half; o[3] = dot3(c, L) + bw;
    o[4] = t1.x; o[5] = t1.y; o[6] = t1.z; memcpy(o + 7, &p, 4);
    o[8] = t2.x; o[9] = t2.y; o[10] = t2.z; memcpy(o + 11, &k, 4);
  }
}

void or_sun_splat(const void *U, const void *SM, const void *desc, const float *layers, const void *gaussians,
                  const void *nodes, const void *parts, uint32_t *acc) {
  const float FIX = 4096.0f, LOWPASS = 0.3f;
  const int MAX_HALF = 32;
  uint32_t n = rdu(U, 1), res = rdu(SM, 1);
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint32_t g = 0; g < n; g++) {
    const char *G = (const char *)gaussians + 64 * (size_t)g;
    uint32_t p = rdu(G, 11) & 0xFFFFu;
    if (p >= 32u) continue;
    uint32_t first = rdu(SM, 28 + 2 * p), cnt = rdu(SM, 29 + 2 * p);
    for (uint32_t l = first; l < first + cnt; l++) {
      const char *d = (const char *)desc + 32 * (size_t)l;
      if (rdu(d, 3) == 0u || rdf(SM, 12 + 4 * rdu(d, 1) + 3) < 0.0f) continue;
      const float *ly = layers + 12 * (size_t)l;
      const char *nd = (const char *)nodes + 96 * (size_t)g;
      const char *xf = (const char *)parts + 128 * (size_t)p;
      m3 R = part_rot(xf);
      v3 x = xf_point(xf, rd3(G, 0));
      v3 w = rd3(SM, 12 + 4 * rdu(d, 1));
      v3 b0 = V3(ly[4], ly[5], ly[6]), b1 = V3(ly[8], ly[9], ly[10]);
      m3 C = leaf_cov(G, R);
      v3 c0 = m3v(C, b0), c1 = m3v(C, b1), cw = m3v(C, w);
      m2 Cbb = {V2(dot3(b0, c0), dot3(b1, c0)), V2(dot3(b0, c1), dot3(b1, c1))};
      v2 cbz = V2(dot3(b0, cw), dot3(b1, cw));
      m2 Cbb_i = inv2(Cbb);
      float sz2 = dot3(w, cw) - dot2(cbz, m2v(Cbb_i, cbz));
      sz2 = sz2 > 1e-12f ? sz2 : 1e-12f;
      float half = ly[2];
      float tex = (float)res / (2.0f * half);
      m2 S = m2s(Cbb, tex * tex);
      float det0 = m2det(S);
      S.c0.x += LOWPASS;
      S.c1.y += LOWPASS;
      float det1 = m2det(S);
      float tau0 = rdf(nd, 7) * dirm(nd, xf_dir_inv(xf, w)) * sqrtf(2.0f * PI_F * sz2) *
                   sqrtf((det0 > 0.0f ? det0 : 0.0f) / det1);
      m2 Si = inv2(S);
      float z0 = ly[3] - dot3(x, w);
      v2 cz = m2v(Cbb_i, cbz);
      v2 gz = V2(cz.x / tex, cz.y / tex);
      v2 ct = V2(((dot3(x, b0) - ly[0]) + half) * tex, ((dot3(x, b1) - ly[1]) + half) * tex);
      int hx = (int)ceilf(3.0f * sqrtf(S.c0.x)), hy = (int)ceilf(3.0f * sqrtf(S.c1.y));
      hx = hx < MAX_HALF ? hx : MAX_HALF;
      hy = hy < MAX_HALF ? hy : MAX_HALF;
      int cx = (int)floorf(ct.x), cy = (int)floorf(ct.y);
      size_t base = (size_t)l * res * res;
      int ires = (int)res;
      for (int y = (cy - hy > 0 ? cy - hy : 0); y <= (cy + hy < ires - 1 ? cy + hy : ires - 1); y++)
        for (int xx = (cx - hx > 0 ? cx - hx : 0); xx <= (cx + hx < ires - 1 ? cx + hx : ires - 1); xx++) {
          v2 dt = V2(((float)xx + 0.5f) - ct.x, ((float)y + 0.5f) - ct.y);
          float q = dot2(dt, m2v(Si, dt));
          if (!(q <= 9.0f)) continue;   /* outside 3 sigma, or a NaN footprint: adds nothing */
          float t = tau0 * exp_map(-0.5f * q);
          t = t < 64.0f ? t : 64.0f;
          float t_fix = t * FIX;
          if (t_fix < 0.5f) continue;
          float z = z0 - dot2(gz, dt);
          z = z > 0.0f ? z : 0.0f;
          size_t i = (base + (size_t)y * res + (size_t)xx) * 3;
          const uint32_t a0 = u32_sat(t_fix + 0.5f), a1 = u32_sat(t_fix * z + 0.5f);
          const uint32_t a2 = u32_sat(t_fix * (z * z + sz2) + 0.5f);
#pragma omp atomic
          acc[i] += a0;
#pragma omp atomic
          acc[i + 1] += a1;
#pragma omp atomic
          acc[i + 2] += a2;
        }
    }
  }
}

void or_sun_resolve(const void *SM, const void *desc, const uint32_t *acc, float *mip) {
  uint32_t nl = rdu(SM, 0), res = rdu(SM, 1), levels = rdu(SM, 2), lt = rdu(SM, 3);
  uint32_t level_off[16];
  for (int k = 0; k < 16; k++) level_off[k] = rdu(SM, 92 + k);
  for (uint32_t l = 0; l < nl; l++) {
    if (rdu((const char *)desc + 32 * (size_t)l, 3) == 0u) continue;
    resolve_layer(acc + 3 * (size_t)l * res * res, 4096.0f, res * res, mip + 4 * (size_t)l * lt);
    down_layer(mip + 4 * (size_t)l * lt, level_off, levels, res);
  }
}

void or_sun_lookup(const void *U, const void *SM, const void *desc, const float *layers, const void *gaussians,
                   const void *parts, const float *mip, float *env_T) {
  uint32_t n = rdu(U, 1), nl = rdu(SM, 0), res = rdu(SM, 1), levels = rdu(SM, 2), lt = rdu(SM, 3);
  uint32_t level_off[16];
  for (int k = 0; k < 16; k++) level_off[k] = rdu(SM, 92 + k);
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint32_t g = 0; g < n; g++) {
    const char *G = (const char *)gaussians + 64 * (size_t)g;
    uint32_t own = rdu(G, 11) & 0xFFFFu;
    v3 P = xf_point((const char *)parts + 128 * (size_t)own, rd3(G, 0));
    float T[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    uint32_t used = 0xFFFFFFFFu;
    for (uint32_t l = 0; l < nl; l++) {
      const float *ly = layers + 12 * (size_t)l;
      uint32_t lp, lk;
      memcpy(&lp, ly + 7, 4);
      memcpy(&lk, ly + 11, 4);
      uint32_t key = lp * 8u + lk;
      if (lp == own || key == used || rdf(SM, 12 + 4 * lk + 3) < 0.0f) continue;
      v3 w = rd3(SM, 12 + 4 * lk);
      float ux = dot3(P, V3(ly[4], ly[5], ly[6])) - ly[0], uy = dot3(P, V3(ly[8], ly[9], ly[10])) - ly[1];
      float inner = (rdu((const char *)desc + 32 * (size_t)l, 2) == 1u ? 0.9f : 1.0f) * ly[2];
      if (fabsf(ux) >= inner || fabsf(uy) >= inner) continue;
      used = key;
      float z = ly[3] - dot3(P, w);
      if (z <= 0.0f) continue;
      float texel = 2.0f * ly[2] / (float)res;
      v2 tc = V2((ux + ly[2]) / texel, (uy + ly[2]) / texel);
      float sig = rdf(SM, 12 + 4 * lk + 3);
      float search = 2.0f * sig * z;
      float x = 2.0f * search / texel;
      float a0[4], s[4];
      const float *mp = mip + 4 * (size_t)l * lt;
      mip_tri(mp, level_off, res, levels, log2_cr(x > 1.0f ? x : 1.0f), tc, a0);
      if (a0[0] < 1e-4f) continue;
      float zb = a0[1] / a0[0];
      float pen = sig * (z - zb > 0.0f ? z - zb : 0.0f);
      x = 2.0f * pen / texel;
      mip_tri(mp, level_off, res, levels, log2_cr(x > 1.0f ? x : 1.0f), tc, s);
      if (s[0] < 1e-5f) continue;
      float mu = s[1] / s[0];
      float var = s[2] / s[0] - mu * mu;
      float sd = sqrtf(var > 1e-8f ? var : 1e-8f);
      T[lk] *= 1.0f - s[0] * phi_((z - mu) / sd);
    }
    memcpy(env_T + 4 * (size_t)g, T, 16);
  }
}

/* ---- env_direct (environment.wgsl:207-229, 266-277), matte ------------------------------------------------------- */

/* Env words: sh 0-63, lights 64 + 8k (dir_std, irr), info 96. */
void or_env_direct(const void *U, const void *ENV, const void *gaussians, const void *nodes, const uint8_t *vgrid,
                   const void *parts, const void *materials, const float *iso_norm, const float *env_T,
                   float *colour) {
  uint32_t n = rdu(U, 1), mcount = rdu(U, 13), nlights = rdu(ENV, 96);
  if (nlights == 0u || (rdu(U, 15) & 1u) == 0u) return;
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint32_t g = 0; g < n; g++) {
    const char *G = (const char *)gaussians + 64 * (size_t)g;
    uint32_t mi = rdu(G, 15) & 0xFFFFu;
    uint32_t mmax = (mcount > 1u ? mcount : 1u) - 1u;
    if (mcount && (rdu((const char *)materials + 32 * (size_t)(mi < mmax ? mi : mmax), 5) & 1u)) continue;
    const char *nd = (const char *)nodes + 96 * (size_t)g;
    const char *xf = (const char *)parts + 128 * (size_t)(rdu(G, 11) & 0xFFFFu);
    float s3 = rdf(G, 8) * rdf(G, 9) * rdf(G, 10);
    v3 acc = V3(0.0f, 0.0f, 0.0f);
    for (uint32_t k = 0; k < (nlights < 4u ? nlights : 4u); k++) {
      v3 wi = xf_dir_inv(xf, rd3(ENV, 64 + 8 * k));
      float v = or_vgrid_lookup(vgrid + 128 * (size_t)g, wi.x, wi.y, wi.z);
      if (v <= 1e-7f) continue;
      float f = env_T[4 * (size_t)g + k] * v * or_footprint(nd, s3 * s3, wi.x, wi.y, wi.z);
      acc = add3(acc, mul3s(rd3(ENV, 68 + 8 * k), f));
    }
    v3 alb = rd3(G, 12);
    float inorm = iso_norm[g];
    colour[4 * (size_t)g + 0] = colour[4 * (size_t)g + 0] + alb.x * acc.x * inorm;
    colour[4 * (size_t)g + 1] = colour[4 * (size_t)g + 1] + alb.y * acc.y * inorm;
    colour[4 * (size_t)g + 2] = colour[4 * (size_t)g + 2] + alb.z * acc.z * inorm;
  }
}
