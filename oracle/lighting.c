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

/* ---- shade_emit (runtime/wgsl/emissive.wgsl:123-150, 853-980), M39 transfer, matte and shiny ----------------- */

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

/* ---- shiny receivers (runtime/wgsl/material_ext.wgsl:10-181) ------------------------------------------------ */

/* material_ext.wgsl:12-14 */
static float sym_det(v3 d, v3 o) {
  return d.x * (d.y * d.z - o.z * o.z) - o.x * (o.x * d.z - o.z * o.y) + o.y * (o.x * o.z - d.y * o.y);
}
/* w^T S^-1 w det S via the adjugate (material_ext.wgsl:17-25) */
static float sym_adj_quad(v3 d, v3 o, v3 w) {
  float a00 = d.y * d.z - o.z * o.z, a11 = d.x * d.z - o.y * o.y, a22 = d.x * d.y - o.x * o.x;
  float a01 = o.y * o.z - o.x * d.z, a02 = o.x * o.z - o.y * d.y, a12 = o.x * o.y - d.x * o.z;
  return a00 * w.x * w.x + a11 * w.y * w.y + a22 * w.z * w.z +
         2.0f * (a01 * w.x * w.y + a02 * w.x * w.z + a12 * w.y * w.z);
}
/* specular flake phase, Heitz et al. 2015 (material_ext.wgsl:27-37) */
static float phase_specular(v3 Sd, v3 So, v3 wi, v3 wo) {
  v3 h = add3(wi, wo);
  float hl = len3(h);
  if (hl < 1e-6f) return 0.0f;
  v3 wh = V3(h.x / hl, h.y / hl, h.z / hl);
  float det = sym_det(Sd, So);
  det = det > 1e-20f ? det : 1e-20f;
  float q = sym_adj_quad(Sd, So, wh) / det;
  float D = 1.0f / (PI_F * sqrtf(det) * q * q);
  float sq = dot3(wi, sym_mul(Sd, So, wi));
  float sigma = sqrtf(sq > 1e-12f ? sq : 1e-12f);
  return D / (4.0f * sigma);
}
/* the phase toward n directions wo[3n] (for quadrature) */
void or_phase_specular_n(const float *Sd, const float *So, const float *wi, const float *wo, uint32_t n, float *out) {
  v3 d = V3(Sd[0], Sd[1], Sd[2]), o = V3(So[0], So[1], So[2]), i = V3(wi[0], wi[1], wi[2]);
#pragma omp parallel for schedule(static)
  for (uint32_t k = 0; k < n; k++) out[k] = phase_specular(d, o, i, V3(wo[3 * k], wo[3 * k + 1], wo[3 * k + 2]));
}
float or_phase_specular(const float *Sd, const float *So, const float *wi, const float *wo) {
  return phase_specular(V3(Sd[0], Sd[1], Sd[2]), V3(So[0], So[1], So[2]), V3(wi[0], wi[1], wi[2]),
                        V3(wo[0], wo[1], wo[2]));
}
/* spec_weight (material_ext.wgsl:39-42): the mirror-flake mix k of a shiny, non-emissive material */
static float spec_weight(const char *mat, uint32_t mflags) {
  if ((mflags & 8u) == 0u || (mflags & 1u) != 0u) return 0.0f;
  return clampf_(unpack_half(rdu(mat, 6) & 0xFFFFu), 0.0f, 1.0f);
}
/* iso_lobe_from, iso_damp, mirror (material_ext.wgsl:56-67) */
static float iso_lobe_from(float k, float p_spec, float damp) { return (1.0f - k) + k * 4.0f * PI_F * p_spec * damp; }
static float iso_damp(float inorm, float area_vis_r) {
  float m = 4.0f * PI_F * inorm * area_vis_r;
  float r = 1.0f / (m > 1e-30f ? m : 1e-30f);
  return r < 1.0f ? r : 1.0f;
}
static v3 mirror(v3 wi, v3 n) { return sub3(mul3s(n, 2.0f * dot3(wi, n)), wi); }

/* ASG lobe fit (material_ext.wgsl:79-80) */
static const float ASG_W[3] = {0.61163f, 0.35083f, 0.0267f};
static const float ASG_A[3] = {2.54696f, 0.6893f, 0.10727f};
/* orthonormal tangent pair, Duff et al. 2017 (material_ext.wgsl:83-88) */
static void onb(v3 n, v3 out[2]) {
  float s = n.z >= 0.0f ? 1.0f : -1.0f;
  float a = -1.0f / (s + n.z);
  float b = n.x * n.y * a;
  out[0] = V3(1.0f + s * n.x * n.x * a, s * b, -s * n.x);
  out[1] = V3(b, s + n.y * n.y * a, -n.y);
}
void or_onb(const float *n, float *out) {
  v3 b[2];
  onb(V3(n[0], n[1], n[2]), b);
  out[0] = b[0].x; out[1] = b[0].y; out[2] = b[0].z; out[3] = b[1].x; out[4] = b[1].y; out[5] = b[1].z;
}
/* symmetric 2x2 as (xx, xy, yy) (material_ext.wgsl:91-115) */
static float s2_det(v3 s) { return s.x * s.z - s.y * s.y; }
static float s2_quad(v3 s, v2 v) { return s.x * v.x * v.x + 2.0f * s.y * v.x * v.y + s.z * v.y * v.y; }
static float s2_quad_inv(v3 s, v2 v) {
  float d = s2_det(s);
  return (s.z * v.x * v.x - 2.0f * s.y * v.x * v.y + s.x * v.y * v.y) / (d > 1e-30f ? d : 1e-30f);
}
static v3 s2_xf(v3 s, v2 t0, v2 t1) {
  v2 a = V2(s.x * t0.x + s.y * t0.y, s.y * t0.x + s.z * t0.y);
  v2 b = V2(s.x * t1.x + s.y * t1.y, s.y * t1.x + s.z * t1.y);
  return V3(dot2(t0, a), dot2(t0, b), dot2(t1, b));
}
static v3 s2_major(v3 s) {
  float h = 0.5f * (s.x - s.z);
  float lam = 0.5f * (s.x + s.z) + sqrtf(h * h + s.y * s.y);
  v2 v = s.x >= s.z ? V2(lam - s.z, s.y) : V2(s.y, lam - s.x);
  float l = sqrtf(dot2(v, v));
  v2 d = l > 1e-20f ? V2(v.x / l, v.y / l) : V2(1.0f, 0.0f);
  return V3(d.x, d.y, lam);
}
static float s2_minor(v3 s) {
  float h = 0.5f * (s.x - s.z);
  return 0.5f * (s.x + s.z) - sqrtf(h * h + s.y * s.y);
}
/* unit a in the tangent plane at c: geodesic angle times direction (material_ext.wgsl:118-124) */
static v2 log_map(v3 c, const v3 cb[2], v3 a) {
  float ca = clampf_(dot3(c, a), -1.0f, 1.0f);
  v3 v = sub3(a, mul3s(c, ca));
  float nv = len3(v);
  if (nv < 1e-9f) return V2(0.0f, 0.0f);
  float k = acosf(ca) / nv;
  return V2(k * dot3(cb[0], v), k * dot3(cb[1], v));
}
static v3 norm3s(v3 a) { float l = len3(a); return V3(a.x / l, a.y / l, a.z / l); }
/* the specular phase integrated over a light ASG (material_ext.wgsl:128-181) */
static float asg_phase(v3 wi, const v3 lb[2], v3 sl, v3 n0, v3 t1, v3 t2, v3 flake, v3 Sd, v3 So, v3 wo) {
  float pc = phase_specular(Sd, So, wi, wo);
  float fm = flake.y > flake.z ? flake.y : flake.z;
  if (flake.x <= fm * 1.0001f) return pc;
  float co = dot3(n0, wo);
  v3 n = co >= 0.0f ? n0 : V3(-n0.x, -n0.y, -n0.z);
  float cos_o = fabsf(co);
  v3 wr = sub3(mul3s(n, 2.0f * cos_o), wo);
  v3 rb[2];
  onb(wr, rb);
  v3 dw1 = add3(mul3s(n, 2.0f * dot3(wo, t1)), mul3s(t1, 2.0f * cos_o));
  v3 dw2 = add3(mul3s(n, 2.0f * dot3(wo, t2)), mul3s(t2, 2.0f * cos_o));
  v2 j0 = V2(dot3(rb[0], dw1), dot3(rb[0], dw2));
  v2 j1 = V2(dot3(rb[1], dw1), dot3(rb[1], dw2));
  v2 ia = V2(flake.y / (flake.x - flake.y), flake.z / (flake.x - flake.z));
  v3 s0 = V3(dot2(V2(j0.x * j0.x, j0.y * j0.y), ia), dot2(V2(j0.x * j1.x, j0.y * j1.y), ia),
             dot2(V2(j1.x * j1.x, j1.y * j1.y), ia));
  float qr = dot3(wr, sym_mul(Sd, So, wr));
  float pk = flake.x * flake.x / (PI_F * sqrtf(flake.x * flake.y * flake.z)) /
             (4.0f * sqrtf(qr > 1e-12f ? qr : 1e-12f));
  v3 cs = add3(wi, wr);
  float cl = len3(cs);
  if (cl < 1e-4f) return pc;
  v3 c = V3(cs.x / cl, cs.y / cl, cs.z / cl);
  v3 cb[2];
  onb(c, cb);
  v2 lr = log_map(c, cb, wr), li = log_map(c, cb, wi);
  v2 dm = V2(lr.x - li.x, lr.y - li.y);
  v3 sL = s2_xf(sl, V2(dot3(cb[0], lb[0]), dot3(cb[0], lb[1])), V2(dot3(cb[1], lb[0]), dot3(cb[1], lb[1])));
  v3 sR = s2_xf(s0, V2(dot3(cb[0], rb[0]), dot3(cb[0], rb[1])), V2(dot3(cb[1], rb[0]), dot3(cb[1], rb[1])));
  float ldm = sqrtf(dot2(dm, dm));
  float ndm = ldm > 1e-9f ? ldm : 1e-9f;
  float a2 = 2.0f * ASG_A[2];
  float d = sqrtf(s2_quad_inv(V3(sL.x + sR.x / a2, sL.y + sR.y / a2, sL.z + sR.z / a2), dm));
  float qa = s2_quad(sL, dm);
  float along = sqrtf(qa > 0.0f ? qa : 0.0f) / ndm;
  float far_ = smoothstep_(1.0f, 2.5f, d) * (1.0f - smoothstep_(0.15f, 0.3f, along / ndm));
  v3 maj = s2_major(sl);
  float mn = s2_minor(s0) / (2.0f * ASG_A[0]);
  float ratio = sqrtf(maj.z / (mn > 1e-30f ? mn : 1e-30f));
  float t = 1.0f - smoothstep_(0.12f, 0.3f, ratio);
  t = far_ > t ? far_ : t;
  float p_asg = 0.0f;
  if (t < 1.0f) {
    for (int k = 0; k < 3; k++) {
      float dk = 2.0f * ASG_A[k];
      v3 sk = V3(sR.x / dk, sR.y / dk, sR.z / dk);
      v3 st = add3(sL, sk);
      float dsk = s2_det(sk), dst = s2_det(st);
      p_asg += ASG_W[k] * sqrtf((dsk > 0.0f ? dsk : 0.0f) / (dst > 1e-30f ? dst : 1e-30f)) *
               expf(-0.5f * s2_quad_inv(st, dm));
    }
    p_asg *= pk;
  }
  float p_pt = pc;
  if (t > 0.0f) {
    float sq = sqrtf(3.0f * maj.z);
    v3 ax = mul3s(add3(mul3s(lb[0], maj.x), mul3s(lb[1], maj.y)), sq);
    p_pt = (2.0f / 3.0f) * pc + (1.0f / 6.0f) * (phase_specular(Sd, So, norm3s(add3(wi, ax)), wo) +
                                                 phase_specular(Sd, So, norm3s(sub3(wi, ax)), wo));
  }
  return mixf(p_asg, p_pt, t);
}
/* asg_phase with its arguments as flat arrays: lb 6 floats, sl 3, n0/t1/t2/flake/Sd/So/wi/wo 3 each */
float or_asg_phase(const float *wi, const float *lb, const float *sl, const float *n0, const float *t1,
                   const float *t2, const float *flake, const float *Sd, const float *So, const float *wo) {
  v3 b[2] = {V3(lb[0], lb[1], lb[2]), V3(lb[3], lb[4], lb[5])};
  return asg_phase(V3(wi[0], wi[1], wi[2]), b, V3(sl[0], sl[1], sl[2]), V3(n0[0], n0[1], n0[2]),
                   V3(t1[0], t1[1], t1[2]), V3(t2[0], t2[1], t2[2]), V3(flake[0], flake[1], flake[2]),
                   V3(Sd[0], Sd[1], Sd[2]), V3(So[0], So[1], So[2]), V3(wo[0], wo[1], wo[2]));
}
/* S = R diag(flake) R^T in the rest frame as (diag, off = xy, xz, yz) (emissive.wgsl:141-150) */
static void flake_matrix(m3 Rq, v3 flake, v3 *Sd, v3 *So) {
  m3 D = {V3(flake.x, 0.0f, 0.0f), V3(0.0f, flake.y, 0.0f), V3(0.0f, 0.0f, flake.z)};
  m3 Sl = m3mul(m3mul(Rq, D), m3t(Rq));
  *Sd = V3(Sl.c0.x, Sl.c1.y, Sl.c2.z);
  *So = V3(Sl.c1.x, Sl.c2.x, Sl.c2.y);
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
    /* a shiny receiver (emissive.wgsl:868-876, 897-930): its flake matrix, the eye in its rest frame, and whether
     * each cut node is taken as a shape (EP.n5 bit 3) */
    float ks = spec_weight(mat, mflags);
    int shape = ks > 0.0f && (rdu(EP, 11) & 8u) != 0u;
    v3 flake = rd3(mat, 0), Sd = V3(0.0f, 0.0f, 0.0f), So = V3(0.0f, 0.0f, 0.0f);
    v3 wo_r = xf_dir_inv(xf, wo);
    if (ks > 0.0f) flake_matrix(Rq, flake, &Sd, &So);
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
      if (ks <= 0.0f) {
        acc = add3(acc, light);
        continue;
      }
      v3 wr = mirror(wi, Rq.c0);
      float damp = iso_damp(iso_norm[g], or_footprint(nd, detC, wr.x, wr.y, wr.z) * or_vgrid_lookup(vg, wr.x, wr.y, wr.z));
      float lobe;
      if (shape) {
        /* the node's footprint seen from the receiver, B C_s B^T / d^2, in a tangent basis at ww */
        v3 bw[2];
        onb(ww, bw);
        v3 cd = rd3(e, 16), cof = rd3(e, 20);
        v3 c0 = sym_mul(cd, cof, bw[0]), c1 = sym_mul(cd, cof, bw[1]);
        v3 sl = V3(dot3(bw[0], c0) / d2, dot3(bw[0], c1) / d2, dot3(bw[1], c1) / d2);
        v3 lb[2] = {xf_dir_inv(xf, bw[0]), xf_dir_inv(xf, bw[1])};
        float p = asg_phase(wi, lb, sl, Rq.c0, Rq.c1, Rq.c2, flake, Sd, So, wo_r);
        lobe = iso_lobe_from(ks, p, damp);
      } else {
        lobe = iso_lobe_from(ks, phase_specular(Sd, So, wi, wo_r), damp);
      }
      acc = add3(acc, mul3s(light, lobe));
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

/* ---- env_direct (environment.wgsl:207-229, 266-307), matte and shiny -------------------------------------------- */

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
    const char *mat = (const char *)materials + 32 * (size_t)(mi < mmax ? mi : mmax);
    float ks = mcount ? spec_weight(mat, rdu(mat, 5)) : 0.0f;
    if (ks > 0.0f) {
      /* shiny (environment.wgsl:278-307): (1 - k) of the matte light plus each light's highlight, its shape the
       * light's ASG of angular std dir_std.w */
      v3 P = xf_point(xf, rd3(G, 0));
      v3 wo = xf_dir_inv(xf, norm3(sub3(rd3(U, 8), P)));
      m3 Rq = quat_matrix(rdf(G, 4), rdf(G, 5), rdf(G, 6), rdf(G, 7));
      v3 flake = rd3(mat, 0), Sd, So;
      flake_matrix(Rq, flake, &Sd, &So);
      float inorm = iso_norm[g];
      v3 diff = V3(0.0f, 0.0f, 0.0f), spec = V3(0.0f, 0.0f, 0.0f);
      for (uint32_t k = 0; k < (nlights < 4u ? nlights : 4u); k++) {
        v3 wi = xf_dir_inv(xf, rd3(ENV, 64 + 8 * k));
        float v = or_vgrid_lookup(vgrid + 128 * (size_t)g, wi.x, wi.y, wi.z);
        if (v <= 1e-7f) continue;
        v3 light = mul3s(rd3(ENV, 68 + 8 * k), env_T[4 * (size_t)g + k] * v * or_footprint(nd, s3 * s3, wi.x, wi.y, wi.z));
        v3 wr = mirror(wi, Rq.c0);
        float damp = iso_damp(inorm, or_footprint(nd, s3 * s3, wr.x, wr.y, wr.z) *
                                         or_vgrid_lookup(vgrid + 128 * (size_t)g, wr.x, wr.y, wr.z));
        float sd = rdf(ENV, 67 + 8 * k);
        float s2 = sd * sd;
        v3 lb[2];
        onb(wi, lb);
        float p = asg_phase(wi, lb, V3(s2, 0.0f, s2), Rq.c0, Rq.c1, Rq.c2, flake, Sd, So, wo);
        diff = add3(diff, mul3s(light, 1.0f - ks));
        spec = add3(spec, mul3s(light, iso_lobe_from(ks, p, damp) - (1.0f - ks)));
      }
      v3 alb = rd3(G, 12);
      v3 a = V3(alb.x * inorm, alb.y * inorm, alb.z * inorm);
      colour[4 * (size_t)g + 0] = colour[4 * (size_t)g + 0] + a.x * (diff.x + spec.x);
      colour[4 * (size_t)g + 1] = colour[4 * (size_t)g + 1] + a.y * (diff.y + spec.y);
      colour[4 * (size_t)g + 2] = colour[4 * (size_t)g + 2] + a.z * (diff.z + spec.z);
      continue;
    }
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
