/* CPU oracle for the sky and bounce through probes: TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Float32 restatement of runtime/wgsl/bounce.wgsl's default path (no sender LOD, no near field, M39 matte with SH2
 * probes: U.iso bits 0 and 1) plus the two passes that feed it from emissive.wgsl and environment.wgsl, evaluated
 * in WGSL expression order without contraction. Cell directions and solid angles come from a table the caller
 * builds in double and rounds once (bounce.wgsl cell_dir / cell_omega); the cell picks use double asin / atan2
 * rounded once, as the device does.
 *
 * Records are raw bytes: Gaussian 64 B, Node 96 B (mean 0, radius 3, cinv_diag 4, kappa 7, cinv_off 8, smax 11,
 * sn_diag 12, part 15, sn_off 16, child_lo 19, child_hi 20, depth 21), Part 128 B (xf 0, xf_prev 12, bound 24,
 * root 28, flags 31), FrameUniforms 64 B (frame 0, gaussian_count 1, part_count 2, material_count 13, iso 15),
 * Env 416 B (sh 0, lights 64, info 96), ProbeParams 9632 B (lo 0, dims.x 3, step 4, node_count 7, block_rad 8,
 * theta 9, tmax 10, near_r 11, alpha 12, force_near 13, g_valid 15, subset 16, phase 17, part_probes 18,
 * sh_full 19, dirs 20, pinv 76, refl 204, dims 208, levels 212, ground 228, vols 1252, vol_info 1380).
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

typedef struct { float x, y, z; } v3;

static const float PI_F = 3.14159265358979f;
#define NCELL 128u
#define HIT_MAX 16u
#define STACK 64u
#define PPK 8u
#define NONE 0xFFFFFFFFu

static v3 V3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static v3 add3(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub3(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 mul3s(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static v3 mul3(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static v3 div3(v3 a, v3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
static float dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static float len3(v3 a) { return sqrtf(dot3(a, a)); }
static float clampf_(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
static float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
static v3 max3s(v3 a, float s) { return V3(fmaxf(a.x, s), fmaxf(a.y, s), fmaxf(a.z, s)); }
static v3 min3(v3 a, v3 b) { return V3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
static v3 clamp3(v3 a, float lo, float hi) { return V3(clampf_(a.x, lo, hi), clampf_(a.y, lo, hi), clampf_(a.z, lo, hi)); }
static float rdf(const void *p, size_t w) { float v; memcpy(&v, (const char *)p + 4 * w, 4); return v; }
static uint32_t rdu(const void *p, size_t w) { uint32_t v; memcpy(&v, (const char *)p + 4 * w, 4); return v; }
static v3 rd3(const void *p, size_t w) { return V3(rdf(p, w), rdf(p, w + 1), rdf(p, w + 2)); }

static const char *REC(const void *base, size_t stride, size_t i) { return (const char *)base + stride * i; }
#define GAUSS(g) REC(gaussians, 64, (g))
#define NODE(n) REC(nodes, 96, (n))
#define PART(q) REC(parts, 128, (q))

/* common.wgsl:79-94 */
static v3 xf_point(const void *xf, v3 p) {
  return V3(dot3(rd3(xf, 0), p) + rdf(xf, 3), dot3(rd3(xf, 4), p) + rdf(xf, 7), dot3(rd3(xf, 8), p) + rdf(xf, 11));
}
static v3 xf_dir_inv(const void *xf, v3 d) {
  return add3(add3(mul3s(rd3(xf, 0), d.x), mul3s(rd3(xf, 4), d.y)), mul3s(rd3(xf, 8), d.z));
}
static v3 xf_point_inv(const void *xf, v3 p) { return xf_dir_inv(xf, sub3(p, V3(rdf(xf, 3), rdf(xf, 7), rdf(xf, 11)))); }
static v3 sym_mul(v3 dg, v3 off, v3 v) {   /* common.wgsl:120-124 */
  return V3((dg.x * v.x + off.x * v.y) + off.y * v.z, (off.x * v.x + dg.y * v.y) + off.z * v.z,
            (off.y * v.x + off.z * v.y) + dg.z * v.z);
}
static float dirm(const void *nd, v3 w) {  /* common.wgsl:127-129 */
  float q = dot3(w, sym_mul(rd3(nd, 12), rd3(nd, 16), w));
  return sqrtf(fmaxf(q, 0.0f));
}
/* common.wgsl:133-152 */
static float seg_geom(const void *nd, v3 P, v3 w, float t0, float t1) {
  v3 d = sub3(rd3(nd, 0), P);
  v3 cw = sym_mul(rd3(nd, 4), rd3(nd, 8), w);
  float a = dot3(cw, w);
  float m = dot3(cw, d) / a;
  v3 e = sub3(d, mul3s(w, m));
  float q = fmaxf(dot3(e, sym_mul(rd3(nd, 4), rd3(nd, 8), e)), 0.0f);
  float s = sqrtf(0.5f * a);
  float x0 = s * (t0 - m), x1 = s * (t1 - m);
  float df;
  if (x0 > 0.0f) df = or_erfc_pos(x0) - or_erfc_pos(x1);
  else if (x1 < 0.0f) df = or_erfc_pos(-x1) - or_erfc_pos(-x0);
  else df = (2.0f - or_erfc_pos(x1)) - or_erfc_pos(-x0);
  return expf(-0.5f * q) * sqrtf(PI_F / (2.0f * a)) * df;
}
static float vgrid_at(const uint8_t *vgrid, uint32_t g, v3 w) { return or_vgrid_lookup(vgrid + 128 * (size_t)g, w.x, w.y, w.z); }
static float footprint(const void *nd, float detC, v3 w) { return or_footprint(nd, detC, w.x, w.y, w.z); }

/* f32 -> f16 bits, round to nearest even (WGSL pack2x16float) */
uint32_t or_f16_bits(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t s = (x >> 16) & 0x8000u, e = (x >> 23) & 0xFFu, m = x & 0x7FFFFFu;
  if (e == 0xFFu) return s | 0x7C00u | (m ? 0x200u : 0u);
  int E = (int)e - 127 + 15;
  if (E >= 31) return s | 0x7C00u;
  if (E <= 0) {                                   /* subnormal half (or zero) */
    if (E < -10) return s;
    uint32_t mm = m | 0x800000u;
    uint32_t shift = (uint32_t)(14 - E);
    uint32_t h = mm >> shift, rem = mm & ((1u << shift) - 1u), half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (h & 1u))) h++;
    return s | h;
  }
  uint32_t h = ((uint32_t)E << 10) | (m >> 13), rem = m & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return s | h;
}
static float half_of(uint32_t h) {
  uint32_t s = (h >> 15) & 1u, e = (h >> 10) & 31u, m = h & 1023u;
  float v = e == 0 ? ldexpf((float)m, -24) : (e == 31 ? (m ? NAN : INFINITY) : ldexpf((float)(m | 1024u), (int)e - 25));
  return s ? -v : v;
}

/* ---- ProbeParams accessors (bounce.wgsl:23-78) ------------------------------------------------------------- */
typedef struct {
  const void *PP, *U, *parts, *nodes;
  const float *cells;      /* 128 x 4: direction xyz, solid angle */
} Ctx;

static uint32_t lat_count(const void *PP) { return rdu(PP, 208) * rdu(PP, 209) * rdu(PP, 210); }
static uint32_t lat_index(const void *PP, uint32_t i, uint32_t j, uint32_t k) { return (i * rdu(PP, 209) + j) * rdu(PP, 210) + k; }
static float lat_level(const void *PP, uint32_t k) { return rdf(PP, 212 + k); }
static float lat_ground(const void *PP, uint32_t i, uint32_t j) { return rdf(PP, 228 + i * rdu(PP, 209) + j); }
static uint32_t part_count(const Ctx *c) { return rdu(c->U, 2); }

static uint32_t vol_of_probe(const Ctx *c, uint32_t r) {   /* bounce.wgsl:61-68 */
  uint32_t n = part_count(c) < 32u ? part_count(c) : 32u;
  for (uint32_t q = 0; q < n; q++) {
    const uint32_t *v = (const uint32_t *)((const char *)c->PP + 4 * (1252 + 4 * q));
    uint32_t vx, vy, vz, vw;
    memcpy(&vx, v, 4); memcpy(&vy, v + 1, 4); memcpy(&vz, v + 2, 4); memcpy(&vw, v + 3, 4);
    if (vy > 0u && r >= vx && r < vx + vy * vz * vw) return q;
  }
  return NONE;
}
static int has_vol(const Ctx *c, uint32_t q) { return q < 32u && rdu(c->PP, 1252 + 4 * q + 1) > 0u; }

typedef struct { v3 c, h; } Frame;
/* bounce.wgsl:256-262 */
static Frame part_probe_frame(const Ctx *c, uint32_t q) {
  const void *parts = c->parts;
  const char *pq = PART(q);
  const void *nodes = c->nodes;
  const char *nd = NODE(rdu(pq, 28));
  v3 a = rd3(nd, 4), o = rd3(nd, 8);
  float det = (a.x * (a.y * a.z - o.z * o.z) - o.x * (o.x * a.z - o.z * o.y)) + o.y * (o.x * o.z - a.y * o.y);
  v3 var = mul3s(V3(a.y * a.z - o.z * o.z, a.x * a.z - o.y * o.y, a.x * a.y - o.x * o.x), 1.0f);
  float dd = fmaxf(det, 1e-30f);
  var = V3(var.x / dd, var.y / dd, var.z / dd);
  v3 s = V3(sqrtf(fmaxf(var.x, 0.0f)), sqrtf(fmaxf(var.y, 0.0f)), sqrtf(fmaxf(var.z, 0.0f)));
  float hb = 0.5f * rdf(pq, 27);
  Frame f = {rd3(nd, 0), min3(mul3s(s, 0.95f), V3(hb, hb, hb))};
  return f;
}
static Frame vol_frame(const Ctx *c, uint32_t q) {   /* bounce.wgsl:70-73 */
  const void *parts = c->parts;
  Frame f = part_probe_frame(c, q);
  float b = rdf(PART(q), 27);
  f.h = min3(mul3s(f.h, 1.6f), V3(b, b, b));
  return f;
}
static int is_part_probe_part(const Ctx *c, uint32_t q) {   /* bounce.wgsl:264-266 */
  const void *parts = c->parts;
  return rdu(c->PP, 18) > 0u && q < rdu(c->PP, 1380) / PPK && (rdu(PART(q), 31) & 1u) == 0u;
}
/* bounce.wgsl:217-244 */
static v3 probe_pos(const Ctx *c, uint32_t p) {
  const void *PP = c->PP, *parts = c->parts;
  uint32_t nl = lat_count(PP);
  if (p >= nl + rdu(PP, 1380)) {
    uint32_t r = p - nl, q = vol_of_probe(c, r);
    if (q == NONE) return V3(0.0f, 0.0f, 0.0f);
    uint32_t dx = rdu(PP, 1252 + 4 * q + 1), dy = rdu(PP, 1252 + 4 * q + 2), dz = rdu(PP, 1252 + 4 * q + 3);
    uint32_t l = r - rdu(PP, 1252 + 4 * q);
    v3 ijk = V3((float)(l / (dy * dz)), (float)((l / dz) % dy), (float)(l % dz));
    Frame vf = vol_frame(c, q);
    v3 den = V3(fmaxf((float)dx - 1.0f, 1.0f), fmaxf((float)dy - 1.0f, 1.0f), fmaxf((float)dz - 1.0f, 1.0f));
    v3 t = div3(mul3(mul3s(vf.h, 2.0f), ijk), den);
    return xf_point(PART(q), add3(sub3(vf.c, vf.h), t));
  }
  if (p >= nl) {
    uint32_t q = (p - nl) / PPK, k = (p - nl) % PPK;
    v3 sgn = V3((float)(k >> 2u) * 2.0f - 1.0f, (float)((k >> 1u) & 1u) * 2.0f - 1.0f, (float)(k & 1u) * 2.0f - 1.0f);
    Frame pf = part_probe_frame(c, q);
    return xf_point(PART(q), add3(pf.c, mul3(sgn, pf.h)));
  }
  uint32_t dy = rdu(PP, 209), dz = rdu(PP, 210);
  uint32_t i = p / (dy * dz), j = (p / dz) % dy, k = p % dz;
  return V3(rdf(PP, 0) + rdf(PP, 4) * (float)i, rdf(PP, 1) + rdf(PP, 5) * (float)j, lat_ground(PP, i, j) + lat_level(PP, k));
}

static v3 cell_dir(const Ctx *c, uint32_t k) { return V3(c->cells[4 * k], c->cells[4 * k + 1], c->cells[4 * k + 2]); }
/* bounce.wgsl:199-204 */
static uint32_t cell_of(v3 w) {
  float el = (float)asin((double)clampf_(w.z, -1.0f, 1.0f));
  float fi = fmaxf((0.5f * PI_F - el) / (PI_F / 8.0f), 0.0f);
  uint32_t i = (uint32_t)fi < 7u ? (uint32_t)fi : 7u;
  float az = (float)atan2((double)w.y, (double)w.x);
  uint32_t j = ((uint32_t)floorf((az + PI_F) / (2.0f * PI_F) * 16.0f)) % 16u;
  return i * 16u + j;
}
/* bounce.wgsl:206-210 */
static void sh2_basis(v3 d, float b[9]) {
  float x = d.x, y = d.y, z = d.z;
  b[0] = 0.282095f; b[1] = 0.488603f * y; b[2] = 0.488603f * z; b[3] = 0.488603f * x;
  b[4] = 1.092548f * x * y; b[5] = 1.092548f * y * z; b[6] = 0.315392f * (3.0f * z * z - 1.0f);
  b[7] = 1.092548f * x * z; b[8] = 0.546274f * (x * x - y * y);
}
/* bounce.wgsl:148-158 */
static v3 sky_radiance(const void *ENV, v3 w) {
  float x = w.x, y = w.y, z = w.z;
  float b[16] = {0.282095f, 0.488603f * y, 0.488603f * z, 0.488603f * x,
                 1.092548f * x * y, 1.092548f * y * z, 0.315392f * (3.0f * z * z - 1.0f), 1.092548f * x * z,
                 0.546274f * (x * x - y * y), 0.590044f * y * (3.0f * x * x - y * y), 2.890611f * x * y * z,
                 0.457046f * y * (5.0f * z * z - 1.0f), 0.373176f * z * (5.0f * z * z - 3.0f),
                 0.457046f * x * (5.0f * z * z - 1.0f), 1.445306f * z * (x * x - y * y),
                 0.590044f * x * (x * x - 3.0f * y * y)};
  v3 c = V3(0.0f, 0.0f, 0.0f);
  for (int k = 0; k < 16; k++) c = add3(c, mul3s(rd3(ENV, 4 * k), b[k]));
  return max3s(c, 0.0f);
}

/* ---- lo_dir: direct_dirs_emit (emissive.wgsl:1012-1016, 1067-1083) + env_dirs (environment.wgsl:311-359) -------
 * The matte receivers' emitter light toward the 14 fixed directions is the same in each (emissive.wgsl:930-937),
 * c0 = the shading's sum before an emitter's own radiance (or_shade_emit's colour for non-emissive materials);
 * one value toward the probes: its mean weighted by footprint x V over DIRS14. With U.iso bit 1 a quarter of the
 * Gaussians (g % 4 == frame % 4) is refreshed; env_dirs adds the distant lights' light to the same ones. */
static const float DIRS14[14][3] = {
    {1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {1e-4f, 0.0f, 1.0f}, {-1.0f, 0.0f, 0.0f}, {0.0f, -1.0f, 0.0f},
    {1e-4f, 0.0f, -1.0f}, {0.57735f, 0.57735f, 0.57735f}, {0.57735f, 0.57735f, -0.57735f},
    {0.57735f, -0.57735f, 0.57735f}, {0.57735f, -0.57735f, -0.57735f}, {-0.57735f, 0.57735f, 0.57735f},
    {-0.57735f, 0.57735f, -0.57735f}, {-0.57735f, -0.57735f, 0.57735f}, {-0.57735f, -0.57735f, -0.57735f}};

void or_bounce_lo_dir(const void *U, const void *ENV, const void *gaussians, const void *nodes, const uint8_t *vgrid,
                      const void *parts_, const void *materials, const float *iso_norm, const float *env_T,
                      const float *c0, float *lo_dir) {
  const void *parts = parts_;
  uint32_t n = rdu(U, 1), mcount = rdu(U, 13), iso = rdu(U, 15), frame = rdu(U, 0);
  int env_on = ENV && rdu(ENV, 96) > 0u && (iso & 1u) != 0u;
#pragma omp parallel for schedule(dynamic, 512)
  for (uint32_t g = 0; g < n; g++) {
    if ((iso & 2u) != 0u && g % 4u != frame % 4u) continue;
    const char *G = GAUSS(g);
    const char *nd = NODE(g);
    const char *xf = PART(rdu(G, 11) & 0xFFFFu);
    float s3 = rdf(G, 8) * rdf(G, 9) * rdf(G, 10);
    v3 c = rd3(c0, 4 * (size_t)g);
    v3 mean = V3(0.0f, 0.0f, 0.0f);
    float ws = 0.0f;
    for (int k = 0; k < 14; k++) {
      v3 u = xf_dir_inv(xf, V3(DIRS14[k][0], DIRS14[k][1], DIRS14[k][2]));
      float w = footprint(nd, s3 * s3, u) * vgrid_at(vgrid, g, u);
      mean = add3(mean, mul3s(c, w));
      ws += w;
    }
    v3 lo = ws > 0.0f ? mul3s(mean, 1.0f) : c;
    if (ws > 0.0f) { float d = fmaxf(ws, 1e-20f); lo = V3(mean.x / d, mean.y / d, mean.z / d); }
    uint32_t mi = rdu(G, 15) & 0xFFFFu, mmax = (mcount > 1u ? mcount : 1u) - 1u;
    int emissive = mcount && (rdu(REC(materials, 32, mi < mmax ? mi : mmax), 5) & 1u);
    if (env_on && !emissive) {
      v3 acc = V3(0.0f, 0.0f, 0.0f);
      uint32_t nl = rdu(ENV, 96) < 4u ? rdu(ENV, 96) : 4u;
      for (uint32_t k = 0; k < nl; k++) {
        v3 wi = xf_dir_inv(xf, rd3(ENV, 64 + 8 * k));
        float v = vgrid_at(vgrid, g, wi);
        if (v <= 1e-7f) continue;
        float f = env_T[4 * (size_t)g + k] * v * footprint(nd, s3 * s3, wi);
        acc = add3(acc, mul3s(rd3(ENV, 68 + 8 * k), f));
      }
      v3 alb = rd3(G, 12);
      float inorm = iso_norm[g];
      lo = add3(lo, V3(alb.x * acc.x * inorm, alb.y * acc.y * inorm, alb.z * acc.z * inorm));
    }
    lo_dir[4 * (size_t)g + 0] = lo.x; lo_dir[4 * (size_t)g + 1] = lo.y; lo_dir[4 * (size_t)g + 2] = lo.z;
    lo_dir[4 * (size_t)g + 3] = 0.0f;
  }
}

/* ---- node emitters (bounce.wgsl:1166-1233) ------------------------------------------------------------------- */

void or_node_leaf(const void *U, const void *PP, const void *gaussians, const void *nodes, const uint8_t *vgrid,
                  const void *parts_, const float *lo_dir, const float *lo_b, float *node_acc) {
  const void *parts = parts_;
  uint32_t n = rdu(U, 1), dark_part = rdu(PP, 207);
#pragma omp parallel for schedule(dynamic, 512)
  for (uint32_t g = 0; g < n; g++) {
    const char *G = GAUSS(g);
    const char *nd = NODE(g);
    const char *xf = PART(rdu(G, 11) & 0xFFFFu);
    float s3 = rdf(G, 8) * rdf(G, 9) * rdf(G, 10);
    int dark = dark_part != 0u && (rdu(G, 11) & 0xFFFFu) == dark_part - 1u;
    for (uint32_t k = 0; k < 14u; k++) {
      v3 wo = xf_dir_inv(xf, rd3(PP, 20 + 4 * k));
      float w = footprint(nd, s3 * s3, wo) * vgrid_at(vgrid, g, wo);
      v3 L = dark ? V3(0.0f, 0.0f, 0.0f) : add3(rd3(lo_dir, 4 * (size_t)g), rd3(lo_b, 4 * (3 * (size_t)g + 1)));
      float *o = node_acc + 4 * (14 * (size_t)g + k);
      o[0] = L.x * w; o[1] = L.y * w; o[2] = L.z * w; o[3] = w;
    }
  }
}

/* one depth of node_up: the nodes up_nodes[r0, r0 + count) */
void or_node_up(const void *U, const void *PP, const void *nodes, const uint32_t *kids, const uint32_t *up_nodes,
                uint32_t depth, uint32_t r0, uint32_t count, float *node_acc) {
  uint32_t G = rdu(U, 1), N = rdu(PP, 7);
  for (uint32_t t = 0; t < count; t++) {
    uint32_t i = up_nodes[r0 + t];
    if (i < G || i >= N) continue;
    const char *nd = NODE(i);
    if (rdu(nd, 21) != depth || rdu(nd, 15) == NONE) continue;
    for (uint32_t k = 0; k < 14u; k++) {
      float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      for (uint32_t j = rdu(nd, 19); j < rdu(nd, 20); j++) {
        const float *a = node_acc + 4 * (14 * (size_t)kids[j] + k);
        s[0] += a[0]; s[1] += a[1]; s[2] += a[2]; s[3] += a[3];
      }
      memcpy(node_acc + 4 * (14 * (size_t)i + k), s, 16);
    }
  }
}

void or_node_sh2(const void *PP, const float *node_acc, uint32_t *node_sh) {
  uint32_t N = rdu(PP, 7);
#pragma omp parallel for schedule(dynamic, 512)
  for (uint32_t i = 0; i < N; i++) {
    v3 L[14];
    for (uint32_t k = 0; k < 14u; k++) {
      const float *a = node_acc + 4 * (14 * (size_t)i + k);
      float d = fmaxf(a[3], 1e-12f);
      L[k] = V3(a[0] / d, a[1] / d, a[2] / d);
    }
    float h[28];
    h[27] = 0.0f;
    for (uint32_t m = 0; m < 9u; m++) {
      v3 c = V3(0.0f, 0.0f, 0.0f);
      for (uint32_t k = 0; k < 14u; k++) c = add3(c, mul3s(L[k], rdf(PP, 76 + m * 14 + k)));
      h[3 * m] = c.x; h[3 * m + 1] = c.y; h[3 * m + 2] = c.z;
    }
    for (uint32_t j = 0; j < 14u; j++) node_sh[14 * (size_t)i + j] = or_f16_bits(h[2 * j]) | (or_f16_bits(h[2 * j + 1]) << 16);
  }
}

/* bounce.wgsl:122-129 */
static v3 node_radiance(const uint32_t *node_sh, uint32_t ni, const float b[9]) {
  float h[28];
  for (uint32_t j = 0; j < 14u; j++) {
    uint32_t w = node_sh[14 * (size_t)ni + j];
    h[2 * j] = half_of(w & 0xFFFFu);
    h[2 * j + 1] = half_of(w >> 16);
  }
  v3 v = V3(0.0f, 0.0f, 0.0f);
  for (uint32_t m = 0; m < 9u; m++) v = add3(v, mul3s(V3(h[3 * m], h[3 * m + 1], h[3 * m + 2]), b[m]));
  return v;
}

/* ---- probe_gather (bounce.wgsl:1255-1415), rays [id0, id1) ---------------------------------------------------
 * No lamps end a ray here (hit_lamps returns tmax: the scenes carry emissive Gaussians, not rect lamps). The
 * gather is taken with PP.alpha in (0, 1] (an EMA; 1 by default); the FIR blend (alpha < 0) is not restated. */
void or_probe_gather(const void *U, const void *PP, const void *ENV, const void *nodes, const uint8_t *vgrid,
                     const void *parts_, const uint32_t *kids, const uint32_t *node_sh, const float *cells,
                     float *probe_data, uint32_t id0, uint32_t id1) {
  Ctx cx = {PP, U, parts_, nodes, cells};
  const Ctx *c = &cx;
  const void *parts = parts_;
  uint32_t G = rdu(U, 1);
  uint32_t np = lat_count(PP), part_probes = rdu(PP, 18);
  float theta = rdf(PP, 9), tmax = rdf(PP, 10), alpha = rdf(PP, 12);
#pragma omp parallel for schedule(dynamic, 64)
  for (uint32_t gid = id0; gid < id1; gid++) {
    uint32_t id = gid * rdu(PP, 16) + rdu(PP, 17);
    if (id >= (np + part_probes) * NCELL) continue;
    uint32_t p = id / NCELL, cell = id % NCELL;
    uint32_t skip = NONE;
    if (p >= np && p < np + rdu(PP, 1380)) {
      skip = (p - np) / PPK;
      if (!is_part_probe_part(c, skip)) continue;
    } else if (p >= np) {
      skip = vol_of_probe(c, p - np);
      if (skip == NONE) continue;
    }
    v3 o = probe_pos(c, p);
    v3 w = cell_dir(c, cell);
    uint32_t dy = rdu(PP, 209), dz = rdu(PP, 210);
    v3 b = V3((float)((p / (dy * dz)) / 2u), (float)(((p / dz) % dy) / 2u), (float)((p % dz) / 2u));
    int own_centre = p >= np || rdu(PP, 211) == 1u;
    v3 cen = own_centre ? o : add3(rd3(PP, 0), mul3(rd3(PP, 4), add3(mul3s(b, 2.0f), V3(0.5f, 0.5f, 0.5f))));
    float t1 = tmax;
    float ht[HIT_MAX], hv[HIT_MAX];
    uint32_t hn[HIT_MAX];
    uint32_t k = 0;
    float brad = (own_centre && rdu(PP, 1381) == 0u) ? 0.0f : rdf(PP, 8);
    for (uint32_t q = 0; q < part_count(c); q++) {
      if (q == skip) continue;
      const char *xf = PART(q);
      v3 Po = xf_point_inv(xf, o), wr = xf_dir_inv(xf, w), cr = xf_point_inv(xf, cen);
      uint32_t stack[STACK];
      stack[0] = rdu(xf, 28);
      uint32_t sp = 1;
      while (sp > 0) {
        sp -= 1;
        uint32_t ni = stack[sp];
        const char *nd = NODE(ni);
        v3 d = sub3(rd3(nd, 0), Po);
        float tt = clampf_(dot3(d, wr), 0.0f, t1);
        v3 e = sub3(d, mul3s(wr, tt));
        float r = rdf(nd, 3);
        if (dot3(e, e) > r * r) continue;
        int use = ni < G;
        if (!use) use = len3(sub3(cr, rd3(nd, 0))) - brad > theta * rdf(nd, 11);
        if (use) {
          float tau = rdf(nd, 7) * dirm(nd, wr) * seg_geom(nd, Po, wr, 0.0f, t1);
          if (tau > 1e-6f) {
            v3 cw = sym_mul(rd3(nd, 4), rd3(nd, 8), wr);
            float den = dot3(cw, wr);
            float tp = fabsf(den) > 1e-30f ? clampf_(dot3(cw, d) / den, 0.0f, t1) : t1;
            uint32_t at = k < HIT_MAX - 1u ? k : HIT_MAX - 1u;
            const uint32_t l = HIT_MAX - 1u;
            if (k == HIT_MAX && tp >= ht[l]) {
              ht[l] = fmaxf(ht[l], tp);
              hv[l] = fminf(hv[l] + tau, 64.0f);
              continue;
            }
            int full = k == HIT_MAX;
            float sv = hv[l], st = ht[l];
            while (at > 0u && ht[at - 1u] > tp) {
              ht[at] = ht[at - 1u]; hv[at] = hv[at - 1u]; hn[at] = hn[at - 1u];
              at -= 1u;
            }
            ht[at] = tp; hv[at] = tau; hn[at] = ni;
            if (full) {
              ht[l] = fmaxf(ht[l], st);
              hv[l] = fminf(hv[l] + sv, 64.0f);
            }
            k = k + 1u < HIT_MAX ? k + 1u : HIT_MAX;
          }
        } else {
          for (uint32_t j = rdu(nd, 19); j < rdu(nd, 20); j++)
            if (sp < STACK) stack[sp++] = kids[j];
        }
      }
    }
    float T = 1.0f, m1 = 0.0f, m2 = 0.0f, first_part = -1.0f, near = 0.0f;
    v3 acc = V3(0.0f, 0.0f, 0.0f);
    float bw[9];
    sh2_basis(V3(-w.x, -w.y, -w.z), bw);
    for (uint32_t i = 0; i < k; i++) {
      if (T <= 0.0f) break;
      float al = 1.0f - expf(-hv[i]);
      acc = add3(acc, mul3s(max3s(node_radiance(node_sh, hn[i], bw), 0.0f), T * al));
      float dep = clampf_(ht[i], 0.0f, t1);
      m1 += T * al * dep;
      m2 += T * al * dep * dep;
      float Tn = T * (1.0f - al);
      if (T >= 0.5f && Tn < 0.5f) {
        uint32_t hp = rdu(NODE(hn[i]), 15);
        first_part = (float)hp;
        near = dep < 0.05f ? 1.0f : 0.0f;
        if (hn[i] < G && vgrid_at(vgrid, hn[i], xf_dir_inv(PART(hp), V3(-w.x, -w.y, -w.z))) < 0.1f) near += 2.0f;
      }
      T = Tn;
    }
    if (ENV && rdu(ENV, 97) != 0u && t1 >= tmax * 0.999f) acc = add3(acc, mul3s(sky_radiance(ENV, w), T));
    m1 += T * t1;
    m2 += T * t1 * t1;
    float *r0 = probe_data + 8 * (size_t)id, *r1 = r0 + 4;
    float z1 = r1[2];
    if (p >= np && cell >= 1u && cell <= 3u) z1 = cell == 1u ? o.x : (cell == 2u ? o.y : o.z);
    r0[0] = mixf(r0[0], acc.x, alpha); r0[1] = mixf(r0[1], acc.y, alpha); r0[2] = mixf(r0[2], acc.z, alpha);
    r0[3] = first_part;
    r1[0] = mixf(r1[0], m1, alpha); r1[1] = mixf(r1[1], m2, alpha);
    r1[2] = z1; r1[3] = near;
  }
}

/* bounce.wgsl:1417-1445 */
void or_probe_valid(const void *U, const void *PP, const void *parts_, float *probe_data) {
  const void *parts = parts_;
  uint32_t np = lat_count(PP), total = np + rdu(PP, 18), pc = rdu(U, 2);
  for (uint32_t p = 0; p < total; p++) {
    float *at = probe_data + 8 * ((size_t)p * NCELL) + 4;
    if (p >= np) { at[2] = 1.0f; continue; }
    uint32_t cnt[17] = {0};
    uint32_t near = 0, back = 0;
    for (uint32_t c = 0; c < NCELL; c++) {
      const float *r0 = probe_data + 8 * ((size_t)p * NCELL + c), *r1 = r0 + 4;
      uint32_t fl = (uint32_t)(fmaxf(r1[3], 0.0f) + 0.5f);
      near += fl & 1u;
      back += (fl >> 1u) & 1u;
      if (r0[3] >= 0.0f) cnt[(uint32_t)r0[3] < 16u ? (uint32_t)r0[3] : 16u] += 1u;
    }
    int enclosed = 0;
    for (uint32_t q = 0; q < (pc < 16u ? pc : 16u); q++)
      if ((rdu(PART(q), 31) & 1u) == 0u && (float)cnt[q] > rdf(PP, 15) * (float)NCELL) enclosed = 1;
    at[2] = (enclosed || near * 2u > NCELL || back * 2u > NCELL) ? 0.0f : 1.0f;
  }
}

/* bounce.wgsl:1028-1041 */
void or_probe_sh2(const void *PP, const float *probe_data, const float *cells, float *probe_sh) {
  uint32_t total = lat_count(PP) + rdu(PP, 18);
  for (uint32_t p = 0; p < total; p++) {
    v3 c9[9];
    for (int m = 0; m < 9; m++) c9[m] = V3(0.0f, 0.0f, 0.0f);
    for (uint32_t c = 0; c < NCELL; c++) {
      v3 L = mul3s(rd3(probe_data, 8 * ((size_t)p * NCELL + c)), cells[4 * c + 3]);
      float y[9];
      sh2_basis(V3(cells[4 * c], cells[4 * c + 1], cells[4 * c + 2]), y);
      for (int m = 0; m < 9; m++) c9[m] = add3(c9[m], mul3s(L, y[m]));
    }
    for (int m = 0; m < 9; m++) {
      float *o = probe_sh + 4 * (9 * (size_t)p + m);
      o[0] = c9[m].x; o[1] = c9[m].y; o[2] = c9[m].z; o[3] = 0.0f;
    }
  }
}

/* bounce.wgsl:993-1026. The near-field grid is the one-entry placeholder the reference binds while near field has
 * never been on (src/lighting.js:731-733, zeros): near_byte reads it for Gaussian 0 (0) and is 1 for the rest. */
void or_iso_xfer(const void *U, const void *PP, const void *gaussians, const void *nodes, const uint8_t *vgrid,
                 const void *parts_, const float *iso_norm, const float *cells, float *xfer) {
  const void *parts = parts_;
  uint32_t n = rdu(U, 1), pc = rdu(U, 2);
  if ((rdu(U, 15) & 3u) != 3u) return;
#pragma omp parallel for schedule(dynamic, 256)
  for (uint32_t g = 0; g < n; g++) {
    const char *G = GAUSS(g);
    uint32_t own = rdu(G, 11) & 0xFFFFu;
    const char *xf = PART(own);
    v3 P = xf_point(xf, rd3(G, 0));
    int stale = rdu(PP, 13) != 0u || (rdu(xf, 31) & 2u) != 0u;
    for (uint32_t q = 0; q < pc && !stale; q++) {
      if (q == own || (rdu(PART(q), 31) & 2u) == 0u) continue;
      const char *pq = PART(q);
      v3 bc = rd3(pq, 24);
      float r = rdf(pq, 27) + rdf(PP, 11);
      stale = len3(sub3(P, xf_point(pq, bc))) < r || len3(sub3(P, xf_point(pq + 48, bc))) < r;
    }
    if (!stale) continue;
    const char *nd = NODE(g);
    float s3 = rdf(G, 8) * rdf(G, 9) * rdf(G, 10);
    float inorm = iso_norm[g];
    float t[9] = {0};
    for (uint32_t c = 0; c < NCELL; c++) {
      v3 w = V3(cells[4 * c], cells[4 * c + 1], cells[4 * c + 2]);
      v3 u = xf_dir_inv(xf, w);
      float v = vgrid_at(vgrid, g, u) * (g == 0u ? 0.0f : 1.0f);
      if (v <= 0.0f) continue;
      float a = inorm * footprint(nd, s3 * s3, u) * v * cells[4 * c + 3];
      float y[9];
      sh2_basis(w, y);
      for (int m = 0; m < 9; m++) t[m] += a * y[m];
    }
    float *o = xfer + 12 * (size_t)g;
    for (int m = 0; m < 9; m++) o[m] = t[m];
    o[9] = o[10] = o[11] = 0.0f;
  }
}

/* ---- probes -> Gaussian (bounce.wgsl:756-885, 1043-1095) ------------------------------------------------------ */
typedef struct { uint32_t pid[8]; float w[8]; } PW;

static float probe_cheb(const Ctx *c, const float *probe_data, uint32_t p, v3 P) {   /* bounce.wgsl:759-768 */
  v3 v = sub3(P, probe_pos(c, p));
  float dist = len3(v);
  float dm = fmaxf(dist, 1e-9f);
  const float *mm = probe_data + 8 * ((size_t)p * NCELL + cell_of(V3(v.x / dm, v.y / dm, v.z / dm))) + 4;
  float var = fmaxf(mm[1] - mm[0] * mm[0], 1e-4f);
  float dd = fmaxf(dist - 0.03f - mm[0], 0.0f);
  if (dd <= 0.0f) return 1.0f;
  return powf(var / (var + dd * dd), 3.0f);
}

static PW probe_fallback(const Ctx *c, const float *probe_data, v3 P, const uint32_t fi[3]) {   /* bounce.wgsl:771-796 */
  const void *PP = c->PP;
  int dm[3] = {(int)rdu(PP, 208), (int)rdu(PP, 209), (int)rdu(PP, 210)};
  PW out;
  for (int k = 0; k < 8; k++) { out.pid[k] = 0; out.w[k] = 0.0f; }
  for (uint32_t j = 0; j < 64u; j++) {
    int ii[3] = {(int)fi[0] + (int)(j >> 4u) - 1, (int)fi[1] + (int)((j >> 2u) & 3u) - 1, (int)fi[2] + (int)(j & 3u) - 1};
    if (ii[0] < 0 || ii[1] < 0 || ii[2] < 0 || ii[0] >= dm[0] || ii[1] >= dm[1] || ii[2] >= dm[2]) continue;
    uint32_t p = lat_index(PP, (uint32_t)ii[0], (uint32_t)ii[1], (uint32_t)ii[2]);
    if (probe_data[8 * ((size_t)p * NCELL) + 6] <= 0.0f) continue;
    float d = fmaxf(len3(sub3(P, probe_pos(c, p))), 0.05f);
    float w = fmaxf(probe_cheb(c, probe_data, p, P), 1e-3f) / (d * d);
    if (w <= out.w[7]) continue;
    uint32_t at = 7;
    while (at > 0u && out.w[at - 1u] < w) { out.w[at] = out.w[at - 1u]; out.pid[at] = out.pid[at - 1u]; at -= 1u; }
    out.w[at] = w; out.pid[at] = p;
  }
  float wsum = 0.0f;
  for (int k = 0; k < 8; k++) wsum += out.w[k];
  for (int k = 0; k < 8; k++) out.w[k] /= fmaxf(wsum, 1e-12f);
  return out;
}

static PW probe_weights(const Ctx *c, const float *probe_data, v3 P, uint32_t own) {   /* bounce.wgsl:813-885 */
  const void *PP = c->PP, *parts = c->parts;
  PW out;
  if (is_part_probe_part(c, own)) {
    Frame pf = part_probe_frame(c, own);
    v3 lp = xf_point_inv(PART(own), P);
    v3 den = max3s(mul3s(pf.h, 2.0f), 1e-6f);
    v3 f = clamp3(div3(add3(sub3(lp, pf.c), pf.h), den), 0.0f, 1.0f);
    for (uint32_t k = 0; k < 8u; k++) {
      uint32_t o[3] = {k >> 2u, (k >> 1u) & 1u, k & 1u};
      float t0 = o[0] ? f.x : 1.0f - f.x, t1 = o[1] ? f.y : 1.0f - f.y, t2 = o[2] ? f.z : 1.0f - f.z;
      out.pid[k] = lat_count(PP) + own * 8u + k;
      out.w[k] = t0 * t1 * t2;
    }
    return out;
  }
  if (has_vol(c, own)) {
    Frame vf = vol_frame(c, own);
    v3 lp = xf_point_inv(PART(own), P);
    uint32_t d[3] = {rdu(PP, 1252 + 4 * own + 1), rdu(PP, 1252 + 4 * own + 2), rdu(PP, 1252 + 4 * own + 3)};
    v3 den = max3s(mul3s(vf.h, 2.0f), 1e-6f);
    v3 f = clamp3(div3(add3(sub3(lp, vf.c), vf.h), den), 0.0f, 1.0f);
    v3 fq = mul3(f, V3((float)d[0] - 1.0f, (float)d[1] - 1.0f, (float)d[2] - 1.0f));
    float fqa[3] = {fq.x, fq.y, fq.z};
    uint32_t fi[3];
    float fr[3];
    for (int a = 0; a < 3; a++) {
      uint32_t fl = (uint32_t)floorf(fqa[a]);
      fi[a] = fl < d[a] - 2u ? fl : d[a] - 2u;
      fr[a] = fqa[a] - (float)fi[a];
    }
    for (uint32_t k = 0; k < 8u; k++) {
      uint32_t o[3] = {k >> 2u, (k >> 1u) & 1u, k & 1u};
      uint32_t ii[3] = {fi[0] + o[0], fi[1] + o[1], fi[2] + o[2]};
      out.pid[k] = lat_count(PP) + rdu(PP, 1252 + 4 * own) + (ii[0] * d[1] + ii[1]) * d[2] + ii[2];
      float t[3];
      for (int a = 0; a < 3; a++) t[a] = o[a] ? fr[a] : 1.0f - fr[a];
      out.w[k] = t[0] * t[1] * t[2];
    }
    return out;
  }
  /* lat_cell (bounce.wgsl:798-811) */
  uint32_t dm[3] = {rdu(PP, 208), rdu(PP, 209), rdu(PP, 210)};
  float fx = clampf_((P.x - rdf(PP, 0)) / rdf(PP, 4), 0.0f, (float)(dm[0] - 1u));
  float fy = clampf_((P.y - rdf(PP, 1)) / rdf(PP, 5), 0.0f, (float)(dm[1] - 1u));
  uint32_t ix = (uint32_t)floorf(fx) < dm[0] - 2u ? (uint32_t)floorf(fx) : dm[0] - 2u;
  uint32_t iy = (uint32_t)floorf(fy) < dm[1] - 2u ? (uint32_t)floorf(fy) : dm[1] - 2u;
  float tx = fx - (float)ix, ty = fy - (float)iy;
  float gnd = mixf(mixf(lat_ground(PP, ix, iy), lat_ground(PP, ix, iy + 1u), ty),
                   mixf(lat_ground(PP, ix + 1u, iy), lat_ground(PP, ix + 1u, iy + 1u), ty), tx);
  float hgt = P.z - gnd;
  uint32_t kz = 0;
  while (kz + 2u < dm[2] && hgt > lat_level(PP, kz + 1u)) kz += 1u;
  float tz = clampf_((hgt - lat_level(PP, kz)) / fmaxf(lat_level(PP, kz + 1u) - lat_level(PP, kz), 1e-6f), 0.0f, 1.0f);
  uint32_t fi[3] = {ix, iy, kz};
  float fr[3] = {tx, ty, tz};
  float wsum = 0.0f;
  for (uint32_t k = 0; k < 8u; k++) {
    uint32_t o[3] = {k >> 2u, (k >> 1u) & 1u, k & 1u};
    uint32_t p = lat_index(PP, fi[0] + o[0], fi[1] + o[1], fi[2] + o[2]);
    out.pid[k] = p;
    float t[3];
    for (int a = 0; a < 3; a++) t[a] = o[a] ? fr[a] : 1.0f - fr[a];
    float valid = probe_data[8 * ((size_t)p * NCELL) + 6];
    float cw = t[0] * t[1] * t[2] * valid;
    out.w[k] = cw * fmaxf(probe_cheb(c, probe_data, p, P), 1e-3f);
    wsum += out.w[k];
  }
  if (wsum <= 0.0f) return probe_fallback(c, probe_data, P, fi);
  for (int k = 0; k < 8; k++) out.w[k] /= fmaxf(wsum, 1e-12f);
  return out;
}

void or_probe_weights(const void *U, const void *PP, const void *nodes, const void *parts_, const float *cells,
                      const float *probe_data, float px, float py, float pz, uint32_t own, uint32_t *pid, float *w) {
  Ctx cx = {PP, U, parts_, nodes, cells};
  PW r = probe_weights(&cx, probe_data, V3(px, py, pz), own);
  memcpy(pid, r.pid, sizeof r.pid);
  memcpy(w, r.w, sizeof r.w);
}

/* bounce.wgsl:1043-1095 for matte materials (spec_weight 0: the materials here carry no specular flakes) */
void or_bounce_sh(const void *U, const void *PP, const void *gaussians, const void *nodes, const void *parts_,
                  const float *cells, const float *probe_data, const float *probe_sh, const float *xfer, float *lo_b) {
  Ctx cx = {PP, U, parts_, nodes, cells};
  const Ctx *c = &cx;
  const void *parts = parts_;
  uint32_t n = rdu(U, 1);
  if ((rdu(U, 15) & 3u) != 3u) return;
#pragma omp parallel for schedule(dynamic, 512)
  for (uint32_t gid = 0; gid < n; gid++) {
    uint32_t g = rdu(PP, 19) != 0u ? gid : gid * rdu(PP, 16) + rdu(PP, 17);
    if (g >= n) continue;
    const char *G = GAUSS(g);
    uint32_t own = rdu(G, 11) & 0xFFFFu;
    v3 P = xf_point(PART(own), rd3(G, 0));
    PW pw = probe_weights(c, probe_data, P, own);
    v3 Lsh[9];
    for (int m = 0; m < 9; m++) Lsh[m] = V3(0.0f, 0.0f, 0.0f);
    for (int k = 0; k < 8; k++) {
      float w = pw.w[k];
      if (w <= 0.0f) continue;
      for (int m = 0; m < 9; m++) Lsh[m] = add3(Lsh[m], mul3s(rd3(probe_sh, 4 * (9 * (size_t)pw.pid[k] + m)), w));
    }
    const float *t = xfer + 12 * (size_t)g;
    v3 s = mul3s(Lsh[0], t[0]);
    for (int m = 1; m < 9; m++) s = add3(s, mul3s(Lsh[m], t[m]));
    v3 alb = rd3(G, 12);
    v3 cb = mul3(alb, max3s(s, 0.0f));
    float *o = lo_b + 12 * (size_t)g;
    o[0] = cb.x; o[1] = cb.y; o[2] = cb.z; o[3] = 0.0f;
    o[4] = cb.x; o[5] = cb.y; o[6] = cb.z; o[7] = 0.0f;
    o[8] = o[9] = o[10] = o[11] = 0.0f;
  }
}

/* probe positions (bounce.wgsl probe_pos), for the host's checks */
void or_probe_pos(const void *U, const void *PP, const void *nodes, const void *parts_, uint32_t count, float *out) {
  Ctx cx = {PP, U, parts_, nodes, NULL};
  for (uint32_t p = 0; p < count; p++) {
    v3 o = probe_pos(&cx, p);
    out[3 * (size_t)p] = o.x; out[3 * (size_t)p + 1] = o.y; out[3 * (size_t)p + 2] = o.z;
  }
}
