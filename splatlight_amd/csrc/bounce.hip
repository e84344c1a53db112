// Sky and bounce through probes on gfx950 (runtime/wgsl/bounce.wgsl, default path: no sender LOD, no near field,
// M39 matte with SH2 probes), plus the passes that feed it (emissive.wgsl direct_dirs_emit, environment.wgsl
// env_dirs). Per frame: the senders' light toward the probes (lo_dir), the tree's node sums (node_leaf, node_up per
// depth, node_sh2), one ray per (probe, cell) walking each part's tree with a theta cut and compositing the hit
// nodes' light front to back with the sky at escape (probe_gather), the probes' validity and SH2 (probe_valid,
// probe_sh2), and per Gaussian the SH2 transfer (iso_xfer, when stale) and the bounce from the 8 surrounding probes
// (bounce_sh) into lo_b, which the frame's colour pass adds with the bounce gain.
//
// One lane per item in every pass. Contraction is off, so every float expression rounds as written in WGSL order;
// the cell directions and solid angles come from a table the host computes in double and rounds once.
#pragma clang fp contract(off)

namespace bn {
using namespace lt;

constexpr uint32_t NCELL = 128u;
constexpr uint32_t HIT_MAX = 16u;   // bounce.wgsl:21
constexpr uint32_t STACK = 64u;     // common.wgsl:72
constexpr uint32_t PPK = 8u;        // probes per movable part (bounce.wgsl:251)
constexpr uint32_t NONE = 0xFFFFFFFFu;

struct ProbeParamsDev {             // bounce.wgsl:23-56 (9632 B, src/lighting.js:173)
  float lo[3]; uint32_t n_axis;
  float step[3]; uint32_t node_count;
  float block_rad, theta, tmax, near_r;
  float alpha; uint32_t force_near; float vfloor, g_valid;
  uint32_t subset, phase, part_probes, sh_full;
  float dirs[14][4];
  float pinv[32][4];                // 9 x 14, row-major in the first 126 floats
  uint32_t refl[4];
  uint32_t dims[4];
  float levels[16];
  float ground[1024];
  uint32_t vols[32][4];
  uint32_t vol_info[4];
  float nodep[256][4];
};
static_assert(sizeof(ProbeParamsDev) == 9632, "ProbeParams");

__device__ __forceinline__ float atan2_cr(float y, float x) { return (float)atan2((double)y, (double)x); }
__device__ __forceinline__ v3 xyz(float4 a) { return V3(a.x, a.y, a.z); }
__device__ __forceinline__ v3 neg3(v3 a) { return V3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ v3 mul3(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ v3 div3(v3 a, v3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ v3 max3s(v3 a, float s) { return V3(fmaxf(a.x, s), fmaxf(a.y, s), fmaxf(a.z, s)); }
__device__ __forceinline__ v3 min3(v3 a, v3 b) { return V3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
__device__ __forceinline__ float clampb(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
__device__ __forceinline__ v3 clamp3(v3 a, float lo, float hi) { return V3(clampb(a.x, lo, hi), clampb(a.y, lo, hi), clampb(a.z, lo, hi)); }
__device__ __forceinline__ v3 xf_point_inv(const Xf &x, v3 p) { return xf_dir_inv(x, sub3(p, x.t)); }
__device__ __forceinline__ Xf load_xf_prev(const float4 *parts, uint32_t p) {
  const float4 a = parts[p * 8 + 3], b = parts[p * 8 + 4], c = parts[p * 8 + 5];
  return {V3(a.x, a.y, a.z), V3(b.x, b.y, b.z), V3(c.x, c.y, c.z), V3(a.w, b.w, c.w)};
}
// Part words: bound 24-27 (float4 6), root 28, g_lo 29, g_hi 30, flags 31 (float4 7)
__device__ __forceinline__ float4 part_bound(const float4 *parts, uint32_t q) { return parts[q * 8 + 6]; }
__device__ __forceinline__ uint32_t part_root(const float4 *parts, uint32_t q) { return __float_as_uint(parts[q * 8 + 7].x); }
__device__ __forceinline__ uint32_t part_flags(const float4 *parts, uint32_t q) { return __float_as_uint(parts[q * 8 + 7].w); }

__device__ __forceinline__ uint32_t lat_count(const ProbeParamsDev *PP) { return PP->dims[0] * PP->dims[1] * PP->dims[2]; }
__device__ __forceinline__ uint32_t lat_index(const ProbeParamsDev *PP, uint32_t i, uint32_t j, uint32_t k) {
  return (i * PP->dims[1] + j) * PP->dims[2] + k;
}
__device__ __forceinline__ float lat_level(const ProbeParamsDev *PP, uint32_t k) { return PP->levels[k]; }
__device__ __forceinline__ float lat_ground(const ProbeParamsDev *PP, uint32_t i, uint32_t j) { return PP->ground[i * PP->dims[1] + j]; }

__device__ __forceinline__ uint32_t vol_of_probe(const ProbeParamsDev *PP, uint32_t part_count, uint32_t r) {   // bounce.wgsl:61-68
  for (uint32_t q = 0; q < min(part_count, 32u); q++) {
    const uint32_t *v = PP->vols[q];
    if (v[1] > 0u && r >= v[0] && r < v[0] + v[1] * v[2] * v[3]) return q;
  }
  return NONE;
}
__device__ __forceinline__ bool has_vol(const ProbeParamsDev *PP, uint32_t q) { return q < 32u && PP->vols[q][1] > 0u; }

struct PFrame { v3 c, h; };
// bounce.wgsl:256-262: the part root's centre and 0.95 sigma per axis, clamped to half the bound
__device__ __forceinline__ PFrame part_probe_frame(const float4 *nodes, const float4 *parts, uint32_t q) {
  const float4 *n = nodes + 6 * (size_t)part_root(parts, q);
  const float4 n0 = n[0], n1 = n[1], n2 = n[2];
  const v3 a = xyz(n1), o = xyz(n2);
  const float det = (a.x * (a.y * a.z - o.z * o.z) - o.x * (o.x * a.z - o.z * o.y)) + o.y * (o.x * o.z - a.y * o.y);
  const float dd = fmaxf(det, 1e-30f);
  const v3 var = V3((a.y * a.z - o.z * o.z) / dd, (a.x * a.z - o.y * o.y) / dd, (a.x * a.y - o.x * o.x) / dd);
  const v3 s = V3(sqrtf(fmaxf(var.x, 0.0f)), sqrtf(fmaxf(var.y, 0.0f)), sqrtf(fmaxf(var.z, 0.0f)));
  const float hb = 0.5f * part_bound(parts, q).w;
  return {xyz(n0), min3(mul3s(s, 0.95f), V3(hb, hb, hb))};
}
__device__ __forceinline__ PFrame vol_frame(const float4 *nodes, const float4 *parts, uint32_t q) {   // bounce.wgsl:70-73
  PFrame f = part_probe_frame(nodes, parts, q);
  const float b = part_bound(parts, q).w;
  f.h = min3(mul3s(f.h, 1.6f), V3(b, b, b));
  return f;
}
__device__ __forceinline__ bool is_part_probe_part(const ProbeParamsDev *PP, const float4 *parts, uint32_t q) {
  return PP->part_probes > 0u && q < PP->vol_info[0] / PPK && (part_flags(parts, q) & 1u) == 0u;
}
// bounce.wgsl:217-244
__device__ v3 probe_pos(const ProbeParamsDev *PP, uint32_t part_count, const float4 *nodes, const float4 *parts,
                        uint32_t p) {
  const uint32_t nl = lat_count(PP);
  if (p >= nl + PP->vol_info[0]) {
    const uint32_t r = p - nl, q = vol_of_probe(PP, part_count, r);
    if (q == NONE) return V3(0.0f, 0.0f, 0.0f);
    const uint32_t dx = PP->vols[q][1], dy = PP->vols[q][2], dz = PP->vols[q][3];
    const uint32_t l = r - PP->vols[q][0];
    const v3 ijk = V3((float)(l / (dy * dz)), (float)((l / dz) % dy), (float)(l % dz));
    const PFrame vf = vol_frame(nodes, parts, q);
    const v3 den = V3(fmaxf((float)dx - 1.0f, 1.0f), fmaxf((float)dy - 1.0f, 1.0f), fmaxf((float)dz - 1.0f, 1.0f));
    return xf_point(load_xf(parts, q), add3(sub3(vf.c, vf.h), div3(mul3(mul3s(vf.h, 2.0f), ijk), den)));
  }
  if (p >= nl) {
    const uint32_t q = (p - nl) / PPK, k = (p - nl) % PPK;
    const v3 sgn = V3((float)(k >> 2u) * 2.0f - 1.0f, (float)((k >> 1u) & 1u) * 2.0f - 1.0f, (float)(k & 1u) * 2.0f - 1.0f);
    const PFrame pf = part_probe_frame(nodes, parts, q);
    return xf_point(load_xf(parts, q), add3(pf.c, mul3(sgn, pf.h)));
  }
  const uint32_t dy = PP->dims[1], dz = PP->dims[2];
  const uint32_t i = p / (dy * dz), j = (p / dz) % dy, k = p % dz;
  return V3(PP->lo[0] + PP->step[0] * (float)i, PP->lo[1] + PP->step[1] * (float)j, lat_ground(PP, i, j) + lat_level(PP, k));
}

// bounce.wgsl:199-204, with asin and atan2 defined as double, rounded once. The float functions (a few ulp) move the
// row and column coordinates by under 1e-5, so where a coordinate lies more than 1e-4 from a cell boundary its
// floor is the double definition's; the double functions run only near a boundary (or for NaN).
__device__ __forceinline__ uint32_t cell_of(v3 w) {
  const float z = clampb(w.z, -1.0f, 1.0f);
  float fi = fmaxf((0.5f * PI_F - asinf(z)) / (PI_F / 8.0f), 0.0f);
  const float ri = fi - floorf(fi);
  if (!(ri > 1e-4f && ri < 1.0f - 1e-4f)) fi = fmaxf((0.5f * PI_F - asin_cr(z)) / (PI_F / 8.0f), 0.0f);
  const uint32_t i = min((uint32_t)fi, 7u);
  float fj = (atan2f(w.y, w.x) + PI_F) / (2.0f * PI_F) * 16.0f;
  const float rj = fj - floorf(fj);
  if (!(rj > 1e-4f && rj < 1.0f - 1e-4f)) fj = (atan2_cr(w.y, w.x) + PI_F) / (2.0f * PI_F) * 16.0f;
  const uint32_t j = ((uint32_t)floorf(fj)) % 16u;
  return i * 16u + j;
}
__device__ __forceinline__ void sh2_basis(v3 d, float b[9]) {   // bounce.wgsl:206-210
  const float x = d.x, y = d.y, z = d.z;
  b[0] = 0.282095f; b[1] = 0.488603f * y; b[2] = 0.488603f * z; b[3] = 0.488603f * x;
  b[4] = 1.092548f * x * y; b[5] = 1.092548f * y * z; b[6] = 0.315392f * (3.0f * z * z - 1.0f);
  b[7] = 1.092548f * x * z; b[8] = 0.546274f * (x * x - y * y);
}
__device__ __forceinline__ v3 sky_radiance(const EnvDev &ENV, v3 w) {   // bounce.wgsl:148-158
  const float x = w.x, y = w.y, z = w.z;
  const float b[16] = {0.282095f, 0.488603f * y, 0.488603f * z, 0.488603f * x,
                       1.092548f * x * y, 1.092548f * y * z, 0.315392f * (3.0f * z * z - 1.0f), 1.092548f * x * z,
                       0.546274f * (x * x - y * y), 0.590044f * y * (3.0f * x * x - y * y), 2.890611f * x * y * z,
                       0.457046f * y * (5.0f * z * z - 1.0f), 0.373176f * z * (5.0f * z * z - 3.0f),
                       0.457046f * x * (5.0f * z * z - 1.0f), 1.445306f * z * (x * x - y * y),
                       0.590044f * x * (x * x - 3.0f * y * y)};
  v3 c = V3(0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int k = 0; k < 16; k++) c = add3(c, mul3s(V3(ENV.sh[k][0], ENV.sh[k][1], ENV.sh[k][2]), b[k]));
  return max3s(c, 0.0f);
}
__device__ __forceinline__ float seg_geom(v3 mean, const NodeL &nd, v3 P, v3 w, float t0, float t1) {   // common.wgsl:133-152
  const v3 d = sub3(mean, P);
  const v3 cw = sym_mul(nd.cinv_d, nd.cinv_o, w);
  const float a = dot3(cw, w);
  const float m = dot3(cw, d) / a;
  const v3 e = sub3(d, mul3s(w, m));
  const float q = fmaxf(dot3(e, sym_mul(nd.cinv_d, nd.cinv_o, e)), 0.0f);
  const float s = sqrtf(0.5f * a);
  const float x0 = s * (t0 - m), x1 = s * (t1 - m);
  float df;
  if (x0 > 0.0f) df = erfc_pos(x0) - erfc_pos(x1);
  else if (x1 < 0.0f) df = erfc_pos(-x1) - erfc_pos(-x0);
  else df = (2.0f - erfc_pos(x1)) - erfc_pos(-x0);
  return expf(-0.5f * q) * sqrtf(PI_F / (2.0f * a)) * df;
}
__device__ __forceinline__ float s3_of(float4 S) { return S.x * S.y * S.z; }

}  // namespace bn

using namespace bn;

// ---- lo_dir: direct_dirs_emit (emissive.wgsl:1012-1016, 1067-1083) + env_dirs (environment.wgsl:311-359) ----------
// A matte receiver's emitter light is the same toward each of the 14 directions (emissive.wgsl:930-937): c0, the
// shading's sum before an emitter's own radiance, saved by the shading pass. One value toward the probes: its mean
// weighted by footprint x V over DIRS14; with U.iso bit 1 a quarter of the Gaussians (g % 4 == frame % 4) per frame,
// to which env_dirs adds the distant lights' light.
__constant__ float kDirs14[14][3] = {
    {1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {1e-4f, 0.0f, 1.0f}, {-1.0f, 0.0f, 0.0f}, {0.0f, -1.0f, 0.0f},
    {1e-4f, 0.0f, -1.0f}, {0.57735f, 0.57735f, 0.57735f}, {0.57735f, 0.57735f, -0.57735f},
    {0.57735f, -0.57735f, 0.57735f}, {0.57735f, -0.57735f, -0.57735f}, {-0.57735f, 0.57735f, 0.57735f},
    {-0.57735f, 0.57735f, -0.57735f}, {-0.57735f, -0.57735f, 0.57735f}, {-0.57735f, -0.57735f, -0.57735f}};

extern "C" __global__ void __launch_bounds__(64)
sl_bounce_lo_dir(FrameUniformsDev U, EnvDev ENV, const float4 *__restrict__ g1, const float4 *__restrict__ g2,
                 const float4 *__restrict__ g3, uint32_t st, const float4 *__restrict__ nodes,
                 const uint8_t *__restrict__ vgrid, const float4 *__restrict__ parts,
                 const uint32_t *__restrict__ materials, const float *__restrict__ iso_norm,
                 const float4 *__restrict__ env_T, const float4 *__restrict__ c0, float4 *__restrict__ lo_dir) {
  const uint32_t t = blockIdx.x * 64u + threadIdx.x;
  const uint32_t g = (U.iso & 2u) != 0u ? t * 4u + U.frame % 4u : t;
  if (g >= U.gaussian_count) return;
  const float4 S = sl_scale(g2, st, g), C = sl_albedo(g1, g2, g3, st, g);
  const NodeL nd = load_node(nodes, g);
  const Xf xf = load_xf(parts, __float_as_uint(S.w) & 0xFFFFu);
  const float s3 = s3_of(S);
  const uint8_t *vg = vgrid + 128 * (size_t)g;
  const float4 cc = c0[g];
  const v3 c = V3(cc.x, cc.y, cc.z);
  v3 mean = V3(0.0f, 0.0f, 0.0f);
  float ws = 0.0f;
  for (int k = 0; k < 14; k++) {
    const v3 u = xf_dir_inv(xf, V3(kDirs14[k][0], kDirs14[k][1], kDirs14[k][2]));
    const float w = footprint(nd, s3 * s3, u) * vgrid_lookup(vg, u);
    mean = add3(mean, mul3s(c, w));
    ws += w;
  }
  v3 lo = c;
  if (ws > 0.0f) {
    const float d = fmaxf(ws, 1e-20f);
    lo = V3(mean.x / d, mean.y / d, mean.z / d);
  }
  const uint32_t mi = __float_as_uint(C.w) & 0xFFFFu, mmax = max(U.material_count, 1u) - 1u;
  const bool emissive = U.material_count && (materials[8 * (size_t)min(mi, mmax) + 5] & 1u);
  if (ENV.info[0] > 0u && (U.iso & 1u) != 0u && !emissive) {
    const float4 Tg = env_T ? env_T[g] : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    const float Tk[4] = {Tg.x, Tg.y, Tg.z, Tg.w};
    v3 acc = V3(0.0f, 0.0f, 0.0f);
    for (uint32_t k = 0; k < min(ENV.info[0], 4u); k++) {
      const v3 wi = xf_dir_inv(xf, V3(ENV.lights[k][0], ENV.lights[k][1], ENV.lights[k][2]));
      const float v = vgrid_lookup(vg, wi);
      if (v <= 1e-7f) continue;
      const float f = Tk[k] * v * footprint(nd, s3 * s3, wi);
      acc = add3(acc, mul3s(V3(ENV.lights[k][4], ENV.lights[k][5], ENV.lights[k][6]), f));
    }
    const float inorm = iso_norm[g];
    lo = add3(lo, V3(C.x * acc.x * inorm, C.y * acc.y * inorm, C.z * acc.z * inorm));
  }
  lo_dir[g] = make_float4(lo.x, lo.y, lo.z, 0.0f);
}

// ---- node emitters (bounce.wgsl:1166-1233) ------------------------------------------------------------------------
namespace bn {
// node_sh2_one (bounce.wgsl:1217-1233): a node's 14 sums (sum w L, sum w) -> the mean light per direction -> SH2 by
// PP.pinv -> 27 halves (round to nearest even) in 14 words
__device__ __forceinline__ void sh2_pack(const ProbeParamsDev *PP, const float4 (&a)[14], uint32_t *o) {
  v3 L[14];
#pragma unroll
  for (int k = 0; k < 14; k++) {
    const float d = fmaxf(a[k].w, 1e-12f);
    L[k] = V3(a[k].x / d, a[k].y / d, a[k].z / d);
  }
  float h[28];
  h[27] = 0.0f;
  const float *pinv = &PP->pinv[0][0];
#pragma unroll
  for (int m = 0; m < 9; m++) {
    v3 c = V3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < 14; k++) c = add3(c, mul3s(L[k], pinv[m * 14 + k]));
    h[3 * m] = c.x; h[3 * m + 1] = c.y; h[3 * m + 2] = c.z;
  }
#pragma unroll
  for (int j = 0; j < 14; j++) {
    const uint32_t lo = __half_as_ushort(__float2half_rn(h[2 * j])), hi = __half_as_ushort(__float2half_rn(h[2 * j + 1]));
    o[j] = lo | (hi << 16);
  }
}
}  // namespace bn

// The leaves' weights toward the 14 directions, A(D_k) V(D_k) (node_leaf's w): they change only with the part's
// pose, so they are kept per leaf (14 floats) and rebuilt when the part table changes.
extern "C" __global__ void __launch_bounds__(64)
sl_leaf_weights(uint32_t G, const ProbeParamsDev *__restrict__ PP, const float4 *__restrict__ g2, uint32_t st,
                const float4 *__restrict__ nodes, const uint8_t *__restrict__ vgrid, const float4 *__restrict__ parts,
                float *__restrict__ wcache) {
  const uint32_t g = blockIdx.x * 64u + threadIdx.x;
  if (g >= G) return;
  const float4 S = sl_scale(g2, st, g);
  const NodeL nd = load_node(nodes, g);
  const Xf xf = load_xf(parts, __float_as_uint(S.w) & 0xFFFFu);
  const float s3 = s3_of(S);
  const uint8_t *vg = vgrid + 128 * (size_t)g;
  for (int k = 0; k < 14; k++) {
    const v3 wo = xf_dir_inv(xf, V3(PP->dirs[k][0], PP->dirs[k][1], PP->dirs[k][2]));
    wcache[14 * (size_t)g + k] = footprint(nd, s3 * s3, wo) * vgrid_lookup(vg, wo);
  }
}

// The node sums without leaf rows: node_acc holds the internal nodes' 14 sums direction-major ([k][node], so a
// depth's consecutive nodes read their children's rows coalesced); a leaf's entries (L w_k, w_k) are formed where
// they are read, from its light toward the probes L (lo_dir + last frame's bounce toward the probes; nothing for the
// PP.refl.w part) and its cached weights. The same float operations in the same order as node_leaf / node_up /
// node_sh2, so the same bits.
namespace bn {
__device__ __forceinline__ v3 leaf_light(const ProbeParamsDev *PP, const float4 *g2, uint32_t st, const float4 *lo_dir,
                                         const float4 *lo_b, uint32_t c) {
  const uint32_t dark = PP->refl[3];
  const uint32_t own = __float_as_uint(sl_scale(g2, st, c).w) & 0xFFFFu;
  if (dark != 0u && own == dark - 1u) return V3(0.0f, 0.0f, 0.0f);
  const float4 a = lo_dir[c], b = lo_b[3 * (size_t)c + 1];
  return V3(a.x + b.x, a.y + b.y, a.z + b.z);
}
}  // namespace bn

// the leaves' SH2 (node_sh2 of node_leaf's sums), one lane per leaf
extern "C" __global__ void __launch_bounds__(64)
sl_leaf_sh(uint32_t G, const ProbeParamsDev *__restrict__ PP, const float4 *__restrict__ g2, uint32_t st,
           const float *__restrict__ wcache, const float4 *__restrict__ lo_dir, const float4 *__restrict__ lo_b,
           uint32_t *__restrict__ node_sh) {
  const uint32_t g = blockIdx.x * 64u + threadIdx.x;
  if (g >= G) return;
  const v3 L = leaf_light(PP, g2, st, lo_dir, lo_b, g);
  float4 s[14];
#pragma unroll
  for (int k = 0; k < 14; k++) {
    const float w = wcache[14 * (size_t)g + k];
    s[k] = make_float4(L.x * w, L.y * w, L.z * w, w);
  }
  sh2_pack(PP, s, node_sh + 14 * (size_t)g);
}

// One child's terms toward its parent's 14 sums: a leaf's light times its weights (node_leaf), or an internal node's
// stored sums
__device__ __forceinline__ void kid_terms(uint32_t G, uint32_t N, const ProbeParamsDev *__restrict__ PP,
                                          const float4 *__restrict__ g2, uint32_t st, uint32_t c,
                                          const float *__restrict__ wcache, const float4 *__restrict__ lo_dir,
                                          const float4 *__restrict__ lo_b, const float4 *__restrict__ node_acc,
                                          float4 (&x)[14]) {
  if (c < G) {
    const v3 L = leaf_light(PP, g2, st, lo_dir, lo_b, c);
#pragma unroll
    for (int k = 0; k < 14; k++) {
      const float w = wcache[14 * (size_t)c + k];
      x[k] = make_float4(L.x * w, L.y * w, L.z * w, w);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 14; k++) x[k] = node_acc[(size_t)k * N + c];
  }
}

// node_up + node_sh2 of one entry of a depth's list, e = (node, child count, first child, second child) — with more
// than two children e.z is the node's first index into kids: the node's sums are its children's terms added in child
// order from zero, then stored with its SH2. Two children (the binary trees' case) load both at once.
__device__ __forceinline__ void node_up_ent(uint32_t G, uint32_t N, const ProbeParamsDev *__restrict__ PP,
                                            const float4 *__restrict__ g2, uint32_t st, const uint32_t *__restrict__ kids,
                                            uint4 e, const float *__restrict__ wcache,
                                            const float4 *__restrict__ lo_dir, const float4 *__restrict__ lo_b,
                                            float4 *__restrict__ node_acc, uint32_t *__restrict__ node_sh) {
  const uint32_t i = e.x, nk = e.y;
  float4 s[14];
  if (nk == 2u) {
    float4 x0[14], x1[14];
    kid_terms(G, N, PP, g2, st, e.z, wcache, lo_dir, lo_b, node_acc, x0);
    kid_terms(G, N, PP, g2, st, e.w, wcache, lo_dir, lo_b, node_acc, x1);
#pragma unroll
    for (int k = 0; k < 14; k++)
      s[k] = make_float4((0.0f + x0[k].x) + x1[k].x, (0.0f + x0[k].y) + x1[k].y, (0.0f + x0[k].z) + x1[k].z,
                         (0.0f + x0[k].w) + x1[k].w);
  } else {
#pragma unroll
    for (int k = 0; k < 14; k++) s[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t j = 0; j < nk; j++) {
      float4 x[14];
      kid_terms(G, N, PP, g2, st, nk == 1u ? e.z : kids[e.z + j], wcache, lo_dir, lo_b, node_acc, x);
#pragma unroll
      for (int k = 0; k < 14; k++)
        s[k] = make_float4(s[k].x + x[k].x, s[k].y + x[k].y, s[k].z + x[k].z, s[k].w + x[k].w);
    }
  }
#pragma unroll
  for (int k = 0; k < 14; k++) node_acc[(size_t)k * N + i] = s[k];
  sh2_pack(PP, s, node_sh + 14 * (size_t)i);
}

// node_up + node_sh2 for one depth (entries up_ent[r0, r0 + count)), one lane per node
extern "C" __global__ void __launch_bounds__(64)
sl_node_up_sh(uint32_t G, uint32_t N, const ProbeParamsDev *__restrict__ PP, const float4 *__restrict__ g2, uint32_t st,
              const uint32_t *__restrict__ kids, const uint4 *__restrict__ up_ent, uint32_t r0, uint32_t count,
              const float *__restrict__ wcache, const float4 *__restrict__ lo_dir, const float4 *__restrict__ lo_b,
              float4 *__restrict__ node_acc, uint32_t *__restrict__ node_sh) {
  const uint32_t t = blockIdx.x * 64u + threadIdx.x;
  if (t >= count) return;
  node_up_ent(G, N, PP, g2, st, kids, up_ent[r0 + t], wcache, lo_dir, lo_b, node_acc, node_sh);
}

// The shallow depths (depth d_hi up to 1, each of at most UP_TAIL_MAX nodes) in one workgroup, deepest first, with a
// workgroup barrier between depths: a depth reads the sums its children's depth stored, and all lanes of a workgroup
// share one CU's vector cache, so workgroup-scope ordering makes them visible. One launch in place of one per depth;
// each depth costs about two dependent memory round trips (its entries, then its children's terms).
constexpr uint32_t UP_TAIL_MAX = 2048u, UP_TAIL_DEPTHS = 32u, UP_TAIL_THREADS = 512u;
struct UpTail {
  uint32_t off[UP_TAIL_DEPTHS + 1], cnt[UP_TAIL_DEPTHS + 1];
};
extern "C" __global__ void __launch_bounds__(UP_TAIL_THREADS)
sl_node_up_tail(uint32_t G, uint32_t N, const ProbeParamsDev *__restrict__ PP, const float4 *__restrict__ g2,
                uint32_t st, const uint32_t *__restrict__ kids, const uint4 *__restrict__ up_ent, UpTail T,
                uint32_t d_hi, const float *__restrict__ wcache, const float4 *__restrict__ lo_dir,
                const float4 *__restrict__ lo_b, float4 *__restrict__ node_acc, uint32_t *__restrict__ node_sh) {
  for (uint32_t d = d_hi; d >= 1u; d--) {
    for (uint32_t t = threadIdx.x; t < T.cnt[d]; t += UP_TAIL_THREADS)
      node_up_ent(G, N, PP, g2, st, kids, up_ent[T.off[d] + t], wcache, lo_dir, lo_b, node_acc, node_sh);
    __syncthreads();
  }
}

// ---- probe_gather (bounce.wgsl:1255-1415) ---------------------------------------------------------------------------
// One lane per (probe, cell) ray, rays of a probe in consecutive lanes. Each part's tree is walked depth first with
// the walk's stack in LDS (64 entries per lane, lane-indexed columns); a node is used whole when it is a leaf or its
// distance from the gather centre, less the block radius, exceeds theta x smax, else its children are pushed. Used
// nodes with optical depth above 1e-6 enter a 16-entry list sorted by depth, held in registers; a full list folds
// the farthest hit into its last entry. The list is then composited front to back with each node's SH2 light, the
// sky added for what escapes to tmax. No lamps end a ray (the scenes carry emissive Gaussians, not rect lamps):
// t1 = tmax. alpha in (0, 1]: the new gather blended into the old (1: taken as is).
// SL: the LDS stack's entries per lane. The reference's walk keeps 64 (a push past them is dropped); the host takes
// 32 when the tree cannot fill more than 32 (no push is ever dropped then either), 64 otherwise. A wave takes 64
// directions of one probe: they walk the same part of the tree, which keeps its nodes in cache (a wave of one
// direction over 64 probes walked alike but ran 20 % slower). Workgroups take the rays from the last probe down:
// the part and volume probes, which sit inside objects and walk the longest (a thousand nodes a ray), start first.
// The walk reads a compact copy of each part's tree (built at set-up): nodes renumbered breadth first per part, so a
// node's children are consecutive and in kids order, 32 B a node: (mean, radius), (smax, first child, child count,
// node id). A visit is one 32 B read; only a used node reads its full record (by node id).
// A used node costs some six times a visit (its footprint, segment geometry and the list insertion), and the lanes of
// a wave use and open nodes at different steps, so the walk queues each lane's used nodes (GQ at most, in LDS) and
// the wave then takes them all, each lane in its own walk order: the hit list sees the same insertions in the same
// order, while both kinds of step run with most lanes busy.
constexpr uint32_t GQ = 64;
template <uint32_t SL>
__global__ void __launch_bounds__(64)
sl_probe_gather(FrameUniformsDev U, const ProbeParamsDev *__restrict__ PP, EnvDev ENV, const float4 *__restrict__ nodes,
                const uint8_t *__restrict__ vgrid, const float4 *__restrict__ parts, const float4 *__restrict__ walk,
                const uint32_t *__restrict__ walk_root, const uint32_t *__restrict__ node_sh,
                const float4 *__restrict__ cells, float4 *__restrict__ probe_data, uint32_t n_threads) {
  __shared__ uint32_t s_stack[SL][64];
  __shared__ uint32_t s_queue[GQ][64];
  const uint32_t lane = threadIdx.x;
  const uint32_t t = (gridDim.x - 1u - blockIdx.x) * 64u + lane;
  if (t >= n_threads) return;
  const uint32_t G = U.gaussian_count;
  const uint32_t np = lat_count(PP);
  const uint32_t id = t * PP->subset + PP->phase;
  if (id >= (np + PP->part_probes) * NCELL) return;
  const uint32_t p = id / NCELL, cell = id % NCELL;
  uint32_t skip = NONE;
  if (p >= np && p < np + PP->vol_info[0]) {
    skip = (p - np) / PPK;
    if (!is_part_probe_part(PP, parts, skip)) return;
  } else if (p >= np) {
    skip = vol_of_probe(PP, U.part_count, p - np);
    if (skip == NONE) return;
  }
  const v3 o = probe_pos(PP, U.part_count, nodes, parts, p);
  const v3 w = xyz(cells[cell]);
  const uint32_t dy = PP->dims[1], dz = PP->dims[2];
  const v3 b = V3((float)((p / (dy * dz)) / 2u), (float)(((p / dz) % dy) / 2u), (float)((p % dz) / 2u));
  const bool own_centre = p >= np || PP->dims[3] == 1u;
  const v3 cen = own_centre ? o : add3(V3(PP->lo[0], PP->lo[1], PP->lo[2]),
                                       mul3(V3(PP->step[0], PP->step[1], PP->step[2]), add3(mul3s(b, 2.0f), V3(0.5f, 0.5f, 0.5f))));
  const float t1 = PP->tmax;
  const float theta = PP->theta;
  const float brad = (own_centre && PP->vol_info[1] == 0u) ? 0.0f : PP->block_rad;
  float ht[HIT_MAX], hv[HIT_MAX];
  uint32_t hn[HIT_MAX];
#pragma unroll
  for (int i = 0; i < (int)HIT_MAX; i++) { ht[i] = 0.0f; hv[i] = 0.0f; hn[i] = 0u; }
  uint32_t k = 0;
  for (uint32_t q = 0; q < U.part_count; q++) {
    if (q == skip) continue;
    const Xf xf = load_xf(parts, q);
    const v3 Po = xf_point_inv(xf, o), wr = xf_dir_inv(xf, w), cr = xf_point_inv(xf, cen);
    s_stack[0][lane] = walk_root[q];
    uint32_t sp = 1;
    for (;;) {
      // the walk: until this lane's stack is empty or its queue of used nodes is full
      uint32_t qn = 0;
      while (sp > 0u && qn < GQ) {
        sp -= 1u;
        const uint32_t wi = s_stack[sp][lane];
        const float4 n0 = walk[2 * (size_t)wi];
        const v3 mean = xyz(n0);
        const v3 d = sub3(mean, Po);
        const float tt = clampb(dot3(d, wr), 0.0f, t1);
        const v3 e = sub3(d, mul3s(wr, tt));
        if (dot3(e, e) > n0.w * n0.w) continue;
        const float4 wb = walk[2 * (size_t)wi + 1];
        const uint32_t ni = __float_as_uint(wb.w);
        bool use = ni < G;
        if (!use) use = len3(sub3(cr, mean)) - brad > theta * wb.x;
        if (use) {
          s_queue[qn][lane] = wi;
          qn += 1u;
        } else {
          const uint32_t fc = __float_as_uint(wb.y), nk = __float_as_uint(wb.z);
          for (uint32_t j = 0; j < nk; j++)
            if (sp < SL) { s_stack[sp][lane] = fc + j; sp += 1u; }
        }
      }
      // the used nodes, in walk order
      for (uint32_t u = 0; u < qn; u++) {
        const uint32_t wi = s_queue[u][lane];
        const v3 mean = xyz(walk[2 * (size_t)wi]);
        const v3 d = sub3(mean, Po);
        const uint32_t ni = __float_as_uint(walk[2 * (size_t)wi + 1].w);
        const float4 *n = nodes + 6 * (size_t)ni;
        const float4 n1 = n[1], n2 = n[2], n3 = n[3], n4 = n[4];
        const NodeL nd = {xyz(n1), xyz(n2), xyz(n3), xyz(n4), n1.w};
        const float tau = nd.kappa * dirm(nd, wr) * seg_geom(mean, nd, Po, wr, 0.0f, t1);
        if (tau > 1e-6f) {
          const v3 cw = sym_mul(nd.cinv_d, nd.cinv_o, wr);
          const float den = dot3(cw, wr);
          const float tp = fabsf(den) > 1e-30f ? clampb(dot3(cw, d) / den, 0.0f, t1) : t1;
          if (k == HIT_MAX && tp >= ht[HIT_MAX - 1u]) {
            ht[HIT_MAX - 1u] = fmaxf(ht[HIT_MAX - 1u], tp);
            hv[HIT_MAX - 1u] = fminf(hv[HIT_MAX - 1u] + tau, 64.0f);
            continue;
          }
          const bool full = k == HIT_MAX;
          const float sv = hv[HIT_MAX - 1u], stt = ht[HIT_MAX - 1u];
          const uint32_t top = min(k, HIT_MAX - 1u);
          // the reference's insertion: walk down from `top` while the entry below is deeper than tp
          uint32_t at = top;
#pragma unroll
          for (int j = (int)HIT_MAX - 1; j >= 1; j--)
            if (at == (uint32_t)j && ht[j - 1] > tp) at = (uint32_t)j - 1u;
#pragma unroll
          for (int i = (int)HIT_MAX - 1; i >= 1; i--)
            if ((uint32_t)i > at && (uint32_t)i <= top) { ht[i] = ht[i - 1]; hv[i] = hv[i - 1]; hn[i] = hn[i - 1]; }
#pragma unroll
          for (int i = 0; i < (int)HIT_MAX; i++)
            if ((uint32_t)i == at) { ht[i] = tp; hv[i] = tau; hn[i] = ni; }
          if (full) {
            ht[HIT_MAX - 1u] = fmaxf(ht[HIT_MAX - 1u], stt);
            hv[HIT_MAX - 1u] = fminf(hv[HIT_MAX - 1u] + sv, 64.0f);
          }
          k = min(k + 1u, HIT_MAX);
        }
      }
      if (!__any(sp > 0u)) break;
    }
  }
  float T = 1.0f, m1 = 0.0f, m2 = 0.0f, first_part = -1.0f, near = 0.0f;
  v3 acc = V3(0.0f, 0.0f, 0.0f);
  float bw[9];
  sh2_basis(neg3(w), bw);
#pragma unroll
  for (int i = 0; i < (int)HIT_MAX; i++) {
    if ((uint32_t)i >= k || T <= 0.0f) break;
    const float al = 1.0f - expf(-hv[i]);
    // node_radiance (bounce.wgsl:122-129): the node's SH2 light along -w
    const uint2 *sh = reinterpret_cast<const uint2 *>(node_sh + 14 * (size_t)hn[i]);
    float hh[28];
#pragma unroll
    for (int j = 0; j < 7; j++) {
      const uint2 v = sh[j];
      hh[4 * j] = unpack_half(v.x & 0xFFFFu); hh[4 * j + 1] = unpack_half(v.x >> 16);
      hh[4 * j + 2] = unpack_half(v.y & 0xFFFFu); hh[4 * j + 3] = unpack_half(v.y >> 16);
    }
    v3 rad = V3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int m = 0; m < 9; m++) rad = add3(rad, mul3s(V3(hh[3 * m], hh[3 * m + 1], hh[3 * m + 2]), bw[m]));
    acc = add3(acc, mul3s(max3s(rad, 0.0f), T * al));
    const float dep = clampb(ht[i], 0.0f, t1);
    m1 += T * al * dep;
    m2 += T * al * dep * dep;
    const float Tn = T * (1.0f - al);
    if (T >= 0.5f && Tn < 0.5f) {
      const uint32_t hp = __float_as_uint(nodes[6 * (size_t)hn[i] + 3].w);
      first_part = (float)hp;
      near = dep < 0.05f ? 1.0f : 0.0f;
      if (hn[i] < G && vgrid_lookup(vgrid + 128 * (size_t)hn[i], xf_dir_inv(load_xf(parts, hp), neg3(w))) < 0.1f)
        near += 2.0f;
    }
    T = Tn;
  }
  if (ENV.info[1] != 0u && t1 >= PP->tmax * 0.999f) acc = add3(acc, mul3s(sky_radiance(ENV, w), T));
  m1 += T * t1;
  m2 += T * t1 * t1;
  const float a = PP->alpha;
  const float4 old0 = probe_data[2 * (size_t)id], old1 = probe_data[2 * (size_t)id + 1];
  float z1 = old1.z;
  if (p >= np && cell >= 1u && cell <= 3u) z1 = cell == 1u ? o.x : (cell == 2u ? o.y : o.z);
  probe_data[2 * (size_t)id] = make_float4(mixf(old0.x, acc.x, a), mixf(old0.y, acc.y, a), mixf(old0.z, acc.z, a), first_part);
  probe_data[2 * (size_t)id + 1] = make_float4(mixf(old1.x, m1, a), mixf(old1.y, m2, a), z1, near);
}

// probe_valid (bounce.wgsl:1417-1445) and probe_sh2 (bounce.wgsl:1028-1041), one wave per probe, two cells a lane.
// Validity: a lattice probe is dropped when most of its rays first hit one movable part, or more than half start near
// a surface or hit a Gaussian from its buried side (counts by ballot); part and volume probes stay valid. SH2 of the
// probe's incoming light over its 128 cells: each cell's weighted light and basis go to LDS, then lane 3m + ch sums
// coefficient m of channel ch over the cells in cell order, the order of the reference's serial loop.
extern "C" __global__ void __launch_bounds__(64)
sl_probe_finish(FrameUniformsDev U, const ProbeParamsDev *__restrict__ PP, const float4 *__restrict__ parts,
                const float4 *__restrict__ cells, float4 *__restrict__ probe_data, float *__restrict__ probe_sh) {
  __shared__ float s_l[NCELL][3];
  __shared__ float s_y[NCELL][9];
  const uint32_t p = blockIdx.x, lane = threadIdx.x;
  const uint32_t np = lat_count(PP);
  if (p >= np + PP->part_probes) return;
  const float4 *rec = probe_data + 2 * (size_t)p * NCELL;
  uint32_t near = 0u, back = 0u, qs[2];
#pragma unroll
  for (uint32_t h = 0; h < 2u; h++) {
    const uint32_t c = lane + 64u * h;
    const float4 r0 = rec[2 * c], r1 = rec[2 * c + 1], cl = cells[c];
    const v3 L = mul3s(xyz(r0), cl.w);
    s_l[c][0] = L.x;
    s_l[c][1] = L.y;
    s_l[c][2] = L.z;
    float y[9];
    sh2_basis(xyz(cl), y);
#pragma unroll
    for (int m = 0; m < 9; m++) s_y[c][m] = y[m];
    const uint32_t fl = (uint32_t)(fmaxf(r1.w, 0.0f) + 0.5f);
    near += (uint32_t)__popcll(__ballot((fl & 1u) != 0u));
    back += (uint32_t)__popcll(__ballot(((fl >> 1u) & 1u) != 0u));
    qs[h] = r0.w >= 0.0f ? min((uint32_t)r0.w, 16u) : 17u;
  }
  float valid = 1.0f;
  if (p < np) {
    bool enclosed = false;
    for (uint32_t q = 0; q < 16u && q < U.part_count; q++) {
      if ((part_flags(parts, q) & 1u) != 0u) continue;
      const uint32_t cnt = (uint32_t)__popcll(__ballot(qs[0] == q)) + (uint32_t)__popcll(__ballot(qs[1] == q));
      if ((float)cnt > PP->g_valid * (float)NCELL) enclosed = true;
    }
    valid = (enclosed || near * 2u > NCELL || back * 2u > NCELL) ? 0.0f : 1.0f;
  }
  if (lane == 0u) probe_data[2 * (size_t)p * NCELL + 1].z = valid;
  __syncthreads();
  if (lane < 27u) {
    const uint32_t m = lane / 3u, ch = lane % 3u;
    float acc = 0.0f;
    for (uint32_t c = 0; c < NCELL; c++) acc += s_l[c][ch] * s_y[c][m];
    probe_sh[36 * (size_t)p + 4 * m + ch] = acc;
  } else if (lane < 36u) {
    probe_sh[36 * (size_t)p + 4 * (lane - 27u) + 3] = 0.0f;
  }
}

// iso_xfer (bounce.wgsl:993-1026): per Gaussian near a part that moved (or on a forced refresh), the SH2 of its
// transfer iso_norm A(w) V(w) near(w) over the 128 world cells. near(w) is the one-entry placeholder grid the
// reference binds while near field has never been on (src/lighting.js:731-733, zeros): 0 for Gaussian 0, 1 else.
extern "C" __global__ void __launch_bounds__(64)
sl_iso_xfer(FrameUniformsDev U, const ProbeParamsDev *__restrict__ PP, const float4 *__restrict__ g0,
            const float4 *__restrict__ g2, uint32_t st, const float4 *__restrict__ nodes,
            const uint8_t *__restrict__ vgrid, const float4 *__restrict__ parts, const float *__restrict__ iso_norm,
            const float4 *__restrict__ cells, float4 *__restrict__ xfer) {
  const uint32_t g = blockIdx.x * 64u + threadIdx.x;
  if (g >= U.gaussian_count || (U.iso & 3u) != 3u) return;
  const float4 A = g0[g], S = sl_scale(g2, st, g);
  const uint32_t own = __float_as_uint(S.w) & 0xFFFFu;
  const Xf xf = load_xf(parts, own);
  const v3 P = xf_point(xf, V3(A.x, A.y, A.z));
  bool stale = PP->force_near != 0u || (part_flags(parts, own) & 2u) != 0u;
  for (uint32_t q = 0; q < U.part_count && !stale; q++) {
    if (q == own || (part_flags(parts, q) & 2u) == 0u) continue;
    const float4 bd = part_bound(parts, q);
    const float r = bd.w + PP->near_r;
    const v3 bc = V3(bd.x, bd.y, bd.z);
    stale = len3(sub3(P, xf_point(load_xf(parts, q), bc))) < r || len3(sub3(P, xf_point(load_xf_prev(parts, q), bc))) < r;
  }
  if (!stale) return;
  const NodeL nd = load_node(nodes, g);
  const float s3 = s3_of(S);
  const float inorm = iso_norm[g];
  const uint8_t *vg = vgrid + 128 * (size_t)g;
  const float nb = g == 0u ? 0.0f : 1.0f;
  float t[9];
#pragma unroll
  for (int m = 0; m < 9; m++) t[m] = 0.0f;
  for (uint32_t c = 0; c < NCELL; c++) {
    const float4 cl = cells[c];
    const v3 wc = xyz(cl);
    const v3 u = xf_dir_inv(xf, wc);
    const float v = vgrid_lookup(vg, u) * nb;
    if (v <= 0.0f) continue;
    const float a = inorm * footprint(nd, s3 * s3, u) * v * cl.w;
    float y[9];
    sh2_basis(wc, y);
#pragma unroll
    for (int m = 0; m < 9; m++) t[m] += a * y[m];
  }
  xfer[3 * (size_t)g] = make_float4(t[0], t[1], t[2], t[3]);
  xfer[3 * (size_t)g + 1] = make_float4(t[4], t[5], t[6], t[7]);
  xfer[3 * (size_t)g + 2] = make_float4(t[8], 0.0f, 0.0f, 0.0f);
}

namespace bn {
struct PW { uint32_t pid[8]; float w[8]; };

// DDGI Chebyshev visibility of world point P from probe p (bounce.wgsl:759-768)
__device__ float probe_cheb(const ProbeParamsDev *PP, uint32_t pc, const float4 *nodes, const float4 *parts,
                            const float4 *probe_data, uint32_t p, v3 P) {
  const v3 v = sub3(P, probe_pos(PP, pc, nodes, parts, p));
  const float dist = len3(v);
  const float dm = fmaxf(dist, 1e-9f);
  const float4 mm = probe_data[2 * ((size_t)p * NCELL + cell_of(V3(v.x / dm, v.y / dm, v.z / dm))) + 1];
  const float var = fmaxf(mm.y - mm.x * mm.x, 1e-4f);
  const float dd = fmaxf(dist - 0.03f - mm.x, 0.0f);
  if (dd <= 0.0f) return 1.0f;
  return powf(var / (var + dd * dd), 3.0f);
}

// no valid probe among the 8: the 8 best valid ones of the surrounding 4x4x4 block (bounce.wgsl:771-796)
__device__ PW probe_fallback(const ProbeParamsDev *PP, uint32_t pc, const float4 *nodes, const float4 *parts,
                             const float4 *probe_data, v3 P, const uint32_t fi[3]) {
  const int dm0 = (int)PP->dims[0], dm1 = (int)PP->dims[1], dm2 = (int)PP->dims[2];
  PW out;
#pragma unroll
  for (int k = 0; k < 8; k++) { out.pid[k] = 0u; out.w[k] = 0.0f; }
  for (uint32_t j = 0; j < 64u; j++) {
    const int i0 = (int)fi[0] + (int)(j >> 4u) - 1, i1 = (int)fi[1] + (int)((j >> 2u) & 3u) - 1, i2 = (int)fi[2] + (int)(j & 3u) - 1;
    if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= dm0 || i1 >= dm1 || i2 >= dm2) continue;
    const uint32_t p = lat_index(PP, (uint32_t)i0, (uint32_t)i1, (uint32_t)i2);
    if (probe_data[2 * ((size_t)p * NCELL) + 1].z <= 0.0f) continue;
    const float d = fmaxf(len3(sub3(P, probe_pos(PP, pc, nodes, parts, p))), 0.05f);
    const float w = fmaxf(probe_cheb(PP, pc, nodes, parts, probe_data, p, P), 1e-3f) / (d * d);
    if (w <= out.w[7]) continue;
    // insertion into the descending top 8 (register arrays: unrolled)
    uint32_t at = 7u;
#pragma unroll
    for (int s = 7; s >= 1; s--)
      if (at == (uint32_t)s && out.w[s - 1] < w) at = (uint32_t)s - 1u;
#pragma unroll
    for (int s = 7; s >= 1; s--)
      if ((uint32_t)s > at) { out.w[s] = out.w[s - 1]; out.pid[s] = out.pid[s - 1]; }
#pragma unroll
    for (int s = 0; s < 8; s++)
      if ((uint32_t)s == at) { out.w[s] = w; out.pid[s] = p; }
  }
  float wsum = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; k++) wsum += out.w[k];
#pragma unroll
  for (int k = 0; k < 8; k++) out.w[k] /= fmaxf(wsum, 1e-12f);
  return out;
}

// the 8 probes around P and their normalised weights (bounce.wgsl:813-885)
__device__ PW probe_weights(const ProbeParamsDev *PP, uint32_t pc, const float4 *nodes, const float4 *parts,
                            const float4 *probe_data, v3 P, uint32_t own) {
  PW out;
  if (is_part_probe_part(PP, parts, own)) {
    const PFrame pf = part_probe_frame(nodes, parts, own);
    const v3 lp = xf_point_inv(load_xf(parts, own), P);
    const v3 f = clamp3(div3(add3(sub3(lp, pf.c), pf.h), max3s(mul3s(pf.h, 2.0f), 1e-6f)), 0.0f, 1.0f);
#pragma unroll
    for (uint32_t k = 0; k < 8u; k++) {
      const float t0 = (k >> 2u) ? f.x : 1.0f - f.x, t1 = ((k >> 1u) & 1u) ? f.y : 1.0f - f.y, t2 = (k & 1u) ? f.z : 1.0f - f.z;
      out.pid[k] = lat_count(PP) + own * 8u + k;
      out.w[k] = t0 * t1 * t2;
    }
    return out;
  }
  if (has_vol(PP, own)) {
    const PFrame vf = vol_frame(nodes, parts, own);
    const v3 lp = xf_point_inv(load_xf(parts, own), P);
    const uint32_t d[3] = {PP->vols[own][1], PP->vols[own][2], PP->vols[own][3]};
    const v3 f = clamp3(div3(add3(sub3(lp, vf.c), vf.h), max3s(mul3s(vf.h, 2.0f), 1e-6f)), 0.0f, 1.0f);
    const float fq[3] = {f.x * ((float)d[0] - 1.0f), f.y * ((float)d[1] - 1.0f), f.z * ((float)d[2] - 1.0f)};
    uint32_t fi[3];
    float fr[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      fi[a] = min((uint32_t)floorf(fq[a]), d[a] - 2u);
      fr[a] = fq[a] - (float)fi[a];
    }
#pragma unroll
    for (uint32_t k = 0; k < 8u; k++) {
      const uint32_t o0 = k >> 2u, o1 = (k >> 1u) & 1u, o2 = k & 1u;
      out.pid[k] = lat_count(PP) + PP->vols[own][0] + ((fi[0] + o0) * d[1] + (fi[1] + o1)) * d[2] + (fi[2] + o2);
      out.w[k] = (o0 ? fr[0] : 1.0f - fr[0]) * (o1 ? fr[1] : 1.0f - fr[1]) * (o2 ? fr[2] : 1.0f - fr[2]);
    }
    return out;
  }
  // lat_cell (bounce.wgsl:798-811): columns bilinear in x, y; the height above the ground among the levels
  const uint32_t dm0 = PP->dims[0], dm1 = PP->dims[1], dm2 = PP->dims[2];
  const float fx = clampb((P.x - PP->lo[0]) / PP->step[0], 0.0f, (float)(dm0 - 1u));
  const float fy = clampb((P.y - PP->lo[1]) / PP->step[1], 0.0f, (float)(dm1 - 1u));
  const uint32_t ix = min((uint32_t)floorf(fx), dm0 - 2u), iy = min((uint32_t)floorf(fy), dm1 - 2u);
  const float tx = fx - (float)ix, ty = fy - (float)iy;
  const float gnd = mixf(mixf(lat_ground(PP, ix, iy), lat_ground(PP, ix, iy + 1u), ty),
                         mixf(lat_ground(PP, ix + 1u, iy), lat_ground(PP, ix + 1u, iy + 1u), ty), tx);
  const float hgt = P.z - gnd;
  uint32_t kz = 0;
  while (kz + 2u < dm2 && hgt > lat_level(PP, kz + 1u)) kz += 1u;
  const float tz = clampb((hgt - lat_level(PP, kz)) / fmaxf(lat_level(PP, kz + 1u) - lat_level(PP, kz), 1e-6f), 0.0f, 1.0f);
  const uint32_t fi[3] = {ix, iy, kz};
  const float fr[3] = {tx, ty, tz};
  float wsum = 0.0f;
#pragma unroll
  for (uint32_t k = 0; k < 8u; k++) {
    const uint32_t o0 = k >> 2u, o1 = (k >> 1u) & 1u, o2 = k & 1u;
    const uint32_t p = lat_index(PP, fi[0] + o0, fi[1] + o1, fi[2] + o2);
    out.pid[k] = p;
    const float valid = probe_data[2 * ((size_t)p * NCELL) + 1].z;
    const float cw = (o0 ? fr[0] : 1.0f - fr[0]) * (o1 ? fr[1] : 1.0f - fr[1]) * (o2 ? fr[2] : 1.0f - fr[2]) * valid;
    out.w[k] = cw * fmaxf(probe_cheb(PP, pc, nodes, parts, probe_data, p, P), 1e-3f);
    wsum += out.w[k];
  }
  if (wsum <= 0.0f) return probe_fallback(PP, pc, nodes, parts, probe_data, P, fi);
#pragma unroll
  for (int k = 0; k < 8; k++) out.w[k] /= fmaxf(wsum, 1e-12f);
  return out;
}
}  // namespace bn

// bounce_sh (bounce.wgsl:1043-1095), matte materials: alb max(t' . L, 0) with L the 8 probes' weighted SH2, written
// to the eye and send slots of lo_b (3 float4 per Gaussian: eye, send, lanes)
extern "C" __global__ void __launch_bounds__(64)
sl_bounce_sh(FrameUniformsDev U, const ProbeParamsDev *__restrict__ PP, const float4 *__restrict__ g0,
             const float4 *__restrict__ g1, const float4 *__restrict__ g2, const float4 *__restrict__ g3, uint32_t st,
             const float4 *__restrict__ nodes, const float4 *__restrict__ parts, const float4 *__restrict__ probe_data,
             const float4 *__restrict__ probe_sh, const float4 *__restrict__ xfer, float4 *__restrict__ lo_b) {
  const uint32_t t = blockIdx.x * 64u + threadIdx.x;
  const uint32_t g = PP->sh_full != 0u ? t : t * PP->subset + PP->phase;
  if (g >= U.gaussian_count || (U.iso & 3u) != 3u) return;
  const float4 A = g0[g], S = sl_scale(g2, st, g), C = sl_albedo(g1, g2, g3, st, g);
  const uint32_t own = __float_as_uint(S.w) & 0xFFFFu;
  const v3 P = xf_point(load_xf(parts, own), V3(A.x, A.y, A.z));
  const PW pw = probe_weights(PP, U.part_count, nodes, parts, probe_data, P, own);
  v3 Lsh[9];
#pragma unroll
  for (int m = 0; m < 9; m++) Lsh[m] = V3(0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const float w = pw.w[k];
    if (w <= 0.0f) continue;
#pragma unroll
    for (int m = 0; m < 9; m++) Lsh[m] = add3(Lsh[m], mul3s(xyz(probe_sh[9 * (size_t)pw.pid[k] + m]), w));
  }
  const float4 t0 = xfer[3 * (size_t)g], t1 = xfer[3 * (size_t)g + 1], t2 = xfer[3 * (size_t)g + 2];
  const float tt[9] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w, t2.x};
  v3 s = mul3s(Lsh[0], tt[0]);
#pragma unroll
  for (int m = 1; m < 9; m++) s = add3(s, mul3s(Lsh[m], tt[m]));
  const v3 cb = mul3(V3(C.x, C.y, C.z), max3s(s, 0.0f));
  lo_b[3 * (size_t)g] = make_float4(cb.x, cb.y, cb.z, 0.0f);
  lo_b[3 * (size_t)g + 1] = make_float4(cb.x, cb.y, cb.z, 0.0f);
  lo_b[3 * (size_t)g + 2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
