// Relighting on gfx950: iso_norm per leaf, the emitter-light cut walk, the emitter and sun opacity maps, and the
// distant lights' direct light (runtime/wgsl/{iso_init,emissive,light_shadow,sun_map,environment}.wgsl, default
// static M39 matte path).
//
// One lane per Gaussian everywhere except the map resolve/down passes (one lane per texel). Contraction is off in
// this file so every float expression rounds as written in WGSL order; the map splats take exp from exp_map (fused
// operations only), so their fixed-point sums (order-independent u32 adds) are bit-identical to any implementation
// using the same definition.
#pragma clang fp contract(off)
#include "sl_device.h"
#include <hip/hip_fp16.h>

namespace lt {

constexpr float PI_F = 3.14159265358979f;
constexpr uint32_t NO_CHILD = 0xFFFFFFFFu;

struct v3 { float x, y, z; };
struct v2 { float x, y; };
struct m2 { v2 c0, c1; };        // column-major
struct m3 { v3 c0, c1, c2; };

__device__ __forceinline__ v3 V3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ v2 V2(float x, float y) { return {x, y}; }
__device__ __forceinline__ v3 add3(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 sub3(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 mul3s(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float dot2(v2 a, v2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ float len3(v3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ v3 norm3(v3 a) { const float l = len3(a); return V3(a.x / l, a.y / l, a.z / l); }
__device__ __forceinline__ v3 cross3(v3 a, v3 b) {
  return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float clampf_(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
// transcendentals whose result picks a cell, a mip level or a cull decision: double, rounded once, so the
// choice does not hang on a library's last-ulp behaviour
__device__ __forceinline__ float log2_cr(float x) { return (float)log2((double)x); }
__device__ __forceinline__ float asin_cr(float x) { return (float)asin((double)x); }
__device__ __forceinline__ float acos_cr(float x) { return (float)acos((double)x); }
// WGSL u32(f32): truncate toward zero, saturate to [0, 2^32 - 1], NaN -> 0, which is exactly what v_cvt_u32_f32
// does (C++'s conversion leaves out-of-range values undefined, so the compiler would add range checks)
__device__ __forceinline__ uint32_t u32_sat(float f) {
  uint32_t r;
  asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(f));
  return r;
}
// e^x for the opacity-map splats: Cody-Waite reduction by ln 2 (split in two, fused) and the degree-7 Taylor
// polynomial on [-ln2/2, ln2/2] in Horner form with explicit fused multiply-adds, so every IEEE operation is
// the same on any host or device that has fma; within 2 ulp of e^x. exp_map clamps x to [-86, 88] (normal
// results); exp_map_in takes x already in that range: the map sums call it with -q/2, where 0 <= q <= 9 up to
// rounding (a positive definite form after the 0.3 texel^2 regulariser; NaN and q > 9 are skipped before).
__device__ __forceinline__ float exp_map_in(float x) {
  const float k = rintf(x * 1.44269504f);
  float r = fmaf(k, -0.693145751953125f, x);
  r = fmaf(k, -1.42860677e-6f, r);
  float p = 1.98412698e-4f;                  // 1/7!
  p = fmaf(p, r, 1.38888889e-3f);            // 1/6!
  p = fmaf(p, r, 8.33333333e-3f);            // 1/5!
  p = fmaf(p, r, 4.16666667e-2f);            // 1/4!
  p = fmaf(p, r, 1.66666667e-1f);            // 1/3!
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}
__device__ __forceinline__ float exp_map(float x) { return exp_map_in(fminf(fmaxf(x, -86.0f), 88.0f)); }
__device__ __forceinline__ v3 m3v(m3 m, v3 v) { return add3(add3(mul3s(m.c0, v.x), mul3s(m.c1, v.y)), mul3s(m.c2, v.z)); }
__device__ __forceinline__ m3 m3mul(m3 a, m3 b) { return {m3v(a, b.c0), m3v(a, b.c1), m3v(a, b.c2)}; }
__device__ __forceinline__ m3 m3t(m3 m) {
  return {V3(m.c0.x, m.c1.x, m.c2.x), V3(m.c0.y, m.c1.y, m.c2.y), V3(m.c0.z, m.c1.z, m.c2.z)};
}
__device__ __forceinline__ v2 m2v(m2 m, v2 v) { return V2(m.c0.x * v.x + m.c1.x * v.y, m.c0.y * v.x + m.c1.y * v.y); }
__device__ __forceinline__ m2 m2mul(m2 a, m2 b) { return {m2v(a, b.c0), m2v(a, b.c1)}; }
__device__ __forceinline__ m2 m2t(m2 m) { return {V2(m.c0.x, m.c1.x), V2(m.c0.y, m.c1.y)}; }
__device__ __forceinline__ m2 m2s(m2 m, float s) { return {V2(m.c0.x * s, m.c0.y * s), V2(m.c1.x * s, m.c1.y * s)}; }
__device__ __forceinline__ float m2det(m2 m) { return m.c0.x * m.c1.y - m.c1.x * m.c0.y; }
__device__ __forceinline__ m2 inv2(m2 m) {   // light_shadow.wgsl:112-115
  const float d = m2det(m);
  return m2s({V2(m.c1.y, -m.c0.y), V2(-m.c1.x, m.c0.x)}, 1.0f / d);
}
__device__ __forceinline__ m3 quat_matrix(float4 q) {   // runtime/wgsl/common.wgsl:218-223
  const float x = q.x, y = q.y, z = q.z, w = q.w;
  return {V3(1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + z * w), 2.0f * (x * z - y * w)),
          V3(2.0f * (x * y - z * w), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + x * w)),
          V3(2.0f * (x * z + y * w), 2.0f * (y * z - x * w), 1.0f - 2.0f * (x * x + y * y))};
}

struct Xf { v3 r0, r1, r2; v3 t; };   // a part's world_from_rest rows [R row | t]
__device__ __forceinline__ Xf load_xf(const float4 *parts, uint32_t p) {
  const float4 a = parts[p * 8 + 0], b = parts[p * 8 + 1], c = parts[p * 8 + 2];
  return {V3(a.x, a.y, a.z), V3(b.x, b.y, b.z), V3(c.x, c.y, c.z), V3(a.w, b.w, c.w)};
}
__device__ __forceinline__ v3 xf_point(const Xf &x, v3 p) {
  return V3(dot3(x.r0, p) + x.t.x, dot3(x.r1, p) + x.t.y, dot3(x.r2, p) + x.t.z);
}
__device__ __forceinline__ v3 xf_dir_inv(const Xf &x, v3 d) {
  return add3(add3(mul3s(x.r0, d.x), mul3s(x.r1, d.y)), mul3s(x.r2, d.z));
}
__device__ __forceinline__ m3 part_rot(const Xf &x) { return m3t({x.r0, x.r1, x.r2}); }

__device__ __forceinline__ float erfc_pos(float z) {   // runtime/wgsl/common.wgsl:113-118
  const float t = 1.0f / (1.0f + 0.5f * z);
  const float p = -z * z - 1.26551223f + t * (1.00002368f + t * (0.37409196f + t * (0.09678418f + t * (-0.18628806f
                  + t * (0.27886807f + t * (-1.13520398f + t * (1.48851587f + t * (-0.82215223f + t * 0.17087277f))))))));
  return t * expf(p);
}
__device__ __forceinline__ float phi_(float x) {
  const float a = fabsf(x) * 0.70710678f;
  const float h = 0.5f * erfc_pos(a);
  return x >= 0.0f ? 1.0f - h : h;
}
__device__ __forceinline__ v3 sym_mul(v3 dg, v3 off, v3 v) {
  return V3((dg.x * v.x + off.x * v.y) + off.y * v.z, (off.x * v.x + dg.y * v.y) + off.z * v.z,
            (off.y * v.x + off.z * v.y) + dg.z * v.z);
}

// Node (96 B): mean 0, radius 3, cinv_diag 4, kappa 7, cinv_off 8, smax 11, sn_diag 12, part 15, sn_off 16
struct NodeL { v3 cinv_d, cinv_o, sn_d, sn_o; float kappa; };
__device__ __forceinline__ NodeL load_node(const float4 *nodes, uint32_t g) {
  const float4 *n = nodes + 6 * (size_t)g;
  const float4 a = n[1], b = n[2], c = n[3], d = n[4];
  return {V3(a.x, a.y, a.z), V3(b.x, b.y, b.z), V3(c.x, c.y, c.z), V3(d.x, d.y, d.z), a.w};
}
__device__ __forceinline__ float dirm(const NodeL &nd, v3 w) {
  const float q = dot3(w, sym_mul(nd.sn_d, nd.sn_o, w));
  return sqrtf(q > 0.0f ? q : 0.0f);
}
// a / b as a times the hardware reciprocal of b (v_rcp_f32, 1 ulp): within the 2.5 ulp WGSL allows a division
// over the normal range, at a tenth of the instructions of the correctly rounded quotient
__device__ __forceinline__ float qdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

__device__ __forceinline__ float ein(float x) {   // runtime/wgsl/common.wgsl:226-239
  if (x < 1.0f) {
    // the reference's series; each division by k is a product with 1/k rounded to float
    float term = x, s = x;
#pragma unroll
    for (int k = 2; k <= 12; k++) {
      const float ik = 1.0f / (float)k;                 // folded at compile time
      term *= -x * ik;
      s += term * ik;
    }
    return s;
  }
  const float p = x * x * x * x + 8.5733287401f * x * x * x + 18.0590169730f * x * x + 8.6347608925f * x + 0.2677737343f;
  const float q = x * x * x * x + 9.5733223454f * x * x * x + 25.6329561486f * x * x + 21.0996530827f * x + 3.9584969228f;
  return logf(x) + 0.5772156649f + qdiv(qdiv(p, q), x * expf(x));
}
__device__ __forceinline__ float footprint(const NodeL &nd, float detC, v3 w) {   // common.wgsl:242-246
  const float a = dot3(w, sym_mul(nd.cinv_d, nd.cinv_o, w));
  const float K = nd.kappa * dirm(nd, w) * sqrtf(qdiv(2.0f * PI_F, a));
  return 2.0f * PI_F * sqrtf(detC * a) * ein(K);
}
__device__ __forceinline__ float vgrid_cell(const uint8_t *v, uint32_t i, uint32_t j) {
  return (float)v[i * 16u + j] * (1.0f / 255.0f);
}
// bilinear between cell centres, so continuous across cell boundaries: the float asin's last-ulp behaviour moves V
// by rounding only (the host restatement rounds asin once from double). cell(i, j): the grid byte scaled by 1/255.
template <typename Cell>
__device__ __forceinline__ float vgrid_lookup_f(Cell cell, v3 w) {   // common.wgsl:162-177
  const float el = asinf(clampf_(w.z, -1.0f, 1.0f));
  const float fi = clampf_((0.5f * PI_F - el) * (8.0f / PI_F) - 0.5f, 0.0f, 7.0f);
  const uint32_t i0 = min((uint32_t)fi, 6u);
  const float ti = fi - (float)i0;
  const float az = (fabsf(w.x) + fabsf(w.y) < 1e-30f) ? 0.0f : atan2f(w.y, w.x);
  const float fj = (az + PI_F) * (8.0f / PI_F) - 0.5f;
  const float j0f = floorf(fj);
  const float tj = fj - j0f;
  const uint32_t j0 = (uint32_t)((int32_t)j0f + 16) % 16u;
  const uint32_t j1 = (j0 + 1u) % 16u;
  const float a = mixf(cell(i0, j0), cell(i0, j1), tj);
  const float b = mixf(cell(i0 + 1u, j0), cell(i0 + 1u, j1), tj);
  return mixf(a, b, ti);
}
__device__ __forceinline__ float vgrid_lookup(const uint8_t *v, v3 w) {
  return vgrid_lookup_f([&](uint32_t i, uint32_t j) { return vgrid_cell(v, i, j); }, w);
}
__device__ __forceinline__ v3 basis_k(v3 w) {
  v3 a = V3(0.0f, 0.0f, 1.0f);
  if (fabsf(w.z) >= 0.999f) a = V3(1.0f, 0.0f, 0.0f);
  return norm3(cross3(a, w));
}
__device__ __forceinline__ m3 leaf_cov(float4 q, float4 s, m3 R) {
  const m3 Rq = quat_matrix(q);
  const m3 S = {V3(s.x * s.x, 0.0f, 0.0f), V3(0.0f, s.y * s.y, 0.0f), V3(0.0f, 0.0f, s.z * s.z)};
  const m3 C = m3mul(m3mul(Rq, S), m3t(Rq));
  return m3mul(m3mul(R, C), m3t(R));
}
__device__ __forceinline__ float unpack_half(uint32_t h) { return __half2float(__ushort_as_half((unsigned short)h)); }

// ---- shiny receivers (runtime/wgsl/material_ext.wgsl:10-181) -----------------------------------------------------
// A material with flags bit 3 mixes mirror flakes (weight k, pack2x16float(k, 0) in pad0) into the matte transfer:
// each light's term takes iso_lobe = (1 - k) + k 4 pi p damp, p the specular flake phase (Heitz et al. 2015) for a
// point light or its integral over the light's angular ASG when lights are shapes (EmitParams.n5 bit 3).
__device__ __forceinline__ float sym_det(v3 d, v3 o) {
  return d.x * (d.y * d.z - o.z * o.z) - o.x * (o.x * d.z - o.z * o.y) + o.y * (o.x * o.z - d.y * o.y);
}
__device__ __forceinline__ float sym_adj_quad(v3 d, v3 o, v3 w) {   // w^T S^-1 w det S via the adjugate
  const float a00 = d.y * d.z - o.z * o.z, a11 = d.x * d.z - o.y * o.y, a22 = d.x * d.y - o.x * o.x;
  const float a01 = o.y * o.z - o.x * d.z, a02 = o.x * o.z - o.y * d.y, a12 = o.x * o.y - d.x * o.z;
  return a00 * w.x * w.x + a11 * w.y * w.y + a22 * w.z * w.z +
         2.0f * (a01 * w.x * w.y + a02 * w.x * w.z + a12 * w.y * w.z);
}
__device__ __forceinline__ float phase_specular(v3 Sd, v3 So, v3 wi, v3 wo) {   // material_ext.wgsl:27-37
  const v3 h = add3(wi, wo);
  const float hl = len3(h);
  if (hl < 1e-6f) return 0.0f;
  const v3 wh = V3(h.x / hl, h.y / hl, h.z / hl);
  const float det = fmaxf(sym_det(Sd, So), 1e-20f);
  const float q = sym_adj_quad(Sd, So, wh) / det;
  const float D = 1.0f / (PI_F * sqrtf(det) * q * q);
  const float sigma = sqrtf(fmaxf(dot3(wi, sym_mul(Sd, So, wi)), 1e-12f));
  return D / (4.0f * sigma);
}
__device__ __forceinline__ float spec_weight(const uint32_t *mat, uint32_t mflags) {   // material_ext.wgsl:39-42
  if ((mflags & 8u) == 0u || (mflags & 1u) != 0u) return 0.0f;
  return clampf_(unpack_half(mat[6] & 0xFFFFu), 0.0f, 1.0f);
}
__device__ __forceinline__ float iso_lobe_from(float k, float p_spec, float damp) {
  return (1.0f - k) + k * 4.0f * PI_F * p_spec * damp;
}
__device__ __forceinline__ float iso_damp(float inorm, float area_vis_r) {
  return fminf(1.0f, 1.0f / fmaxf(4.0f * PI_F * inorm * area_vis_r, 1e-30f));
}
__device__ __forceinline__ v3 mirror(v3 wi, v3 n) { return sub3(mul3s(n, 2.0f * dot3(wi, n)), wi); }
__device__ __forceinline__ void onb(v3 n, v3 &b0, v3 &b1) {   // Duff et al. 2017 (material_ext.wgsl:83-88)
  const float s = n.z >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (s + n.z);
  const float b = n.x * n.y * a;
  b0 = V3(1.0f + s * n.x * n.x * a, s * b, -s * n.x);
  b1 = V3(b, s + n.y * n.y * a, -n.y);
}
__device__ __forceinline__ float s2_det(v3 s) { return s.x * s.z - s.y * s.y; }
__device__ __forceinline__ float s2_quad(v3 s, v2 v) { return s.x * v.x * v.x + 2.0f * s.y * v.x * v.y + s.z * v.y * v.y; }
__device__ __forceinline__ float s2_quad_inv(v3 s, v2 v) {
  return (s.z * v.x * v.x - 2.0f * s.y * v.x * v.y + s.x * v.y * v.y) / fmaxf(s2_det(s), 1e-30f);
}
__device__ __forceinline__ v3 s2_xf(v3 s, v2 t0, v2 t1) {
  const v2 a = V2(s.x * t0.x + s.y * t0.y, s.y * t0.x + s.z * t0.y);
  const v2 b = V2(s.x * t1.x + s.y * t1.y, s.y * t1.x + s.z * t1.y);
  return V3(dot2(t0, a), dot2(t0, b), dot2(t1, b));
}
__device__ __forceinline__ v3 s2_major(v3 s) {
  const float h = 0.5f * (s.x - s.z);
  const float lam = 0.5f * (s.x + s.z) + sqrtf(h * h + s.y * s.y);
  const v2 v = s.x >= s.z ? V2(lam - s.z, s.y) : V2(s.y, lam - s.x);
  const float l = sqrtf(dot2(v, v));
  return l > 1e-20f ? V3(v.x / l, v.y / l, lam) : V3(1.0f, 0.0f, lam);
}
__device__ __forceinline__ float s2_minor(v3 s) {
  const float h = 0.5f * (s.x - s.z);
  return 0.5f * (s.x + s.z) - sqrtf(h * h + s.y * s.y);
}
__device__ __forceinline__ v2 log_map(v3 c, v3 cb0, v3 cb1, v3 a) {   // material_ext.wgsl:118-124
  const float ca = clampf_(dot3(c, a), -1.0f, 1.0f);
  const v3 v = sub3(a, mul3s(c, ca));
  const float nv = len3(v);
  if (nv < 1e-9f) return V2(0.0f, 0.0f);
  const float k = acosf(ca) / nv;
  return V2(k * dot3(cb0, v), k * dot3(cb1, v));
}
__device__ __forceinline__ float smooth_(float e0, float e1, float x) {
  const float t = clampf_((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
// the specular phase integrated over a light ASG centred on wi, covariance sl (xx, xy, yy) in basis lb0, lb1
// (material_ext.wgsl:128-181); n0, t1, t2 the receiver's axes, flake their eigenvalues
__device__ __noinline__ float asg_phase(v3 wi, v3 lb0, v3 lb1, v3 sl, v3 n0, v3 t1, v3 t2, v3 flake, v3 Sd, v3 So,
                                        v3 wo) {
  const float pc = phase_specular(Sd, So, wi, wo);
  if (flake.x <= fmaxf(flake.y, flake.z) * 1.0001f) return pc;
  const float co = dot3(n0, wo);
  const v3 n = co >= 0.0f ? n0 : V3(-n0.x, -n0.y, -n0.z);
  const float cos_o = fabsf(co);
  const v3 wr = sub3(mul3s(n, 2.0f * cos_o), wo);
  v3 rb0, rb1;
  onb(wr, rb0, rb1);
  const v3 dw1 = add3(mul3s(n, 2.0f * dot3(wo, t1)), mul3s(t1, 2.0f * cos_o));
  const v3 dw2 = add3(mul3s(n, 2.0f * dot3(wo, t2)), mul3s(t2, 2.0f * cos_o));
  const v2 j0 = V2(dot3(rb0, dw1), dot3(rb0, dw2));
  const v2 j1 = V2(dot3(rb1, dw1), dot3(rb1, dw2));
  const v2 ia = V2(flake.y / (flake.x - flake.y), flake.z / (flake.x - flake.z));
  const v3 s0 = V3(dot2(V2(j0.x * j0.x, j0.y * j0.y), ia), dot2(V2(j0.x * j1.x, j0.y * j1.y), ia),
                   dot2(V2(j1.x * j1.x, j1.y * j1.y), ia));
  const float pk = flake.x * flake.x / (PI_F * sqrtf(flake.x * flake.y * flake.z)) /
                   (4.0f * sqrtf(fmaxf(dot3(wr, sym_mul(Sd, So, wr)), 1e-12f)));
  const v3 cs = add3(wi, wr);
  const float cl = len3(cs);
  if (cl < 1e-4f) return pc;
  const v3 c = V3(cs.x / cl, cs.y / cl, cs.z / cl);
  v3 cb0, cb1;
  onb(c, cb0, cb1);
  const v2 lr = log_map(c, cb0, cb1, wr), li = log_map(c, cb0, cb1, wi);
  const v2 dm = V2(lr.x - li.x, lr.y - li.y);
  const v3 sL = s2_xf(sl, V2(dot3(cb0, lb0), dot3(cb0, lb1)), V2(dot3(cb1, lb0), dot3(cb1, lb1)));
  const v3 sR = s2_xf(s0, V2(dot3(cb0, rb0), dot3(cb0, rb1)), V2(dot3(cb1, rb0), dot3(cb1, rb1)));
  const float ndm = fmaxf(sqrtf(dot2(dm, dm)), 1e-9f);
  const float a2 = 2.0f * 0.10727f;
  const float d = sqrtf(s2_quad_inv(V3(sL.x + sR.x / a2, sL.y + sR.y / a2, sL.z + sR.z / a2), dm));
  const float along = sqrtf(fmaxf(s2_quad(sL, dm), 0.0f)) / ndm;
  const float far_ = smooth_(1.0f, 2.5f, d) * (1.0f - smooth_(0.15f, 0.3f, along / ndm));
  const v3 maj = s2_major(sl);
  const float ratio = sqrtf(maj.z / fmaxf(s2_minor(s0) / (2.0f * 2.54696f), 1e-30f));
  const float t = fmaxf(far_, 1.0f - smooth_(0.12f, 0.3f, ratio));
  float p_asg = 0.0f;
  if (t < 1.0f) {
    const float aw[3] = {0.61163f, 0.35083f, 0.0267f}, aa[3] = {2.54696f, 0.6893f, 0.10727f};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float dk = 2.0f * aa[k];
      const v3 sk = V3(sR.x / dk, sR.y / dk, sR.z / dk);
      const v3 st = add3(sL, sk);
      p_asg += aw[k] * sqrtf(fmaxf(s2_det(sk), 0.0f) / fmaxf(s2_det(st), 1e-30f)) * expf(-0.5f * s2_quad_inv(st, dm));
    }
    p_asg *= pk;
  }
  float p_pt = pc;
  if (t > 0.0f) {
    const v3 ax = mul3s(add3(mul3s(lb0, maj.x), mul3s(lb1, maj.y)), sqrtf(3.0f * maj.z));
    p_pt = (2.0f / 3.0f) * pc + (1.0f / 6.0f) * (phase_specular(Sd, So, norm3(add3(wi, ax)), wo) +
                                                 phase_specular(Sd, So, norm3(sub3(wi, ax)), wo));
  }
  return mixf(p_asg, p_pt, t);
}
// S = R diag(flake) R^T in the rest frame as (diag, off = xy, xz, yz) (emissive.wgsl:141-150)
__device__ __forceinline__ void flake_matrix(const m3 &Rq, v3 flake, v3 &Sd, v3 &So) {
  const m3 D = {V3(flake.x, 0.0f, 0.0f), V3(0.0f, flake.y, 0.0f), V3(0.0f, 0.0f, flake.z)};
  const m3 Sl = m3mul(m3mul(Rq, D), m3t(Rq));
  Sd = V3(Sl.c0.x, Sl.c1.y, Sl.c2.z);
  So = V3(Sl.c1.x, Sl.c2.x, Sl.c2.y);
}

}  // namespace lt

// Uniform layouts (the bytes the host writes)
struct FrameUniformsDev {   // runtime/wgsl/common.wgsl:39-47
  uint32_t frame, gaussian_count, part_count, lamp_count;
  float theta, alpha_min, alpha_part;
  uint32_t slot_count;
  float eye[3];
  uint32_t stratum_step, light_samples_axis, material_count, history, iso;
};
struct EmitParamsDev {      // runtime/wgsl/emissive.wgsl:34-53
  uint32_t count, root;
  float scale, theta;
  uint32_t mode, node_count, skip_part, one_sided;
  float normal[3];
  uint32_t n5;
  float edge_soft;
  uint32_t hier, stride, lights;
};
struct LSParamsDev {        // runtime/wgsl/light_shadow.wgsl:26-38
  float center[3]; uint32_t res;
  float normal[3]; uint32_t levels;
  float t1[3]; uint32_t layer_texels;
  float t2[3]; uint32_t dirty;
  float sigma_l; uint32_t skip_part, hemi, part_count;
  float cut_theta; uint32_t tiles, tile_cap; float refine;
  float lc_diag[3]; uint32_t max_lobes;
  float lc_off[3]; uint32_t budget;
  uint32_t slices; float d0, dlog; uint32_t combine;
  uint32_t level_off[16];
  float layers[32];
};
struct SunParamsDev {       // runtime/wgsl/sun_map.wgsl:24-32
  uint32_t n_layers, res, levels, layer_texels, n_lights, n_movable, dirty, range_count;
  float margin, shutter, cut_k;
  uint32_t node_count;
  float lights[4][4];
  uint32_t part_layers[64];
  uint32_t level_off[16];
};
struct EnvDev {             // runtime/wgsl/environment.wgsl:21-30 (416 B buffer)
  float sh[16][4];
  float lights[4][8];
  uint32_t info[4];
  uint32_t pad[4];
};
struct SunLayerDesc { uint32_t part, light, kind, dirty; float c[2], half, pad; };   // sun_map.wgsl:36
struct SunLayer { float c[2]; float half, zref; float t1[3]; uint32_t part; float t2[3]; uint32_t light; };
static_assert(sizeof(FrameUniformsDev) == 64, "FrameUniforms");
static_assert(sizeof(EmitParamsDev) == 64, "EmitParams");
static_assert(sizeof(LSParamsDev) == 336, "LSParams");
static_assert(sizeof(SunParamsDev) == 432, "SunParams");
static_assert(sizeof(EnvDev) == 416, "Env");
static_assert(sizeof(SunLayerDesc) == 32, "LayerDesc");
static_assert(sizeof(SunLayer) == 48, "SunLayer");

using namespace lt;

// ---- iso_init (runtime/wgsl/iso_init.wgsl:29-56) ---------------------------------------------------------------

extern "C" __global__ void __launch_bounds__(64)
sl_iso_init(uint32_t n, const float4 *__restrict__ g2, uint32_t st, const float4 *__restrict__ nodes,
            const uint8_t *__restrict__ vgrid, float *__restrict__ iso_norm) {
  // the 128 cells' directions and solid angles, the same for every Gaussian: computed once per workgroup into LDS
  __shared__ float4 s_cell[128];
  for (uint32_t c = threadIdx.x; c < 128u; c += 64u) {
    const uint32_t i = c / 16u, j = c % 16u;
    const float el = 0.5f * PI_F - ((float)i + 0.5f) * PI_F / 8.0f;
    const float az = -PI_F + ((float)j + 0.5f) * 2.0f * PI_F / 16.0f;
    const float top = 0.5f * PI_F - (float)i * PI_F / 8.0f;
    const float om = (2.0f * PI_F / 16.0f) * (sinf(top) - sinf(top - PI_F / 8.0f));
    s_cell[c] = make_float4(cosf(el) * cosf(az), cosf(el) * sinf(az), sinf(el), om);
  }
  __syncthreads();
  const uint32_t g = blockIdx.x * 64u + threadIdx.x;
  if (g >= n) return;
  const float4 s = sl_scale(g2, st, g);
  const NodeL nd = load_node(nodes, g);
  const float s3 = s.x * s.y * s.z;
  const uint4 *vg = (const uint4 *)(vgrid + 128 * (size_t)g);   // the grid's cache line, read 16 B at a time
  float S = 0.0f, S0 = 0.0f;
  for (uint32_t q = 0; q < 8u; q++) {
    const uint4 v4 = vg[q];
    const uint32_t words[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
    for (uint32_t k = 0; k < 16u; k++) {
      const float4 cw = s_cell[16u * q + k];
      const float v = (float)((words[k >> 2] >> (8u * (k & 3u))) & 0xFFu) * (1.0f / 255.0f);
      const float a = footprint(nd, s3 * s3, V3(cw.x, cw.y, cw.z)) * cw.w;
      S += a * v;
      S0 += a;
    }
  }
  const float m = S > 0.02f * S0 ? S : 0.02f * S0;
  iso_norm[g] = 1.0f / (m > 1e-30f ? m : 1e-30f);
}

// ---- shade_emit (runtime/wgsl/emissive.wgsl:123-150, 853-980), M39 transfer, matte and shiny -----------------------

// A shiny receiver's set-up (emissive.wgsl:141-150, 868-876): its mirror-flake weight, whether the cut's nodes are
// taken as shapes (EmitParams.n5 bit 3), its axes, flake matrix and the eye in its rest frame, and its iso_norm.
struct Shiny {
  float ks, inorm;
  bool shape;
  v3 n, t1, t2, flake, Sd, So, wo;
};
__device__ __forceinline__ Shiny shiny_receiver(const uint32_t *mat, uint32_t mflags, uint32_t n5, const m3 &Rq,
                                                v3 wo_rest, float inorm) {
  Shiny r;
  r.ks = spec_weight(mat, mflags);
  r.shape = r.ks > 0.0f && (n5 & 8u) != 0u;
  r.n = Rq.c0;
  r.t1 = Rq.c1;
  r.t2 = Rq.c2;
  r.flake = V3(__uint_as_float(mat[0]), __uint_as_float(mat[1]), __uint_as_float(mat[2]));
  r.Sd = r.So = V3(0.0f, 0.0f, 0.0f);
  if (r.ks > 0.0f) flake_matrix(Rq, r.flake, r.Sd, r.So);
  r.wo = wo_rest;
  r.inorm = inorm;
  return r;
}
// iso_lobe of one cut node's light on a shiny receiver (emissive.wgsl:897-930): damped toward the mirror direction,
// the node a shape (its covariance words 16-22 seen from the receiver, B C_s B^T / d^2) or a point
template <typename Cell>
__device__ __forceinline__ float shiny_lobe(const Shiny &r, const NodeL &nd, float detC, Cell vcell, const Xf &xf,
                                            v3 wi, v3 ww, float d2, const float4 *e) {
  const v3 wr = mirror(wi, r.n);
  const float damp = iso_damp(r.inorm, footprint(nd, detC, wr) * vgrid_lookup_f(vcell, wr));
  if (r.shape) {
    v3 b0, b1;
    onb(ww, b0, b1);
    const float4 e4 = e[4], e5 = e[5];
    const v3 cd = V3(e4.x, e4.y, e4.z), co = V3(e5.x, e5.y, e5.z);
    const v3 c0 = sym_mul(cd, co, b0), c1 = sym_mul(cd, co, b1);
    const v3 sl = V3(dot3(b0, c0) / d2, dot3(b0, c1) / d2, dot3(b1, c1) / d2);
    const float p = asg_phase(wi, xf_dir_inv(xf, b0), xf_dir_inv(xf, b1), sl, r.n, r.t1, r.t2, r.flake, r.Sd, r.So,
                              r.wo);
    return iso_lobe_from(r.ks, p, damp);
  }
  return iso_lobe_from(r.ks, phase_specular(r.Sd, r.So, wi, r.wo), damp);
}

// Rows of a per-Gaussian array copied into a visiting order: row t of dst = row order[t] of src, rows of `words`
// 16 B words; one lane per word, so a wave writes contiguous memory.
extern "C" __global__ void __launch_bounds__(256)
sl_gather_rows(uint32_t n, const uint32_t *__restrict__ order, const uint4 *__restrict__ src,
               uint4 *__restrict__ dst, uint32_t words) {
  const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
  if (i >= (size_t)n * words) return;
  const uint32_t t = (uint32_t)(i / words), k = (uint32_t)(i - (size_t)t * words);
  dst[i] = src[(size_t)order[t] * words + k];
}
extern "C" __global__ void __launch_bounds__(256)
sl_gather_f32(uint32_t n, const uint32_t *__restrict__ order, const float *__restrict__ src, float *__restrict__ dst) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t < n) dst[t] = src[order[t]];
}

// One cut walk per lane. STACK: the walk's stack depth in LDS, 48 as the reference's; used for trees whose full
// walk can hold more than 16 entries (sl_shade_emit_wave below takes the others). Lane t shades Gaussian order[t]
// (spatially coherent waves); its inputs r0..r3, nodes, vgrid and iso_norm are copies in that visiting order (row t),
// so a wave reads them contiguously; lamp_T and the outputs are indexed by Gaussian.
template <uint32_t STACK, bool SHINY>
__global__ void __launch_bounds__(64)
sl_shade_emit(FrameUniformsDev U, EmitParamsDev EP, const uint32_t *__restrict__ order,
              const float4 *__restrict__ emit_nodes, uint32_t n_nodes,
              const float4 *__restrict__ r0, const float4 *__restrict__ r1, const float4 *__restrict__ r2,
              const float4 *__restrict__ r3, uint32_t st, const float4 *__restrict__ nodes,
              const uint8_t *__restrict__ vgrid, const float4 *__restrict__ parts, const float *__restrict__ lamp_T,
              const uint32_t *__restrict__ materials, const float *__restrict__ iso_norm,
              float4 *__restrict__ colour, float4 *__restrict__ c0_out) {
  const uint32_t t = blockIdx.x * 64u + threadIdx.x;
  if (t >= U.gaussian_count) return;
  const uint32_t g = order[t];
  const float4 A = r0[t], Q = sl_quat(r1, st, t), S = sl_scale(r2, st, t), C = sl_albedo(r1, r2, r3, st, t);
  const uint32_t mi = __float_as_uint(C.w) & 0xFFFFu;
  const uint32_t mmax = max(U.material_count, 1u) - 1u;
  const uint32_t *mat = materials + 8 * (size_t)min(mi, mmax);
  const uint32_t mflags = U.material_count ? mat[5] : 0u;
  if (EP.count == 0u && (mflags & 1u) == 0u) {
    colour[g] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    if (c0_out) c0_out[g] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  const Xf xf = load_xf(parts, __float_as_uint(S.w) & 0xFFFFu);
  const v3 P = xf_point(xf, V3(A.x, A.y, A.z));
  const m3 Rq = quat_matrix(Q);
  const float s3 = S.x * S.y * S.z;
  const float detC = s3 * s3;
  const v3 wo = norm3(sub3(V3(U.eye[0], U.eye[1], U.eye[2]), P));
  const NodeL nd = load_node(nodes, t);
  Shiny sh{};
  if constexpr (SHINY) sh = shiny_receiver(mat, mflags, EP.n5, Rq, xf_dir_inv(xf, wo), iso_norm[t]);
  // the Gaussian's 128 B visibility grid, read once into LDS as 32 words per lane in a [word][lane] layout (lanes
  // on distinct banks whatever cells they read): the walk looks up 4 cells per used node, and re-reading them from
  // memory costs a cache line each time once other waves have evicted it
  __shared__ uint32_t s_vg[32][64];
  uint32_t *vcol = &s_vg[0][threadIdx.x];
  {
    const uint4 *src = (const uint4 *)(vgrid + 128 * (size_t)t);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const uint4 v = src[q];
      vcol[64 * (4 * q)] = v.x;
      vcol[64 * (4 * q + 1)] = v.y;
      vcol[64 * (4 * q + 2)] = v.z;
      vcol[64 * (4 * q + 3)] = v.w;
    }
  }
  auto vcell = [&](uint32_t i, uint32_t j) {
    const uint32_t c = i * 16u + j;
    return (float)((vcol[64 * (c >> 2)] >> (8u * (c & 3u))) & 0xFFu) * (1.0f / 255.0f);
  };
  const v3 en = V3(EP.normal[0], EP.normal[1], EP.normal[2]);
  v3 acc = V3(0.0f, 0.0f, 0.0f);
  // the cut walk's stack in LDS, one column per lane, node indices as u16 (the host caps the tree at 65535 nodes;
  // a child index beyond that becomes 0xFFFF, which the walk skips)
  __shared__ uint16_t s_stack[STACK][64];
  uint16_t *stack = &s_stack[0][threadIdx.x];
  stack[0] = (uint16_t)min(EP.root, 0xFFFFu);
  uint32_t sp = 1;
  while (sp > 0) {
    sp -= 1;
    const uint32_t ni = stack[64 * sp];
    if (ni >= n_nodes) continue;                      // a malformed tree: the node does not exist
    const float4 *e = emit_nodes + 6 * (size_t)ni;
    const float4 e0 = e[0], e1 = e[1], e2 = e[2], e3 = e[3];
    const uint32_t clo = __float_as_uint(e2.w);
    bool use;
    if (clo == NO_CHILD || EP.mode == 0u) {
      use = clo == NO_CHILD;
    } else {
      const v3 d = sub3(P, V3(e0.x, e0.y, e0.z));
      use = dot3(d, d) > EP.theta * EP.theta * e0.w * e0.w;
    }
    if (!use) {
      if (sp + 2u <= STACK) {
        stack[64 * sp] = (uint16_t)min(clo, 0xFFFFu);
        stack[64 * (sp + 1)] = (uint16_t)min(__float_as_uint(e3.w), 0xFFFFu);
        sp += 2;
      }
      continue;
    }
    const v3 d = sub3(V3(e0.x, e0.y, e0.z), P);
    const float d2 = dot3(d, d);
    if (d2 < 1e-8f) continue;
    const float sd = sqrtf(d2), isd = __builtin_amdgcn_rcpf(sd);
    // shiny receivers divide exactly: the ASG phase differences two nearly equal directions near the mirror one
    const v3 ww = SHINY ? V3(d.x / sd, d.y / sd, d.z / sd) : V3(d.x * isd, d.y * isd, d.z * isd);
    const v3 wi = xf_dir_inv(xf, ww);
    const float vi = vgrid_lookup_f(vcell, wi);
    if (vi <= 1e-7f) continue;
    const float fp = footprint(nd, detC, wi);
    // intensity(e, -ww) (emissive.wgsl:123-130)
    const v3 w = V3(-ww.x, -ww.y, -ww.z);
    float s = 1.0f;
    if (EP.one_sided != 0u || (__float_as_uint(e1.w) & 2u) != 0u) {
      const float c = dot3(w, en);
      if (EP.edge_soft > 0.0f) {
        const float t = clampf_((c - -EP.edge_soft) / (EP.edge_soft - -EP.edge_soft), 0.0f, 1.0f);
        s = t * t * (3.0f - 2.0f * t);
      } else {
        s = c > 0.0f ? 1.0f : 0.0f;
      }
    }
    const float q = dot3(w, sym_mul(V3(e1.x, e1.y, e1.z), V3(e2.x, e2.y, e2.z), w));
    const float ik = s * EP.scale * sqrtf(q > 0.0f ? q : 0.0f);
    const float k = qdiv(fp, d2);
    v3 light = V3(e3.x * ik * vi * k, e3.y * ik * vi * k, e3.z * ik * vi * k);
    if constexpr (SHINY) {
      if (sh.ks > 0.0f) light = mul3s(light, shiny_lobe(sh, nd, detC, vcell, xf, wi, ww, d2, e));
    }
    acc = add3(acc, light);
  }
  const float T = lamp_T ? lamp_T[g] : 1.0f;
  const float tk = T * iso_norm[t];
  v3 c = V3(C.x * acc.x * tk, C.y * acc.y * tk, C.z * acc.z * tk);
  if (c0_out) c0_out[g] = make_float4(c.x, c.y, c.z, 0.0f);
  if ((mflags & 1u) != 0u && ((mflags & 4u) == 0u || dot3(xf_dir_inv(xf, wo), Rq.c0) > 0.0f)) {
    c = add3(c, V3(unpack_half(mat[6] & 0xFFFFu), unpack_half(mat[6] >> 16), unpack_half(mat[7] & 0xFFFFu)));
  }
  colour[g] = make_float4(c.x, c.y, c.z, 1.0f);
}

// The same shading with one cut walk per wave instead of one per lane, for trees whose full walk never holds more
// than 16 stack entries (then no cut can either: a cut's stack is a subset of the full walk's when it reaches a
// node). The wave walks the union of its lanes' cuts in the lanes' own
// depth-first order (children pushed low then high, the high one popped first): a node is expanded when any active
// lane expands it, and a lane that uses a node sits out that node's subtree (it is active again once the walk pops
// below the node's stack slot). Each lane so meets its own used nodes in its own order and sums them exactly as
// sl_shade_emit does. The node index, the node's words and the stack (16 x 16-bit entries in four 64-bit words) are
// the same for every lane, so they live in scalar registers; only the V-grid uses LDS (8 KiB per wave). Inputs in
// visiting order as for sl_shade_emit.
template <bool SHINY>
__global__ void __launch_bounds__(64)
sl_shade_emit_wave(FrameUniformsDev U, EmitParamsDev EP, const uint32_t *__restrict__ order,
                   const float4 *__restrict__ emit_nodes, uint32_t n_nodes,
                   const float4 *__restrict__ r0, const float4 *__restrict__ r1, const float4 *__restrict__ r2,
                   const float4 *__restrict__ r3, uint32_t st, const float4 *__restrict__ nodes,
                   const uint8_t *__restrict__ vgrid, const float4 *__restrict__ parts,
                   const float *__restrict__ lamp_T, const uint32_t *__restrict__ materials,
                   const float *__restrict__ iso_norm, float4 *__restrict__ colour, float4 *__restrict__ c0_out) {
  const uint32_t t = blockIdx.x * 64u + threadIdx.x;
  const bool live = t < U.gaussian_count;
  const uint32_t ti = live ? t : 0u, g = order[ti];   // a lane past the count reads row 0 and sits out
  const float4 A = r0[ti], Q = sl_quat(r1, st, ti), S = sl_scale(r2, st, ti), C = sl_albedo(r1, r2, r3, st, ti);
  const uint32_t mi = __float_as_uint(C.w) & 0xFFFFu;
  const uint32_t mmax = max(U.material_count, 1u) - 1u;
  const uint32_t *mat = materials + 8 * (size_t)min(mi, mmax);
  const uint32_t mflags = U.material_count ? mat[5] : 0u;
  bool lane_on = live;
  if (live && EP.count == 0u && (mflags & 1u) == 0u) {
    colour[g] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    if (c0_out) c0_out[g] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    lane_on = false;
  }
  const Xf xf = load_xf(parts, __float_as_uint(S.w) & 0xFFFFu);
  const v3 P = xf_point(xf, V3(A.x, A.y, A.z));
  const m3 Rq = quat_matrix(Q);
  const float s3 = S.x * S.y * S.z;
  const float detC = s3 * s3;
  const v3 wo = norm3(sub3(V3(U.eye[0], U.eye[1], U.eye[2]), P));
  const NodeL nd = load_node(nodes, ti);
  Shiny sh{};
  if constexpr (SHINY) sh = shiny_receiver(mat, mflags, EP.n5, Rq, xf_dir_inv(xf, wo), iso_norm[ti]);
  __shared__ uint32_t s_vg[32][64];                    // as in sl_shade_emit
  uint32_t *vcol = &s_vg[0][threadIdx.x];
  {
    const uint4 *src = (const uint4 *)(vgrid + 128 * (size_t)ti);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const uint4 v = src[q];
      vcol[64 * (4 * q)] = v.x;
      vcol[64 * (4 * q + 1)] = v.y;
      vcol[64 * (4 * q + 2)] = v.z;
      vcol[64 * (4 * q + 3)] = v.w;
    }
  }
  auto vcell = [&](uint32_t i, uint32_t j) {
    const uint32_t c = i * 16u + j;
    return (float)((vcol[64 * (c >> 2)] >> (8u * (c & 3u))) & 0xFFu) * (1.0f / 255.0f);
  };
  const v3 en = V3(EP.normal[0], EP.normal[1], EP.normal[2]);
  v3 acc = V3(0.0f, 0.0f, 0.0f);
  // the wave's stack: entry 0 (the top) in the low 16 bits of w0
  uint64_t w0 = min(EP.root, 0xFFFFu), w1 = 0, w2 = 0, w3 = 0;
  uint32_t sp = 1;
  uint32_t floor_at = 0xFFFFFFFFu;                     // this lane sits out pops at stack slots >= floor_at
  while (sp > 0) {
    const uint32_t ni = (uint32_t)(w0 & 0xFFFFu);
    w0 = (w0 >> 16) | (w1 << 48);
    w1 = (w1 >> 16) | (w2 << 48);
    w2 = (w2 >> 16) | (w3 << 48);
    w3 >>= 16;
    sp -= 1;
    const uint32_t at = sp;                            // the popped node's slot
    const bool act = lane_on && at < floor_at;
    if (act) floor_at = 0xFFFFFFFFu;
    if (__ballot(act) == 0ull) continue;
    if (ni >= n_nodes) continue;                      // a malformed tree: the node does not exist
    const float4 *e = emit_nodes + 6 * (size_t)ni;
    const float4 e0 = e[0], e1 = e[1], e2 = e[2], e3 = e[3];
    const uint32_t clo = __float_as_uint(e2.w);
    bool use;
    if (clo == NO_CHILD || EP.mode == 0u) {
      use = clo == NO_CHILD;
    } else {
      const v3 d = sub3(P, V3(e0.x, e0.y, e0.z));
      use = dot3(d, d) > EP.theta * EP.theta * e0.w * e0.w;
    }
    if (__ballot(act && !use) != 0ull && at + 2u <= 16u) {
      const uint64_t lo = min(clo, 0xFFFFu), hi = min(__float_as_uint(e3.w), 0xFFFFu);
      w3 = (w3 << 32) | (w2 >> 32);
      w2 = (w2 << 32) | (w1 >> 32);
      w1 = (w1 << 32) | (w0 >> 32);
      w0 = (w0 << 32) | (lo << 16) | hi;               // hi on top, as the per-lane walk pushes it last
      sp += 2;
    }
    const bool contrib = act && use;
    if (contrib) floor_at = at;                        // sit out this node's subtree
    if (__ballot(contrib) == 0ull) continue;
    if (contrib) {
      const v3 d = sub3(V3(e0.x, e0.y, e0.z), P);
      const float d2 = dot3(d, d);
      if (d2 >= 1e-8f) {
        const float sd = sqrtf(d2), isd = __builtin_amdgcn_rcpf(sd);
        const v3 ww = SHINY ? V3(d.x / sd, d.y / sd, d.z / sd) : V3(d.x * isd, d.y * isd, d.z * isd);   // as above
        const v3 wi = xf_dir_inv(xf, ww);
        const float vi = vgrid_lookup_f(vcell, wi);
        if (vi > 1e-7f) {
          const float fp = footprint(nd, detC, wi);
          // intensity(e, -ww) (emissive.wgsl:123-130)
          const v3 w = V3(-ww.x, -ww.y, -ww.z);
          float s = 1.0f;
          if (EP.one_sided != 0u || (__float_as_uint(e1.w) & 2u) != 0u) {
            const float c = dot3(w, en);
            if (EP.edge_soft > 0.0f) {
              const float tt = clampf_((c - -EP.edge_soft) / (EP.edge_soft - -EP.edge_soft), 0.0f, 1.0f);
              s = tt * tt * (3.0f - 2.0f * tt);
            } else {
              s = c > 0.0f ? 1.0f : 0.0f;
            }
          }
          const float q = dot3(w, sym_mul(V3(e1.x, e1.y, e1.z), V3(e2.x, e2.y, e2.z), w));
          const float ik = s * EP.scale * sqrtf(q > 0.0f ? q : 0.0f);
          const float k = qdiv(fp, d2);
          v3 light = V3(e3.x * ik * vi * k, e3.y * ik * vi * k, e3.z * ik * vi * k);
          if constexpr (SHINY) {
            if (sh.ks > 0.0f) light = mul3s(light, shiny_lobe(sh, nd, detC, vcell, xf, wi, ww, d2, e));
          }
          acc = add3(acc, light);
        }
      }
    }
  }
  if (!lane_on) return;
  const float T = lamp_T ? lamp_T[g] : 1.0f;
  const float tk = T * iso_norm[ti];
  v3 c = V3(C.x * acc.x * tk, C.y * acc.y * tk, C.z * acc.z * tk);
  if (c0_out) c0_out[g] = make_float4(c.x, c.y, c.z, 0.0f);
  if ((mflags & 1u) != 0u && ((mflags & 4u) == 0u || dot3(xf_dir_inv(xf, wo), Rq.c0) > 0.0f)) {
    c = add3(c, V3(unpack_half(mat[6] & 0xFFFFu), unpack_half(mat[6] >> 16), unpack_half(mat[7] & 0xFFFFu)));
  }
  colour[g] = make_float4(c.x, c.y, c.z, 1.0f);
}

// ---- opacity maps, shared: resolve, mips (light_shadow.wgsl:290-321, sun_map.wgsl:228-261) ----------------------

// layer l's texel i -> (alpha, alpha mu, alpha m2, 0); layers with dirty_mask bit 0 (bit l) skipped
extern "C" __global__ void __launch_bounds__(256)
sl_opacity_resolve(const uint32_t *__restrict__ acc, float4 *__restrict__ mip, uint32_t n_layers, uint32_t texels,
                   uint32_t layer_texels, float fix, const uint32_t *__restrict__ dirty) {
  const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
  if (i >= (size_t)texels * n_layers) return;
  const uint32_t l = (uint32_t)(i / texels);
  if (!dirty[l]) return;
  const float s0 = (float)acc[i * 3] / fix;
  float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (s0 > 1e-5f) {
    // 1 - e^-s0 rounded once from double: the f32 form cancels for small s0 (a 1-ulp exp difference becomes
    // 1e-3 relative at s0 = 1.5e-5, and the moments divide by s0)
    const float a = (float)(-expm1(-(double)s0));
    o = make_float4(a, a * (float)acc[i * 3 + 1] / (fix * s0), a * (float)acc[i * 3 + 2] / (fix * s0), 0.0f);
  }
  mip[(size_t)l * layer_texels + (i - (size_t)l * texels)] = o;
}

extern "C" __global__ void __launch_bounds__(256)
sl_opacity_down(float4 *__restrict__ mip, uint32_t n_layers, uint32_t rk, uint32_t off_src, uint32_t off_dst,
                uint32_t layer_texels, const uint32_t *__restrict__ dirty) {
  const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
  const size_t n = (size_t)rk * rk;
  if (i >= n * n_layers) return;
  const uint32_t l = (uint32_t)(i / n);
  if (!dirty[l]) return;
  const uint32_t j = (uint32_t)(i - (size_t)l * n);
  const uint32_t y = j / rk, x = j - y * rk, pr = rk * 2u;
  const size_t src = (size_t)l * layer_texels + off_src + (size_t)(2u * y) * pr + 2u * x;
  const float4 a = mip[src], b = mip[src + 1], c = mip[src + pr], d = mip[src + pr + 1];
  mip[(size_t)l * layer_texels + off_dst + j] =
      make_float4(0.25f * (((a.x + b.x) + c.x) + d.x), 0.25f * (((a.y + b.y) + c.y) + d.y),
                  0.25f * (((a.z + b.z) + c.z) + d.z), 0.25f * (((a.w + b.w) + c.w) + d.w));
}

// The same resolve and mips, fused: one 256-thread workgroup per 32x32 texel region of a layer (blockIdx.y is the
// layer; small workgroups, so they find CUs beside another stream's kernels). Each thread resolves a 2x2 quad of
// level 0 and averages it into level 1; levels 2-5 of the region come from LDS. Same arithmetic and order as
// sl_opacity_resolve / sl_opacity_down, so the levels are bit-identical. Needs res a multiple of 32; levels past 5
// are left to sl_opacity_down / sl_opacity_tail.
struct LevelOffs { uint32_t off[16]; };
__device__ __forceinline__ float4 resolve_texel(const uint32_t *__restrict__ acc, size_t i, float fix) {
  const float s0 = (float)acc[i * 3] / fix;
  float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (s0 > 1e-5f) {
    const float a = (float)(-expm1(-(double)s0));
    o = make_float4(a, a * (float)acc[i * 3 + 1] / (fix * s0), a * (float)acc[i * 3 + 2] / (fix * s0), 0.0f);
  }
  return o;
}
__device__ __forceinline__ float4 avg4(float4 a, float4 b, float4 c, float4 d) {
  return make_float4(0.25f * (((a.x + b.x) + c.x) + d.x), 0.25f * (((a.y + b.y) + c.y) + d.y),
                     0.25f * (((a.z + b.z) + c.z) + d.z), 0.25f * (((a.w + b.w) + c.w) + d.w));
}
// levels k0 .. k1 - 1 of a square block held in LDS at `src` (side `side`, the level k0 - 1 values), written to
// lm at level offsets lo.off with row length res >> k and the block's origin (bx, by) in level-k0-1 texels / 2^..
__device__ __forceinline__ void lds_levels(float4 *src, float4 *dst, uint32_t side, uint32_t k0, uint32_t k1,
                                           uint32_t res, uint32_t bx, uint32_t by, const LevelOffs &lo, float4 *lm) {
  for (uint32_t k = k0; k < k1; k++) {
    __syncthreads();
    const uint32_t ns = side >> 1;
    for (uint32_t j = threadIdx.x; j < ns * ns; j += blockDim.x) {
      const uint32_t tx = j % ns, ty = j / ns;
      const float4 *p = src + (2u * ty) * side + 2u * tx;
      const float4 w = avg4(p[0], p[1], p[side], p[side + 1]);
      dst[j] = w;
      lm[lo.off[k] + (size_t)(by * ns + ty) * (res >> k) + (bx * ns + tx)] = w;
    }
    float4 *t = src;
    src = dst;
    dst = t;
    side = ns;
  }
}
// touched (when given): one bit per 16x16 tile of the layers, set by the sums for every tile they gave records; a
// clear bit's tile holds the zero sums its layer was cleared to, whose resolve is zero, so its sums are not read.
extern "C" __global__ void __launch_bounds__(256)
sl_opacity_pyramid(const uint32_t *__restrict__ acc, float4 *__restrict__ mip, uint32_t res, uint32_t levels,
                   LevelOffs lo, uint32_t layer_texels, float fix, const uint32_t *__restrict__ dirty,
                   const uint32_t *__restrict__ touched) {
  __shared__ float4 s1[16 * 16];
  __shared__ float4 s2[8 * 8];
  const uint32_t l = blockIdx.y;
  if (!dirty[l]) return;
  const uint32_t regions_x = res / 32u, rx = blockIdx.x % regions_x, ry = blockIdx.x / regions_x;
  const uint32_t qx = threadIdx.x & 15u, qy = threadIdx.x >> 4;
  const uint32_t x = rx * 32u + 2u * qx, y = ry * 32u + 2u * qy;
  const uint32_t *lacc = acc + (size_t)l * res * res * 3;
  float4 *lm = mip + (size_t)l * layer_texels;
  const size_t i00 = (size_t)y * res + x;
  bool live = true;                                  // this thread's 2x2 texels lie in one 16x16 tile
  if (touched != nullptr) {
    const uint32_t tx = res / 16u, tile = (l * tx + (y >> 4)) * tx + (x >> 4);
    live = ((touched[tile >> 5] >> (tile & 31u)) & 1u) != 0u;
  }
  const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float4 a = live ? resolve_texel(lacc, i00, fix) : z, b = live ? resolve_texel(lacc, i00 + 1, fix) : z;
  const float4 c = live ? resolve_texel(lacc, i00 + res, fix) : z;
  const float4 d = live ? resolve_texel(lacc, i00 + res + 1, fix) : z;
  lm[i00] = a;
  lm[i00 + 1] = b;
  lm[i00 + res] = c;
  lm[i00 + res + 1] = d;
  if (levels < 2u) return;
  const float4 v = avg4(a, b, c, d);
  lm[lo.off[1] + (size_t)(y >> 1) * (res >> 1) + (x >> 1)] = v;
  s1[threadIdx.x] = v;
  lds_levels(s1, s2, 16u, 2u, min(levels, 6u), res, rx, ry, lo, lm);
}

// Levels k_first .. levels - 1 of every dirty layer, one workgroup per layer: level k_first (side at most 32) from
// level k_first - 1 in mip, the rest through LDS. Same arithmetic and order as sl_opacity_down.
extern "C" __global__ void __launch_bounds__(256)
sl_opacity_tail(float4 *__restrict__ mip, uint32_t res, uint32_t k_first, uint32_t levels, LevelOffs lo,
                uint32_t layer_texels, const uint32_t *__restrict__ dirty) {
  __shared__ float4 sa[32 * 32];
  __shared__ float4 sb[16 * 16];
  const uint32_t l = blockIdx.x;
  if (!dirty[l]) return;
  float4 *lm = mip + (size_t)l * layer_texels;
  const uint32_t side = res >> k_first, pr = side * 2u;
  for (uint32_t j = threadIdx.x; j < side * side; j += blockDim.x) {
    const uint32_t tx = j % side, ty = j / side;
    const float4 *p = lm + lo.off[k_first - 1] + (size_t)(2u * ty) * pr + 2u * tx;
    const float4 w = avg4(p[0], p[1], p[pr], p[pr + 1]);
    sa[j] = w;
    lm[lo.off[k_first] + j] = w;
  }
  lds_levels(sa, sb, side, k_first + 1u, levels, res, 0u, 0u, lo, lm);
}

// ---- opacity-map splats as tiled sums -------------------------------------------------------------------------
//
// The reference splats each Gaussian with one thread looping over its texel box and three u32 atomics per texel
// (light_shadow.wgsl:191-286, sun_map.wgsl:127-219). Here a prep pass turns each (Gaussian, layer) into a 64 B
// record and its box of 16x16 map tiles, the frame path's pair emission and onesweep sort bin the (tile, record)
// pairs by tile without contended atomics, and workgroups sum each tile's records' fixed-point moments in
// registers, one lane per texel, writing the tile once. The per-texel
// arithmetic is the reference's, unchanged; u32 sums are order-independent, so the maps are bit-identical to
// the atomic form whatever order the records arrive in.

constexpr int MT = 16;   // map tile side, texels

// One Gaussian's splat onto one map layer: everything the per-texel sum needs.
struct MapRec {
  float ctx, cty, s00, s01;   // centre in texels; inverse footprint, column 0
  float s10, s11, tau0, zc;   // column 1; peak optical depth; depth at the centre
  float gzx, gzy, sz2;        // depth change per texel; depth variance along the light
  uint32_t layer;
  int32_t x0, x1, y0, y1;     // texel box clamped to the map; x0 > x1 when the record adds nothing
};
static_assert(sizeof(MapRec) == 64, "MapRec");

__device__ __forceinline__ MapRec empty_rec() {
  MapRec r = {};
  r.x0 = 1; r.x1 = 0; r.y0 = 1; r.y1 = 0;
  return r;
}

// A record's tile box in the layers' stacked tile grid (layer l's rows follow layer l - 1's), packed as the frame
// path's pair emission reads it: (tx0 | ty0 << 16, tx1 | ty1 << 16); the grid's tile id ty * tiles_x + tx is then
// l * tiles_x * tiles_y + (local row) * tiles_x + tx, the map tile index. Empty records: (0xffffffff, 0).
__device__ __forceinline__ uint2 map_tbox(const MapRec &r, uint32_t tiles_y) {
  if (r.x0 > r.x1 || r.y0 > r.y1) return make_uint2(0xffffffffu, 0u);
  const uint32_t yb = r.layer * tiles_y;
  return make_uint2((uint32_t)(r.x0 / MT) | ((yb + (uint32_t)(r.y0 / MT)) << 16),
                    (uint32_t)(r.x1 / MT) | ((yb + (uint32_t)(r.y1 / MT)) << 16));
}

constexpr uint32_t MAP_STAGE = 192;   // records staged in LDS per round
constexpr uint32_t MAP_CHUNK = 1024;  // tile-sorted pairs per ticket

// One texel's fixed-point moments of record r, added to (a0, a1, a2) (light_shadow.wgsl:237-284,
// sun_map.wgsl:171-217: the same per-texel arithmetic).
__device__ __forceinline__ void map_texel(const MapRec &r, float dx, float py, float fix, uint32_t &a0, uint32_t &a1,
                                          uint32_t &a2) {
  const v2 dt = V2(dx, py - r.cty);
  const m2 Si = {V2(r.s00, r.s01), V2(r.s10, r.s11)};
  const float q = dot2(dt, m2v(Si, dt));
  if (!(q <= 9.0f)) return;                         // outside 3 sigma, or a NaN footprint: adds nothing
  float t = r.tau0 * exp_map_in(-0.5f * q);
  t = t < 64.0f ? t : 64.0f;
  const float t_fix = t * fix;
  if (t_fix < 0.5f) return;
  float z = r.zc + dot2(V2(r.gzx, r.gzy), dt);
  z = z > 0.0f ? z : 0.0f;
  a0 += u32_sat(t_fix + 0.5f);
  a1 += u32_sat(t_fix * z + 0.5f);
  a2 += u32_sat(t_fix * (z * z + r.sz2) + 0.5f);
}

// Whether record r may add something (q <= 9) at a texel of the texel-centre rectangle [xa, xb] x [ya, yb]
// (inside its box): false only when the minimum of its quadratic form over the rectangle exceeds 9 by a margin.
// The rectangle's corners are the same float differences px - ctx the sums compute, and rounding is monotonic, so
// every texel's offsets lie inside it; the margin (1e-4 of the largest term magnitude on the rectangle, plus 1e-3)
// is some 100x the float rounding of both the per-texel q and this minimum. true whenever the form is not positive
// definite or not finite, so a rectangle with a contributing texel is never ruled out.
struct MapForm {
  float a, b, c, nba, nbc;   // -b/a, -b/c: where the form's minimum lies along a row (v fixed) or a column (u fixed)
  bool pd;
  __device__ __forceinline__ explicit MapForm(const MapRec &r)
      : a(r.s00), b(0.5f * (r.s01 + r.s10)), c(r.s11) {
    pd = a > 0.0f && c > 0.0f && a * c - b * b > 0.0f && a < 1e30f && c < 1e30f;
    nba = pd ? -b * __builtin_amdgcn_rcpf(a) : 0.0f;
    nbc = pd ? -b * __builtin_amdgcn_rcpf(c) : 0.0f;
  }
};
__device__ __forceinline__ bool map_rect_may_hit(const MapRec &r, const MapForm &f, int xa, int xb, int ya, int yb) {
  if (!f.pd) return true;
  const float a = f.a, b = f.b, c = f.c;
  const float u0 = ((float)xa + 0.5f) - r.ctx, u1 = ((float)xb + 0.5f) - r.ctx;
  const float v0 = ((float)ya + 0.5f) - r.cty, v1 = ((float)yb + 0.5f) - r.cty;
  auto q = [&](float u, float v) { return (a * u * u + 2.0f * b * u * v) + c * v * v; };
  float m;
  if (u0 <= 0.0f && u1 >= 0.0f && v0 <= 0.0f && v1 >= 0.0f) {
    m = 0.0f;                                            // the centre lies inside
  } else {
    // the minimum lies on an edge: on u = const the best v is -b u / c, on v = const the best u is -b v / a (the
    // reciprocals' rounding moves a candidate by ulps, which raises q there by far less than the margin below)
    m = fminf(fminf(q(u0, fminf(fmaxf(f.nbc * u0, v0), v1)), q(u1, fminf(fmaxf(f.nbc * u1, v0), v1))),
              fminf(q(fminf(fmaxf(f.nba * v0, u0), u1), v0), q(fminf(fmaxf(f.nba * v1, u0), u1), v1)));
  }
  const float um = fmaxf(fabsf(u0), fabsf(u1)), vm = fmaxf(fabsf(v0), fabsf(v1));
  const float big = a * um * um + (fabsf(r.s01) + fabsf(r.s10)) * um * vm + c * vm * vm;
  return m <= 9.0f + 1e-4f * big + 1e-3f;
}

// Lane layouts of a 16x16 tile's 256 texels as four units of 64 (one wave each): 0 bands of 16x4 (unit k: rows 4k..4k+3;
// lane: column lane & 15, row lane >> 4 of the band), 1 the transpose, columns of 4x16, 2 quads of 8x8 (unit k:
// columns 8 (k & 1).., rows 8 (k >> 1)..; lane: column lane & 7, row lane >> 3). Tile-local texel of a lane in unit k:
template <int LAY>
__device__ __forceinline__ void map_unit_texel(uint32_t lane, int k, int &cx, int &cy) {
  if (LAY == 0) {
    cx = (int)(lane & 15u);
    cy = 4 * k + (int)(lane >> 4);
  } else if (LAY == 1) {
    cx = 4 * k + (int)(lane >> 4);
    cy = (int)(lane & 15u);
  } else {
    cx = 8 * (k & 1) + (int)(lane & 7u);
    cy = 8 * (k >> 1) + (int)(lane >> 3);
  }
}

// The layout record r is evaluated in over the 16x16 tile at texel (tx0, ty0), and which of its 4 units r may touch:
// mask bits 0-3, layout in bits 4-5. The layout is the one whose units the record's box (within the tile) overlaps
// fewest (bands on ties); each overlapped unit is kept unless the rectangle test rules it out.
__device__ __forceinline__ uint32_t map_unit_mask(const MapRec &r, int tx0, int ty0) {
  const int xa = max(r.x0 - tx0, 0), xb = min(r.x1 - tx0, 15), ya = max(r.y0 - ty0, 0), yb = min(r.y1 - ty0, 15);
  if (xa > xb || ya > yb) return 0u;
  const int nb = (yb >> 2) - (ya >> 2) + 1, nv = (xb >> 2) - (xa >> 2) + 1;
  const int nq = ((yb >> 3) - (ya >> 3) + 1) * ((xb >> 3) - (xa >> 3) + 1);
  const int lay = (nv < nb && nv <= nq) ? 1 : (nq < nb && nq < nv) ? 2 : 0;
  const int nu = lay == 0 ? nb : lay == 1 ? nv : nq;
  if (nu == 1) {                                     // one unit holds the box: the tile's own test kept it
    const int k = lay == 0 ? ya >> 2 : lay == 1 ? xa >> 2 : 2 * (ya >> 3) + (xa >> 3);
    return (1u << k) | ((uint32_t)lay << 4);
  }
  const MapForm f(r);
  uint32_t mask = 0;
#pragma unroll 1
  for (int k = 0; k < 4; k++) {
    const int ux0 = lay == 0 ? 0 : lay == 1 ? 4 * k : 8 * (k & 1);
    const int uy0 = lay == 0 ? 4 * k : lay == 1 ? 0 : 8 * (k >> 1);
    const int uw = lay == 0 ? 15 : lay == 1 ? 3 : 7, uh = lay == 0 ? 3 : lay == 1 ? 15 : 7;
    const int rxa = max(xa, ux0), rxb = min(xb, ux0 + uw), rya = max(ya, uy0), ryb = min(yb, uy0 + uh);
    if (rxa <= rxb && rya <= ryb && map_rect_may_hit(r, f, tx0 + rxa, tx0 + rxb, ty0 + rya, ty0 + ryb)) mask |= 1u << k;
  }
  return mask | ((uint32_t)lay << 4);
}

// Which tiles of record r's tile box (at most 8x8) it may touch: bit (row * box width + column), the order
// sl_pairs_emit walks a box in. Empty records: 0.
__device__ __forceinline__ uint64_t map_tile_mask(const MapRec &r) {
  if (r.x0 > r.x1 || r.y0 > r.y1) return 0ull;
  const MapForm f(r);
  const int tx0 = r.x0 / MT, tx1 = r.x1 / MT, ty0 = r.y0 / MT, ty1 = r.y1 / MT, bw = tx1 - tx0 + 1;
  uint64_t mask = 0;
  for (int ty = ty0; ty <= ty1; ty++)
    for (int tx = tx0; tx <= tx1; tx++) {
      const int xa = max(r.x0, tx * MT), xb = min(r.x1, tx * MT + MT - 1);
      const int ya = max(r.y0, ty * MT), yb = min(r.y1, ty * MT + MT - 1);
      if (map_rect_may_hit(r, f, xa, xb, ya, yb)) mask |= 1ull << ((ty - ty0) * bw + (tx - tx0));
    }
  return mask;
}

// The wave's staged records [g0, g1) of one layout, each over its units of the tile, skipping the units its mask
// rules out; the wave's sums are then added into s_acc (integer adds, order-independent).
template <int LAY>
__device__ __forceinline__ void map_group_sums(const MapRec *w_rec, uint32_t g0, uint32_t g1, uint32_t lane, int tx0,
                                               int ty0, float fix, uint32_t *s_acc) {
  if (g0 >= g1) return;                              // the same for every lane
  uint32_t a[4][3] = {};
  for (uint32_t j = g0; j < g1; j++) {
    const MapRec r = w_rec[j];
    const uint32_t mask = r.layer;                   // the same for every lane
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (!((mask >> k) & 1u)) continue;
      int cx, cy;
      map_unit_texel<LAY>(lane, k, cx, cy);
      const int x = tx0 + cx, y = ty0 + cy;
      if (x >= r.x0 && x <= r.x1 && y >= r.y0 && y <= r.y1)
        map_texel(r, ((float)x + 0.5f) - r.ctx, (float)y + 0.5f, fix, a[k][0], a[k][1], a[k][2]);
    }
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    int cx, cy;
    map_unit_texel<LAY>(lane, k, cx, cy);
#pragma unroll
    for (int c = 0; c < 3; c++) atomicAdd(&s_acc[c * 256 + cy * 16 + cx], a[k][c]);
  }
}

// Per-tile sums of the records list[s0, s1) of the 16x16 tile at texel (tx0, ty0). Wave w takes records w, w + 4, ...
// of the run and stages MAP_WSTAGE of them at a time in its own part of LDS (one load per lane), each with its layout
// and unit mask for this tile (in place of its layer, which the sums do not read), grouped by layout (slots ranked
// by ballot); each group is then evaluated in its layout (map_group_sums). A wave reads only what it staged, so the
// rounds need no workgroup barrier. On return s_acc[c][t] holds texel t's sum c (t = row * 16 + column) and the
// caller clears it after reading.
constexpr uint32_t MAP_WSTAGE = MAP_STAGE / 4u;
__device__ __forceinline__ void map_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ void map_tile_sums(const MapRec *__restrict__ recs, const uint32_t *__restrict__ list,
                                              uint32_t s0, uint32_t s1, int tx0, int ty0, float fix, MapRec *s_rec,
                                              uint32_t *s_acc) {
  static_assert(MAP_WSTAGE <= 64u, "one record per lane");
  const uint32_t lane = threadIdx.x & 63u, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  MapRec *w_rec = s_rec + wave * MAP_WSTAGE;
  const uint32_t n = s1 - s0, mine = n > wave ? (n - wave + 3u) / 4u : 0u;
  __syncthreads();                                   // the last tile's sums are read and cleared
  for (uint32_t b = 0; b < mine; b += MAP_WSTAGE) {
    const uint32_t nb = min(MAP_WSTAGE, mine - b);
    const bool act = lane < nb;
    MapRec r = act ? recs[list[s0 + wave + 4u * (b + lane)]] : empty_rec();
    const uint32_t m = act ? map_unit_mask(r, tx0, ty0) : 0u;
    const uint32_t lay = act ? m >> 4 : 3u;
    r.layer = m & 15u;
    const uint64_t b0 = __ballot(lay == 0u), b1 = __ballot(lay == 1u), b2 = __ballot(lay == 2u);
    const uint32_t c0 = (uint32_t)__popcll(b0), c1 = (uint32_t)__popcll(b1);
    if (act) {
      const uint64_t bm = lay == 0u ? b0 : lay == 1u ? b1 : b2;
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(bm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u));
      w_rec[(lay >= 1u ? c0 : 0u) + (lay >= 2u ? c1 : 0u) + rank] = r;
    }
    map_wave_sync();
    map_group_sums<0>(w_rec, 0u, c0, lane, tx0, ty0, fix, s_acc);
    map_group_sums<1>(w_rec, c0, c0 + c1, lane, tx0, ty0, fix, s_acc);
    map_group_sums<2>(w_rec, c0 + c1, nb, lane, tx0, ty0, fix, s_acc);
    map_wave_sync();                                 // read before the next round overwrites them
  }
  __syncthreads();
}

// One workgroup per chunk of MAP_CHUNK consecutive entries of the tile-sorted pair list (the grid covers the
// pair capacity; workgroups past the pair count exit at once). An ordinary grid rather than a persistent one, so
// kernels of another stream get CUs as these workgroups retire. Inside a chunk, each run of equal tile ids is one
// tile's records; a tile whose run lies inside one chunk is stored, a tile split across chunks (a crowded tile is
// shared by several workgroups) is merged with integer atomic adds (order-independent). acc of the dirty layers is
// zero on entry; tiles with no records are not touched. stats[0] = pairs, stats[1] += tiles with records (zero on
// entry). LDS is under 20 KiB per workgroup, so 8 workgroups (32 waves) fit a CU's 160 KiB. KIND: 0 the emitter
// map (fixed point x65536, light_shadow.wgsl), 1 the sun maps (x4096, sun_map.wgsl), one instance each.
template <uint32_t KIND>
__global__ void __launch_bounds__(256, 6)
sl_map_accumulate(const MapRec *__restrict__ recs, const uint32_t *__restrict__ keys,
                  const uint32_t *__restrict__ list, const FrameCounters *counters, uint32_t res, uint32_t tiles_x,
                  uint32_t tpl, uint32_t *__restrict__ acc, uint32_t *__restrict__ stats, uint32_t *touched) {
  constexpr float fix = KIND == 0 ? 65536.0f : 4096.0f;
  __shared__ MapRec s_rec[MAP_STAGE];
  __shared__ uint32_t s_keys[MAP_CHUNK];
  __shared__ uint32_t s_acc[3 * 256];
  const uint32_t total = __hip_atomic_load(&counters->pairs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (blockIdx.x == 0 && threadIdx.x == 0) stats[0] = total;
  const uint32_t lo = blockIdx.x * MAP_CHUNK;
  if (lo >= total) return;
  const uint32_t hi = min(lo + MAP_CHUNK, total), n = hi - lo;
  const uint32_t prev = lo > 0u ? keys[lo - 1u] : 0xffffffffu, next = hi < total ? keys[hi] : 0xffffffffu;
  for (uint32_t c = 0; c < 3; c++) s_acc[c * 256 + threadIdx.x] = 0u;
  for (uint32_t j = threadIdx.x; j < n; j += 256u) s_keys[j] = keys[lo + j];
  __syncthreads();
  uint32_t used = 0;
  for (uint32_t s0 = 0; s0 < n;) {
    const uint32_t tile = s_keys[s0];
    uint32_t lo2 = s0 + 1u, hi2 = n;                   // the run's end: the first key past `tile`
    while (lo2 < hi2) {
      const uint32_t mid = (lo2 + hi2) >> 1;
      if (s_keys[mid] <= tile) lo2 = mid + 1u; else hi2 = mid;
    }
    const uint32_t s1 = lo2;
    const bool starts = s0 > 0u || prev != tile, ends = s1 < n || next != tile;
    used += starts;
    if (threadIdx.x == 0u) atomicOr(touched + (tile >> 5), 1u << (tile & 31u));
    const uint32_t layer = tile / tpl, rem = tile - layer * tpl, ty = rem / tiles_x, tx = rem - ty * tiles_x;
    map_tile_sums(recs, list, lo + s0, lo + s1, (int)(tx * MT), (int)(ty * MT), fix, s_rec, s_acc);
    // texel threadIdx.x of the tile: its three sums, then cleared for the next tile (read and cleared by this
    // thread only; the next adds come after at least one more barrier)
    const uint32_t x = tx * MT + (threadIdx.x & 15u), y = ty * MT + (threadIdx.x >> 4);
    const uint32_t a0 = s_acc[threadIdx.x], a1 = s_acc[256 + threadIdx.x], a2 = s_acc[512 + threadIdx.x];
    s_acc[threadIdx.x] = 0u;
    s_acc[256 + threadIdx.x] = 0u;
    s_acc[512 + threadIdx.x] = 0u;
    if (x < res && y < res) {
      uint32_t *k = acc + ((size_t)layer * res * res + (size_t)y * res + (size_t)x) * 3;
      if (starts && ends) {
        k[0] = a0;
        k[1] = a1;
        k[2] = a2;
      } else {
        if (a0) atomicAdd(k, a0);
        if (a1) atomicAdd(k + 1, a1);
        if (a2) atomicAdd(k + 2, a2);
      }
    }
    s0 = s1;
  }
  if (threadIdx.x == 0 && used) atomicAdd(stats + 1, used);
}

namespace lt {
__device__ __forceinline__ float4 mip_at(const float4 *mip, const uint32_t *level_off, uint32_t res, uint32_t k, v2 tc) {
  const uint32_t rk = max(res >> k, 1u);
  const float sc = (float)(1u << k);
  const float cx = clampf_(tc.x / sc - 0.5f, 0.0f, (float)(rk - 1u)), cy = clampf_(tc.y / sc - 0.5f, 0.0f, (float)(rk - 1u));
  const uint32_t i0x = (uint32_t)floorf(cx), i0y = (uint32_t)floorf(cy);
  const uint32_t i1x = min(i0x + 1u, rk - 1u), i1y = min(i0y + 1u, rk - 1u);
  const float fx = cx - (float)i0x, fy = cy - (float)i0y;
  const float4 *b = mip + level_off[k];
  const float4 p00 = b[i0y * rk + i0x], p01 = b[i0y * rk + i1x], p10 = b[i1y * rk + i0x], p11 = b[i1y * rk + i1x];
  float4 o;
  o.x = mixf(mixf(p00.x, p01.x, fx), mixf(p10.x, p11.x, fx), fy);
  o.y = mixf(mixf(p00.y, p01.y, fx), mixf(p10.y, p11.y, fx), fy);
  o.z = mixf(mixf(p00.z, p01.z, fx), mixf(p10.z, p11.z, fx), fy);
  o.w = mixf(mixf(p00.w, p01.w, fx), mixf(p10.w, p11.w, fx), fy);
  return o;
}
__device__ __forceinline__ float4 mip_tri(const float4 *mip, const uint32_t *level_off, uint32_t res, uint32_t levels,
                                          float l, v2 tc) {
  const float lc = clampf_(l, 0.0f, (float)(levels - 1u));
  const uint32_t k0 = (uint32_t)floorf(lc);
  const uint32_t k1 = min(k0 + 1u, levels - 1u);
  const float4 a = mip_at(mip, level_off, res, k0, tc), b = mip_at(mip, level_off, res, k1, tc);
  const float t = lc - (float)k0;
  return make_float4(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t), mixf(a.w, b.w, t));
}

// octahedral map coordinates in [-1, 1]^2 around the light's normal (light_shadow.wgsl:64-76)
__device__ __forceinline__ v2 oct_uv(const LSParamsDev &LS, v3 w) {
  const v3 l = V3(dot3(w, V3(LS.t1[0], LS.t1[1], LS.t1[2])), dot3(w, V3(LS.t2[0], LS.t2[1], LS.t2[2])),
                  dot3(w, V3(LS.normal[0], LS.normal[1], LS.normal[2])));
  if (LS.hemi != 0u) {
    const float z = l.z > 0.0f ? l.z : 0.0f;
    float den = (fabsf(l.x) + fabsf(l.y)) + z;
    den = den > 1e-12f ? den : 1e-12f;
    const v2 p = V2(l.x / den, l.y / den);
    return V2(p.x + p.y, p.x - p.y);
  }
  float den = (fabsf(l.x) + fabsf(l.y)) + fabsf(l.z);
  den = den > 1e-12f ? den : 1e-12f;
  v2 p = V2(l.x / den, l.y / den);
  if (l.z < 0.0f)
    p = V2((1.0f - fabsf(p.y)) * (p.x >= 0.0f ? 1.0f : -1.0f), (1.0f - fabsf(p.x)) * (p.y >= 0.0f ? 1.0f : -1.0f));
  return p;
}
__device__ __forceinline__ m2 oct_jac(const LSParamsDev &LS, v3 w, v3 b0, v3 b1) {
  const float e = 1e-3f;
  const v2 u0 = oct_uv(LS, w);
  const v2 ua = oct_uv(LS, norm3(add3(w, mul3s(b0, e))));
  const v2 ub = oct_uv(LS, norm3(add3(w, mul3s(b1, e))));
  return {V2((ua.x - u0.x) / e, (ua.y - u0.y) / e), V2((ub.x - u0.x) / e, (ub.y - u0.y) / e)};
}
}  // namespace lt

// ---- A: emitter light opacity map (light_shadow.wgsl:191-286, 347-392) -------------------------------------------

// the emitter map record of the Gaussian in row g of the given streams and nodes (its own part's layer); z = r + gz . dt
__device__ MapRec lsm_rec(uint32_t g, const LSParamsDev &LS, const float4 *__restrict__ g0,
                          const float4 *__restrict__ g1, const float4 *__restrict__ g2, uint32_t st,
                          const float4 *__restrict__ nodes, const float4 *__restrict__ parts) {
  const float4 A = g0[g], Q = sl_quat(g1, st, g), S4 = sl_scale(g2, st, g);
  const uint32_t p = __float_as_uint(S4.w) & 0xFFFFu;
  if (p >= LS.part_count || p == LS.skip_part || (LS.dirty & (1u << p)) == 0u) return empty_rec();
  const NodeL nd = load_node(nodes, g);
  const Xf xf = load_xf(parts, p);
  const m3 R = part_rot(xf);
  const v3 m = xf_point(xf, V3(A.x, A.y, A.z));
  const v3 center = V3(LS.center[0], LS.center[1], LS.center[2]), nrm = V3(LS.normal[0], LS.normal[1], LS.normal[2]);
  const v3 dv = sub3(m, center);
  const float r = len3(dv);
  if (r < 1e-4f) return empty_rec();
  const v3 w = V3(dv.x / r, dv.y / r, dv.z / r);
  const m3 C = leaf_cov(Q, S4, R);
  if (LS.hemi != 0u && dot3(dv, nrm) < -3.0f * sqrtf(dot3(nrm, m3v(C, nrm)))) return empty_rec();
  const v3 b0 = basis_k(w), b1 = cross3(w, b0);
  const v3 c0 = m3v(C, b0), c1 = m3v(C, b1), cw = m3v(C, w);
  const m2 Cbb = {V2(dot3(b0, c0), dot3(b0, c1)), V2(dot3(b0, c1), dot3(b1, c1))};
  const v2 cbz = V2(dot3(b0, cw), dot3(b1, cw));
  const float czz = dot3(w, cw);
  const m2 Cbb_i = inv2(Cbb);
  float sz2 = czz - dot2(cbz, m2v(Cbb_i, cbz));
  sz2 = sz2 > 1e-12f ? sz2 : 1e-12f;
  float tau0 = nd.kappa * dirm(nd, xf_dir_inv(xf, w)) * sqrtf(2.0f * PI_F * sz2);
  const float half_res = 0.5f * (float)LS.res;
  const m2 J = m2s(oct_jac(LS, w, b0, b1), half_res);
  const m2 Sang = m2s(Cbb, 1.0f / (r * r));
  m2 S = m2mul(m2mul(J, Sang), m2t(J));
  const float det0 = m2det(S);
  S.c0.x += 0.3f;
  S.c1.y += 0.3f;
  const float det1 = m2det(S);
  tau0 *= sqrtf((det0 > 0.0f ? det0 : 0.0f) / det1);
  const m2 Si = inv2(S);
  const v2 cz = m2v(Cbb_i, cbz);
  const v2 gz = m2v(m2t(inv2(J)), V2(cz.x * r, cz.y * r));
  const v2 uv = oct_uv(LS, w);
  const v2 ct = V2((uv.x * 0.5f + 0.5f) * (float)LS.res, (uv.y * 0.5f + 0.5f) * (float)LS.res);
  const int hx = min((int)ceilf(3.0f * sqrtf(S.c0.x)), 48), hy = min((int)ceilf(3.0f * sqrtf(S.c1.y)), 48);
  const int cx = (int)floorf(ct.x), cy = (int)floorf(ct.y);
  if (tau0 * 65536.0f < 0.5f) return empty_rec();   // no texel reaches half a fixed-point unit
  const int res = (int)LS.res;
  MapRec o;
  o.ctx = ct.x; o.cty = ct.y;
  o.s00 = Si.c0.x; o.s01 = Si.c0.y; o.s10 = Si.c1.x; o.s11 = Si.c1.y;
  o.tau0 = tau0; o.zc = r; o.gzx = gz.x; o.gzy = gz.y; o.sz2 = sz2;
  o.layer = p;
  o.x0 = max(cx - hx, 0); o.x1 = min(cx + hx, res - 1);
  o.y0 = max(cy - hy, 0); o.y1 = min(cy + hy, res - 1);
  return o;
}

// Record slot t from row t of the streams and leaf nodes in the shading's visiting order (spatially coherent), so a
// map tile's records, listed in slot order, lie close together in memory; the sums do not depend on slot order.
extern "C" __global__ void __launch_bounds__(64)
sl_lsm_prep(uint32_t n, LSParamsDev LS, const float4 *__restrict__ s0, const float4 *__restrict__ s1,
            const float4 *__restrict__ s2, uint32_t st, const float4 *__restrict__ nodes,
            const float4 *__restrict__ parts, MapRec *__restrict__ recs, uint2 *__restrict__ tbox,
            uint64_t *__restrict__ tmask) {
  const uint32_t t = blockIdx.x * 64u + threadIdx.x;
  if (t >= n) return;
  const MapRec r = lsm_rec(t, LS, s0, s1, s2, st, nodes, parts);
  const uint2 b = map_tbox(r, (LS.res + MT - 1) / MT);
  tbox[t] = b;
  if (b.x != 0xffffffffu) {   // an empty record is in no tile's list and never read
    recs[t] = r;
    tmask[t] = map_tile_mask(r);
  }
}

extern "C" __global__ void __launch_bounds__(64)
sl_lsm_lookup(uint32_t n, LSParamsDev LS, const float4 *__restrict__ g0, const float4 *__restrict__ g2,
              uint32_t st, const float4 *__restrict__ parts, const float4 *__restrict__ mip,
              float *__restrict__ T_out) {
  const uint32_t g = blockIdx.x * 64u + threadIdx.x;
  if (g >= n) return;
  const float4 A = g0[g], S4 = sl_scale(g2, st, g);
  const uint32_t own = __float_as_uint(S4.w) & 0xFFFFu;
  const v3 P = xf_point(load_xf(parts, own), V3(A.x, A.y, A.z));
  const v3 center = V3(LS.center[0], LS.center[1], LS.center[2]);
  const v3 dv = sub3(P, center);
  const float z = len3(dv);
  if (z < 1e-4f) { T_out[g] = 1.0f; return; }
  const v3 w = V3(dv.x / z, dv.y / z, dv.z / z);
  const v3 b0 = basis_k(w), b1 = cross3(w, b0);
  const m2 J = oct_jac(LS, w, b0, b1);
  const float dj = fabsf(m2det(J));
  const float tex_ang = (2.0f / (float)LS.res) / sqrtf(dj > 1e-12f ? dj : 1e-12f);
  const v2 uv = oct_uv(LS, w);
  const v2 tc = V2((uv.x * 0.5f + 0.5f) * (float)LS.res, (uv.y * 0.5f + 0.5f) * (float)LS.res);
  float T = 1.0f;
  for (uint32_t p = 0; p < LS.part_count; p++) {
    if (p == own || p == LS.skip_part) continue;
    const float4 bs = parts[p * 8 + 6];
    const v3 bc = sub3(xf_point(load_xf(parts, p), V3(bs.x, bs.y, bs.z)), center);
    const float dc = len3(bc);
    const float zn = dc - bs.w > 0.02f ? dc - bs.w : 0.02f;
    if (z <= zn) continue;
    const float search = 2.0f * LS.sigma_l * (z - zn) / (zn * z);
    if (dc > bs.w) {
      const float ang = acos_cr(clampf_(dot3(V3(bc.x / dc, bc.y / dc, bc.z / dc), w), -1.0f, 1.0f));
      if (ang > asin_cr(bs.w / dc) + search + 2.0f * tex_ang) continue;
    }
    float x = 2.0f * search / tex_ang;
    const float ls = log2_cr(x > 1.0f ? x : 1.0f);
    const float4 *mp = mip + (size_t)p * LS.layer_texels;
    const float4 a0 = mip_tri(mp, LS.level_off, LS.res, LS.levels, ceilf(ls), tc);
    if (a0.x < 1e-4f) continue;
    float zb = a0.y / a0.x;
    zb = zb > 0.02f ? zb : 0.02f;
    const float pen = LS.sigma_l * (z - zb > 0.0f ? z - zb : 0.0f) / (zb * z);
    x = 2.0f * pen / tex_ang;
    const float4 s = mip_tri(mp, LS.level_off, LS.res, LS.levels, log2_cr(x > 1.0f ? x : 1.0f), tc);
    if (s.x < 1e-5f) continue;
    const float mu = s.y / s.x;
    const float var = s.z / s.x - mu * mu;
    const float sd = sqrtf(var > 1e-8f ? var : 1e-8f);
    T *= 1.0f - s.x * phi_((z - mu) / sd);
  }
  T_out[g] = T;
}

// ---- sun maps (sun_map.wgsl:91-219, 263-324) ----------------------------------------------------------------------

extern "C" __global__ void __launch_bounds__(64)
sl_sun_frame(SunParamsDev SM, const SunLayerDesc *__restrict__ desc, const float4 *__restrict__ parts,
             SunLayer *__restrict__ layers) {
  const uint32_t l = blockIdx.x * 64u + threadIdx.x;
  if (l >= SM.n_layers) return;
  const SunLayerDesc d = desc[l];
  const v3 L = V3(SM.lights[d.light][0], SM.lights[d.light][1], SM.lights[d.light][2]);
  const v3 t1 = basis_k(L), t2 = cross3(L, t1);
  const float4 b = parts[d.part * 8 + 6];
  const v3 c = xf_point(load_xf(parts, d.part), V3(b.x, b.y, b.z));
  SunLayer o;
  o.c[0] = dot3(c, t1);
  o.c[1] = dot3(c, t2);
  o.half = b.w + SM.margin;
  if (d.kind == 1u) { o.c[0] = d.c[0]; o.c[1] = d.c[1]; o.half = d.half; }
  o.zref = dot3(c, L) + b.w;
  o.t1[0] = t1.x; o.t1[1] = t1.y; o.t1[2] = t1.z; o.part = d.part;
  o.t2[0] = t2.x; o.t2[1] = t2.y; o.t2[2] = t2.z; o.light = d.light;
  layers[l] = o;
}

// The records of row t (the streams and leaf nodes in the shading's visiting order) on the layers of its part: slot
// j n + t for the j-th layer, empty when the layer is clean, its light is off or nothing reaches half a fixed-point
// unit; z = z0 - gz . dt, stored as z0 + (-gz) . dt (the same float result: negation is exact).
extern "C" __global__ void __launch_bounds__(64)
sl_sun_prep(uint32_t n, SunParamsDev SM, const SunLayerDesc *__restrict__ desc, const SunLayer *__restrict__ layers,
            const float4 *__restrict__ s0, const float4 *__restrict__ s1, const float4 *__restrict__ s2,
            uint32_t st, const float4 *__restrict__ nodes, const float4 *__restrict__ parts, uint32_t max_layers,
            MapRec *__restrict__ recs, uint2 *__restrict__ tbox, uint64_t *__restrict__ tmask) {
  const uint32_t t = blockIdx.x * 64u + threadIdx.x;
  if (t >= n) return;
  // slot of the row's j-th layer: j n + t (layer-major, rows in visiting order: a map tile's records of one layer
  // lie close together in memory)
  MapRec *out = recs + t;
  uint2 *obox = tbox + t;
  uint64_t *omask = tmask + t;
  const float4 A = s0[t], Q = sl_quat(s1, st, t), S4 = sl_scale(s2, st, t);
  const uint32_t p = __float_as_uint(S4.w) & 0xFFFFu;
  const uint32_t first = p < 32u ? SM.part_layers[2 * p] : 0u, cnt = p < 32u ? min(SM.part_layers[2 * p + 1], max_layers) : 0u;
  for (uint32_t j = cnt; j < max_layers; j++) obox[(size_t)j * n] = make_uint2(0xffffffffu, 0u);   // in no list
  if (cnt == 0u) return;
  const Xf xf = load_xf(parts, p);
  const m3 R = part_rot(xf);
  const NodeL nd = load_node(nodes, t);
  const v3 x = xf_point(xf, V3(A.x, A.y, A.z));
  const m3 C = leaf_cov(Q, S4, R);
  const int res = (int)SM.res;
  const uint32_t tiles_y = (SM.res + MT - 1) / MT;
  for (uint32_t j = 0; j < cnt; j++) {
    const uint32_t l = first + j;
    const SunLayerDesc d = desc[l];
    if (d.dirty == 0u || SM.lights[d.light][3] < 0.0f) {
      obox[(size_t)j * n] = make_uint2(0xffffffffu, 0u);
      continue;
    }
    const SunLayer ly = layers[l];
    const v3 w = V3(SM.lights[d.light][0], SM.lights[d.light][1], SM.lights[d.light][2]);
    const v3 b0 = V3(ly.t1[0], ly.t1[1], ly.t1[2]), b1 = V3(ly.t2[0], ly.t2[1], ly.t2[2]);
    const v3 c0 = m3v(C, b0), c1 = m3v(C, b1), cw = m3v(C, w);
    const m2 Cbb = {V2(dot3(b0, c0), dot3(b1, c0)), V2(dot3(b0, c1), dot3(b1, c1))};
    const v2 cbz = V2(dot3(b0, cw), dot3(b1, cw));
    const m2 Cbb_i = inv2(Cbb);
    float sz2 = dot3(w, cw) - dot2(cbz, m2v(Cbb_i, cbz));
    sz2 = sz2 > 1e-12f ? sz2 : 1e-12f;
    const float tex = (float)SM.res / (2.0f * ly.half);
    m2 S = m2s(Cbb, tex * tex);
    const float det0 = m2det(S);
    S.c0.x += 0.3f;
    S.c1.y += 0.3f;
    const float det1 = m2det(S);
    const float tau0 = nd.kappa * dirm(nd, xf_dir_inv(xf, w)) * sqrtf(2.0f * PI_F * sz2) *
                       sqrtf((det0 > 0.0f ? det0 : 0.0f) / det1);
    const m2 Si = inv2(S);
    const float z0 = ly.zref - dot3(x, w);
    const v2 cz = m2v(Cbb_i, cbz);
    const v2 gz = V2(cz.x / tex, cz.y / tex);
    const v2 ct = V2(((dot3(x, b0) - ly.c[0]) + ly.half) * tex, ((dot3(x, b1) - ly.c[1]) + ly.half) * tex);
    const int hx = min((int)ceilf(3.0f * sqrtf(S.c0.x)), 32), hy = min((int)ceilf(3.0f * sqrtf(S.c1.y)), 32);
    const int cx = (int)floorf(ct.x), cy = (int)floorf(ct.y);
    if (tau0 * 4096.0f < 0.5f) {   // no texel reaches half a fixed-point unit
      obox[(size_t)j * n] = make_uint2(0xffffffffu, 0u);
      continue;
    }
    MapRec o;
    o.ctx = ct.x; o.cty = ct.y;
    o.s00 = Si.c0.x; o.s01 = Si.c0.y; o.s10 = Si.c1.x; o.s11 = Si.c1.y;
    o.tau0 = tau0; o.zc = z0; o.gzx = -gz.x; o.gzy = -gz.y; o.sz2 = sz2;
    o.layer = l;
    o.x0 = max(cx - hx, 0); o.x1 = min(cx + hx, res - 1);
    o.y0 = max(cy - hy, 0); o.y1 = min(cy + hy, res - 1);
    const uint2 b = map_tbox(o, tiles_y);
    obox[(size_t)j * n] = b;
    if (b.x != 0xffffffffu) {   // an empty record is in no tile's list and never read
      out[(size_t)j * n] = o;
      omask[(size_t)j * n] = map_tile_mask(o);
    }
  }
}

extern "C" __global__ void __launch_bounds__(64)
sl_sun_lookup(uint32_t n, SunParamsDev SM, const SunLayerDesc *__restrict__ desc, const SunLayer *__restrict__ layers,
              const float4 *__restrict__ g0, const float4 *__restrict__ g2, uint32_t st,
              const float4 *__restrict__ parts, const float4 *__restrict__ mip, float4 *__restrict__ env_T) {
  const uint32_t g = blockIdx.x * 64u + threadIdx.x;
  if (g >= n) return;
  const float4 A = g0[g], S4 = sl_scale(g2, st, g);
  const uint32_t own = __float_as_uint(S4.w) & 0xFFFFu;
  const v3 P = xf_point(load_xf(parts, own), V3(A.x, A.y, A.z));
  float T[4] = {1.0f, 1.0f, 1.0f, 1.0f};
  uint32_t used = 0xFFFFFFFFu;
  for (uint32_t l = 0; l < SM.n_layers; l++) {
    const SunLayer ly = layers[l];
    const uint32_t key = ly.part * 8u + ly.light;
    if (ly.part == own || key == used || SM.lights[ly.light][3] < 0.0f) continue;
    const v3 w = V3(SM.lights[ly.light][0], SM.lights[ly.light][1], SM.lights[ly.light][2]);
    const float ux = dot3(P, V3(ly.t1[0], ly.t1[1], ly.t1[2])) - ly.c[0];
    const float uy = dot3(P, V3(ly.t2[0], ly.t2[1], ly.t2[2])) - ly.c[1];
    const float inner = (desc[l].kind == 1u ? 0.9f : 1.0f) * ly.half;
    if (fabsf(ux) >= inner || fabsf(uy) >= inner) continue;
    used = key;
    const float z = ly.zref - dot3(P, w);
    if (z <= 0.0f) continue;
    const float texel = 2.0f * ly.half / (float)SM.res;
    const v2 tc = V2((ux + ly.half) / texel, (uy + ly.half) / texel);
    const float sig = SM.lights[ly.light][3];
    const float search = 2.0f * sig * z;
    float xv = 2.0f * search / texel;
    const float4 *mp = mip + (size_t)l * SM.layer_texels;
    const float4 a0 = mip_tri(mp, SM.level_off, SM.res, SM.levels, log2_cr(xv > 1.0f ? xv : 1.0f), tc);
    if (a0.x < 1e-4f) continue;
    const float zb = a0.y / a0.x;
    const float pen = sig * (z - zb > 0.0f ? z - zb : 0.0f);
    xv = 2.0f * pen / texel;
    const float4 s = mip_tri(mp, SM.level_off, SM.res, SM.levels, log2_cr(xv > 1.0f ? xv : 1.0f), tc);
    if (s.x < 1e-5f) continue;
    const float mu = s.y / s.x;
    const float var = s.z / s.x - mu * mu;
    const float sd = sqrtf(var > 1e-8f ? var : 1e-8f);
    T[ly.light] *= 1.0f - s.x * phi_((z - mu) / sd);
  }
  env_T[g] = make_float4(T[0], T[1], T[2], T[3]);
}

// ---- env_direct (environment.wgsl:207-229, 266-277), matte -------------------------------------------------------

template <bool SHINY>
__global__ void __launch_bounds__(64)
sl_env_direct(FrameUniformsDev U, EnvDev ENV, const float4 *__restrict__ g0, const float4 *__restrict__ g1,
              const float4 *__restrict__ g2,
              const float4 *__restrict__ g3, uint32_t st, const float4 *__restrict__ nodes, const uint8_t *__restrict__ vgrid, const float4 *__restrict__ parts,
              const uint32_t *__restrict__ materials, const float *__restrict__ iso_norm,
              const float4 *__restrict__ env_T, float4 *__restrict__ colour) {
  const uint32_t g = blockIdx.x * 64u + threadIdx.x;
  if (g >= U.gaussian_count || ENV.info[0] == 0u || (U.iso & 1u) == 0u) return;
  const float4 S4 = sl_scale(g2, st, g), C = sl_albedo(g1, g2, g3, st, g);
  const uint32_t mi = __float_as_uint(C.w) & 0xFFFFu;
  const uint32_t mmax = max(U.material_count, 1u) - 1u;
  const uint32_t *mat = materials + 8 * (size_t)min(mi, mmax);
  if (U.material_count && (mat[5] & 1u)) return;
  const NodeL nd = load_node(nodes, g);
  const Xf xf = load_xf(parts, __float_as_uint(S4.w) & 0xFFFFu);
  const float s3 = S4.x * S4.y * S4.z;
  const float4 Tg = env_T ? env_T[g] : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
  const float Tk[4] = {Tg.x, Tg.y, Tg.z, Tg.w};
  if constexpr (SHINY) {
    const float ks = U.material_count ? spec_weight(mat, mat[5]) : 0.0f;
    if (ks > 0.0f) {
      // shiny (environment.wgsl:278-307): (1 - k) of the matte light plus each light's highlight, its shape the
      // light's ASG of angular std dir_std.w
      const float4 A = g0[g], Q = sl_quat(g1, st, g);
      const v3 P = xf_point(xf, V3(A.x, A.y, A.z));
      const v3 wo = xf_dir_inv(xf, norm3(sub3(V3(U.eye[0], U.eye[1], U.eye[2]), P)));
      const m3 Rq = quat_matrix(Q);
      const v3 flake = V3(__uint_as_float(mat[0]), __uint_as_float(mat[1]), __uint_as_float(mat[2]));
      v3 Sd, So;
      flake_matrix(Rq, flake, Sd, So);
      const float inorm = iso_norm[g];
      const uint8_t *vg = vgrid + 128 * (size_t)g;
      v3 diff = V3(0.0f, 0.0f, 0.0f), spec = V3(0.0f, 0.0f, 0.0f);
      const uint32_t nl = min(ENV.info[0], 4u);
      for (uint32_t k = 0; k < nl; k++) {
        const v3 wi = xf_dir_inv(xf, V3(ENV.lights[k][0], ENV.lights[k][1], ENV.lights[k][2]));
        const float v = vgrid_lookup(vg, wi);
        if (v <= 1e-7f) continue;
        const v3 light = mul3s(V3(ENV.lights[k][4], ENV.lights[k][5], ENV.lights[k][6]),
                               Tk[k] * v * footprint(nd, s3 * s3, wi));
        const v3 wr = mirror(wi, Rq.c0);
        const float damp = iso_damp(inorm, footprint(nd, s3 * s3, wr) * vgrid_lookup(vg, wr));
        const float sd = ENV.lights[k][3];
        const float s2 = sd * sd;
        v3 lb0, lb1;
        onb(wi, lb0, lb1);
        const float p = asg_phase(wi, lb0, lb1, V3(s2, 0.0f, s2), Rq.c0, Rq.c1, Rq.c2, flake, Sd, So, wo);
        diff = add3(diff, mul3s(light, 1.0f - ks));
        spec = add3(spec, mul3s(light, iso_lobe_from(ks, p, damp) - (1.0f - ks)));
      }
      const v3 a = V3(C.x * inorm, C.y * inorm, C.z * inorm);
      const float4 c = colour[g];
      colour[g] = make_float4(c.x + a.x * (diff.x + spec.x), c.y + a.y * (diff.y + spec.y),
                              c.z + a.z * (diff.z + spec.z), c.w);
      return;
    }
  }
  v3 acc = V3(0.0f, 0.0f, 0.0f);
  const uint32_t nl = min(ENV.info[0], 4u);
  for (uint32_t k = 0; k < nl; k++) {
    const v3 wi = xf_dir_inv(xf, V3(ENV.lights[k][0], ENV.lights[k][1], ENV.lights[k][2]));
    const float v = vgrid_lookup(vgrid + 128 * (size_t)g, wi);
    if (v <= 1e-7f) continue;
    const float f = Tk[k] * v * footprint(nd, s3 * s3, wi);
    acc = add3(acc, mul3s(V3(ENV.lights[k][4], ENV.lights[k][5], ENV.lights[k][6]), f));
  }
  const float inorm = iso_norm[g];
  const float4 c = colour[g];
  colour[g] = make_float4(c.x + C.x * acc.x * inorm, c.y + C.y * acc.y * inorm, c.z + C.z * acc.z * inorm, c.w);
}
