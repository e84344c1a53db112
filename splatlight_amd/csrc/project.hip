// project: pose, cull and project every splat slot (shaders/project.wgsl:84-370) on gfx950.
//
// One lane per slot. Gaussian attributes are SoA float4 streams (mu|kappa, quat, scale|part_surface,
// albedo|flags), so each of the four loads of a wavefront reads one contiguous 1 KiB run. Outputs per slot: the
// 48 B Splat in the reference's layout, the 16-bit depth key (or INVALID), and the splat's inclusive tile (16 or 32 px)
// box packed as (tx0 | ty0 << 16, tx1 | ty1 << 16) for the pair emission.
//
// Contraction is off in this file: every float expression rounds as written, in the WGSL evaluation order
// (dot products left to right, m * v as m[0] v.x + m[1] v.y + m[2] v.z). Division and sqrt are IEEE correctly
// rounded and the two logs (depth key, cut radius) are the correctly rounded float log, so every splat word,
// the 16-bit key and the tile box are bit-identical to the CPU restatement's.
#pragma clang fp contract(off)
#include "sl_device.h"

namespace {

__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// quat_to_mat (shaders/common.wgsl:83-89): m[col][row]
__device__ __forceinline__ void quat_to_mat(float4 q, float m[3][3]) {
  const float x = q.x, y = q.y, z = q.z, w = q.w;
  m[0][0] = 1.0f - 2.0f * (y * y + z * z); m[0][1] = 2.0f * (x * y + w * z); m[0][2] = 2.0f * (x * z - w * y);
  m[1][0] = 2.0f * (x * y - w * z); m[1][1] = 1.0f - 2.0f * (x * x + z * z); m[1][2] = 2.0f * (y * z + w * x);
  m[2][0] = 2.0f * (x * z + w * y); m[2][1] = 2.0f * (y * z - w * x); m[2][2] = 1.0f - 2.0f * (x * x + y * y);
}

// The 16-bit back-to-front depth key (shaders/project.wgsl:366-367) with log taken as the correctly rounded
// float log (double log, rounded once).
__device__ __forceinline__ uint32_t depth_key(float depth, float near_, float log_far_near) {
  const float num = (float)log((double)(depth / near_));
  float t = num / log_far_near;
  t = fminf(fmaxf(t, 0.0f), 1.0f);
  const float s = t * 65535.0f + 0.5f;
  const uint32_t k = 65535u - (uint32_t)s;
  return k < 0xfffeu ? k : 0xfffeu;
}

}  // namespace

// The splat colour of a visible slot (project.wgsl:317-350): the lit colour plus gain x bounce in the default view,
// else the albedo or the part palette, scaled.
// node: an LOD node's lit colour (its bounce already included, project.wgsl:325), else null.
__device__ __forceinline__ void splat_colour(const FrameDev &f, uint32_t material, float4 C4,
                                             const float4 *__restrict__ colour, const float4 *__restrict__ lo_b,
                                             uint32_t i, float col[3], const float4 *__restrict__ node) {
  col[0] = C4.x; col[1] = C4.y; col[2] = C4.z;
  if (material != SL_EMITTER_MATERIAL) {
    if (f.view_mode == 3u) {
      if (node) {
        const float4 nc = *node;
        col[0] = nc.x;
        col[1] = nc.y;
        col[2] = nc.z;
      } else {
        const float4 cv = colour ? colour[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 lb = lo_b ? lo_b[3 * (size_t)i] : make_float4(0.f, 0.f, 0.f, 0.f);
        col[0] = cv.x + f.gain * lb.x;
        col[1] = cv.y + f.gain * lb.y;
        col[2] = cv.z + f.gain * lb.z;
      }
      if (material < 8u) { col[0] *= f.palette[material][0]; col[1] *= f.palette[material][1]; col[2] *= f.palette[material][2]; }
    } else if (material < 8u && f.palette[material][3] > 0.5f) {
      col[0] = f.palette[material][0]; col[1] = f.palette[material][1]; col[2] = f.palette[material][2];
    }
    col[0] *= f.albedo_scale; col[1] *= f.albedo_scale; col[2] *= f.albedo_scale;
  }
}

// One lane per global slot of a batch; the 256-slot block belongs to one view. A valid slot's key is
// (view << 16) | the 16-bit depth key, so one stable sort orders views, then depth, then slot; slots past a view's
// count (padding to 256) get INVALID. COLOUR: the splats' colour words and the per-block colour max are written
// here; without it they are left to sl_splat_colour (so the geometry need not wait for the lighting), the colour
// words are 0 and the partials' colour max is 0.
template <bool COLOUR, bool BLUR>
__global__ void __launch_bounds__(256)
sl_project_kernel(const ViewDev *__restrict__ views, const uint32_t *__restrict__ slot_base, uint32_t n_views,
                  Splat *__restrict__ splats, uint32_t *__restrict__ keys, uint2 *__restrict__ tbox,
                  uint2 *__restrict__ partials) {
  __shared__ uint2 s_part[256 / SL_WAVE];
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t v = n_views > 1u ? view_of(slot_base, n_views, blockIdx.x * blockDim.x) : 0u;
  const ViewDev &vd = views[v];
  const FrameDev &f = vd.f;
  const float4 *__restrict__ g0 = vd.g0, *__restrict__ g1 = vd.g1, *__restrict__ g2 = vd.g2, *__restrict__ g3 = vd.g3;
  const float4 *__restrict__ e0 = vd.e0, *__restrict__ e1 = vd.e1, *__restrict__ e2 = vd.e2, *__restrict__ e3 = vd.e3;
  const float4 *__restrict__ parts = vd.parts, *__restrict__ colour = vd.colour, *__restrict__ lo_b = vd.lo_b;
  const uint32_t i = gi - vd.slot_base;
  bool ok = i < f.nslots && i != f.count;
  // the LOD's cut (project.wgsl:90-109): node slots read the node streams, only flagged slots are drawn
  bool is_node = false;
  uint32_t src = i;
  if (ok && vd.lod_flag) {
    if (i >= vd.node_slot0) {
      src = i - vd.node_slot0;
      is_node = true;
      ok = vd.lod_flag[f.count + src] == SL_LOD_DRAW;
    } else if (i < f.count) {
      ok = vd.lod_flag[i] == SL_LOD_DRAW;
    }
  }
  float4 A, Q4, S4, C4;
  float mu[3], rot[3][3];
  uint32_t material = SL_EMITTER_MATERIAL;
  float keep = 1.0f;
  float vel[3] = {0.0f, 0.0f, 0.0f};   // BLUR: world displacement over the exposure (project.wgsl:138-149)
  if (ok) {
    if (i > f.count && !is_node) {                                    // emitters (project.wgsl:123-127)
      const uint32_t k = i - f.count - 1u;
      A = e0[k]; Q4 = e1[k]; S4 = e2[k]; C4 = e3[k];
      mu[0] = A.x; mu[1] = A.y; mu[2] = A.z;
      quat_to_mat(Q4, rot);
    } else {                                                          // pose (project.wgsl:128-148)
      const float4 *__restrict__ s0 = is_node ? vd.n0 : g0, *__restrict__ s1 = is_node ? vd.n1 : g1;
      const float4 *__restrict__ s2 = is_node ? vd.n2 : g2, *__restrict__ s3 = is_node ? vd.n3 : g3;
      A = s0[src]; Q4 = sl_quat(s1, vd.storage, src); S4 = sl_scale(s2, vd.storage, src);
      C4 = sl_albedo(s1, s2, s3, vd.storage, src);
      uint32_t part = __float_as_uint(S4.w) & 0xffffu;
      part = part < SL_MAX_PARTS - 1u ? part : SL_MAX_PARTS - 1u;
      const float4 x0 = parts[part * 8 + 0], x1 = parts[part * 8 + 1], x2 = parts[part * 8 + 2];
      const uint32_t pflags = __float_as_uint(parts[part * 8 + 7].w);
      keep = f.part_keep[part];
      ok = (pflags & SL_PART_HIDDEN) == 0u && keep > 0.0f;
      mu[0] = ((x0.x * A.x + x0.y * A.y) + x0.z * A.z) + x0.w * 1.0f;
      mu[1] = ((x1.x * A.x + x1.y * A.y) + x1.z * A.z) + x1.w * 1.0f;
      mu[2] = ((x2.x * A.x + x2.y * A.y) + x2.z * A.z) + x2.w * 1.0f;
      float q[3][3];
      quat_to_mat(Q4, q);
      const float xr[3][3] = {{x0.x, x0.y, x0.z}, {x1.x, x1.y, x1.z}, {x2.x, x2.y, x2.z}};   // xr[row][col]
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int r = 0; r < 3; r++) rot[j][r] = (xr[r][0] * q[j][0] + xr[r][1] * q[j][1]) + xr[r][2] * q[j][2];
      if (BLUR && f.shutter > 0.0f && (pflags & 2u) != 0u) {
        // the chord from the previous pose (xf_prev, words 12-23) over the shutter, centred mid-exposure
        const float4 p0 = parts[part * 8 + 3], p1 = parts[part * 8 + 4], p2 = parts[part * 8 + 5];
        const float prev[3] = {((p0.x * A.x + p0.y * A.y) + p0.z * A.z) + p0.w * 1.0f,
                               ((p1.x * A.x + p1.y * A.y) + p1.z * A.z) + p1.w * 1.0f,
                               ((p2.x * A.x + p2.y * A.y) + p2.z * A.z) + p2.w * 1.0f};
#pragma unroll
        for (int r = 0; r < 3; r++) vel[r] = f.shutter * (mu[r] - prev[r]);
#pragma unroll
        for (int r = 0; r < 3; r++) mu[r] -= 0.5f * vel[r];
      }
      material = f.part_material[part];
    }
  }
  float tv[3], depth = 0.0f;
  if (ok) {
#pragma unroll
    for (int r = 0; r < 3; r++)
      tv[r] = ((f.view[r] * mu[0] + f.view[4 + r] * mu[1]) + f.view[8 + r] * mu[2]) + f.view[12 + r] * 1.0f;
    depth = -tv[2];
    ok = !(depth < f.near_ || depth > f.far_);                       // project.wgsl:199-200
  }
  Splat out;
  uint32_t key = SL_INVALID_KEY;
  uint2 box = make_uint2(0xffffffffu, 0u);
  float cmax = 0.0f;
  if (ok) {
    // 3D covariance in view space W R S^2 R^T W^T (project.wgsl:202-206)
    float wr[3][3], m[3][3], cov[3][3];
    const float sc[3] = {S4.x, S4.y, S4.z};
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int r = 0; r < 3; r++)
        wr[j][r] = (f.view[r] * rot[j][0] + f.view[4 + r] * rot[j][1]) + f.view[8 + r] * rot[j][2];
    const float s2[3] = {sc[0] * sc[0], sc[1] * sc[1], sc[2] * sc[2]};
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int r = 0; r < 3; r++) m[j][r] = wr[j][r] * sc[j];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int r = 0; r < 3; r++) cov[j][r] = (m[0][r] * m[0][j] + m[1][r] * m[1][j]) + m[2][r] * m[2][j];
    const bool blurred = BLUR && (vel[0] != 0.0f || vel[1] != 0.0f || vel[2] != 0.0f);
    float cov_s[3][3];                                               // the sharp footprint (project.wgsl:208-210)
    if (BLUR) {
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int r = 0; r < 3; r++) cov_s[j][r] = cov[j][r];
    }
    if (blurred) {
      float wv[3];
#pragma unroll
      for (int r = 0; r < 3; r++) wv[r] = (f.view[r] * vel[0] + f.view[4 + r] * vel[1]) + f.view[8 + r] * vel[2];
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int r = 0; r < 3; r++) cov[j][r] = cov[j][r] + (wv[r] * wv[j]) * (1.0f / 12.0f);
    }

    // EWA Jacobian with the 1.3x tangent clamp (project.wgsl:212-231)
    const float iz = 1.0f / depth;
    const float px = f.fx * tv[0] * iz;
    const float py = f.fy * tv[1] * iz;
    const float limx = 1.3f * f.vw * 0.5f, limy = 1.3f * f.vh * 0.5f;
    const float cx = clampf(tv[0] * iz, -limx / f.fx, limx / f.fx) * depth;
    const float cy = clampf(tv[1] * iz, -limy / f.fy, limy / f.fy) * depth;
    const float j0[3] = {f.fx * iz, 0.0f, f.fx * cx * iz * iz};
    const float j1[3] = {0.0f, f.fy * iz, f.fy * cy * iz * iz};
    float cj0[3], cj1[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
      cj0[r] = (cov[0][r] * j0[0] + cov[1][r] * j0[1]) + cov[2][r] * j0[2];
      cj1[r] = (cov[0][r] * j1[0] + cov[1][r] * j1[1]) + cov[2][r] * j1[2];
    }
    float a = (j0[0] * cj0[0] + j0[1] * cj0[1]) + j0[2] * cj0[2];
    const float b = (j0[0] * cj1[0] + j0[1] * cj1[1]) + j0[2] * cj1[2];
    float c = (j1[0] * cj1[0] + j1[1] * cj1[1]) + j1[2] * cj1[2];
    const float det0 = a * c - b * b;
    a += f.lowpass;
    c += f.lowpass;
    const float det = a * c - b * b;
    ok = !(det <= 0.0f);                                              // project.wgsl:232
    // peak optical depth tau0 = kappa sqrt(2 pi / q) sqrt(det0 / det) (project.wgsl:237-261)
    const float len = sqrtf((tv[0] * tv[0] + tv[1] * tv[1]) + tv[2] * tv[2]);
    const float d[3] = {tv[0] / len, tv[1] / len, tv[2] / len};
    float dl[3];
#pragma unroll
    for (int k = 0; k < 3; k++) dl[k] = (wr[k][0] * d[0] + wr[k][1] * d[1]) + wr[k][2] * d[2];
    float q = ((dl[0] * dl[0]) * (1.0f / s2[0]) + (dl[1] * dl[1]) * (1.0f / s2[1])) +
              (dl[2] * dl[2]) * (1.0f / s2[2]);
    float kappa = A.w;
    float tau0;
    if (blurred) {
      // Sherman-Morrison for d^T (S + v v^T / 12)^-1 d and the mass-keeping density (project.wgsl:242-251)
      const float q_s = q;
      float vl[3];
#pragma unroll
      for (int k = 0; k < 3; k++) vl[k] = (rot[k][0] * vel[0] + rot[k][1] * vel[1]) + rot[k][2] * vel[2];
      const float mv2 = ((vl[0] * vl[0]) * (1.0f / s2[0]) + (vl[1] * vl[1]) * (1.0f / s2[1])) +
                        (vl[2] * vl[2]) * (1.0f / s2[2]);
      const float dv = ((dl[0] * vl[0]) * (1.0f / s2[0]) + (dl[1] * vl[1]) * (1.0f / s2[1])) +
                       (dl[2] * vl[2]) * (1.0f / s2[2]);
      q -= dv * dv / (12.0f + mv2);
      kappa /= sqrtf(1.0f + mv2 / 12.0f);
      tau0 = kappa * sqrtf(6.283185307f / q) * sqrtf(fmaxf(det0, 0.0f) / det);
      if (f.blur_cover) {
        // coverage cap: the peak alpha at most the sharp splat's x sqrt(det_s / det) (project.wgsl:253-268); exp
        // and log correctly rounded, as the host restatement defines them
        float sj0[3], sj1[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
          sj0[r] = (cov_s[0][r] * j0[0] + cov_s[1][r] * j0[1]) + cov_s[2][r] * j0[2];
          sj1[r] = (cov_s[0][r] * j1[0] + cov_s[1][r] * j1[1]) + cov_s[2][r] * j1[2];
        }
        const float a_s = ((j0[0] * sj0[0] + j0[1] * sj0[1]) + j0[2] * sj0[2]) + f.lowpass;
        const float b_s = (j0[0] * sj1[0] + j0[1] * sj1[1]) + j0[2] * sj1[2];
        const float c_s = ((j1[0] * sj1[0] + j1[1] * sj1[1]) + j1[2] * sj1[2]) + f.lowpass;
        const float det0_s = (a_s - f.lowpass) * (c_s - f.lowpass) - b_s * b_s;
        const float det_s = a_s * c_s - b_s * b_s;
        const float tau_s = A.w * sqrtf(6.283185307f / q_s) * sqrtf(fmaxf(det0_s, 0.0f) / det_s);
        const float ap = (1.0f - (float)exp((double)-tau_s)) * sqrtf(det_s / det);
        const float cap = -(float)log((double)fmaxf(1.0f - ap, 1e-6f));
        tau0 = fminf(tau0, cap);
      }
    } else {
      tau0 = kappa * sqrtf(6.283185307f / q) * sqrtf(fmaxf(det0, 0.0f) / det);
    }
    tau0 *= keep;
    // cut radius r: alpha >= 1/255 (project.wgsl:290-295); tau_min = -log(1 - 1/255) correctly rounded
    const float tau_min = 0x1.018258p-8f;
    ok = ok && !(tau0 <= tau_min);
    const float r = sqrtf(2.0f * (float)log((double)(tau0 / tau_min)));   // correctly rounded float log
    // stable eigen axes of the 2x2 covariance (project.wgsl:297-311)
    const float mid = 0.5f * (a + c);
    const float disc = sqrtf(0.25f * (a - c) * (a - c) + b * b);
    const float l0 = mid + disc;
    const float l1 = fmaxf(det / l0, 1e-6f);
    float e0x = 1.0f, e0y = 0.0f;
    if (a >= c) {
      if (fabsf(b) > 0.0f || l0 > a) {
        const float vx = l0 - c, vy = b, n = sqrtf(vx * vx + vy * vy);
        e0x = vx / n; e0y = vy / n;
      }
    } else {
      const float vx = b, vy = l0 - a, n = sqrtf(vx * vx + vy * vy);
      e0x = vx / n; e0y = vy / n;
    }
    const float e1x = -e0y, e1y = e0x;
    const float tnx = 2.0f / f.vw, tny = 2.0f / f.vh;
    const float sl0 = sqrtf(l0), sl1 = sqrtf(l1);
    const float a0x = e0x * sl0 * r * tnx, a0y = e0y * sl0 * r * tny;
    const float a1x = e1x * sl1 * r * tnx, a1y = e1y * sl1 * r * tny;
    ok = ok && !(sl0 * r < 0.1f);
    const float cenx = px * tnx, ceny = py * tny;                     // project.wgsl:313-315
    const float ex = fabsf(a0x) + fabsf(a1x), ey = fabsf(a0y) + fabsf(a1y);
    ok = ok && !(fabsf(cenx) - ex > 1.0f || fabsf(ceny) - ey > 1.0f);
    if (ok) {
      float col[3] = {0.0f, 0.0f, 0.0f};
      if (COLOUR) splat_colour(f, material, C4, colour, lo_b, i, col, is_node ? vd.node_colour + src : nullptr);
      out.center_axis0 = make_float4(cenx, ceny, a0x, a0y);
      out.axis1_r_tau = make_float4(a1x, a1y, r, tau0);
      // SPLAT_GLASS (shaders/lens.wgsl:21) in the depth word's low bits (project.wgsl:167-168)
      const uint32_t mat_idx = min(__float_as_uint(C4.w) & 0xffffu, 31u);
      const uint32_t glass = material != SL_EMITTER_MATERIAL && !is_node ? (f.glass_mask >> mat_idx) & 1u : 0u;
      out.color_depth = make_uint4(__float_as_uint(col[0]), __float_as_uint(col[1]), __float_as_uint(col[2]),
                                   (__float_as_uint(depth) & ~3u) | glass);
      key = (v << 16) | depth_key(depth, f.near_, f.log_far_near);
      if (COLOUR) cmax = fmaxf(fmaxf(fabsf(col[0]), fabsf(col[1])), fabsf(col[2]));
      // tile box of the extent box in pixel centres (the pair definition)
      const float hw = 0.5f * (float)f.width, hh = 0.5f * (float)f.height;
      float fx0 = ceilf((cenx - ex + 1.0f) * hw - 0.5f), fx1 = floorf((cenx + ex + 1.0f) * hw - 0.5f);
      float fy0 = ceilf((1.0f - (ceny + ey)) * hh - 0.5f), fy1 = floorf((1.0f - (ceny - ey)) * hh - 0.5f);
      fx0 = fx0 < 0.0f ? 0.0f : fx0;
      fy0 = fy0 < 0.0f ? 0.0f : fy0;
      fx1 = fx1 > (float)(f.width - 1) ? (float)(f.width - 1) : fx1;
      fy1 = fy1 > (float)(f.height - 1) ? (float)(f.height - 1) : fy1;
      if (fx0 <= fx1 && fy0 <= fy1) {
        const uint32_t sh = f.tile_shift;
        const uint32_t tx0 = (uint32_t)fx0 >> sh, ty0 = (uint32_t)fy0 >> sh;
        const uint32_t tx1 = (uint32_t)fx1 >> sh, ty1 = (uint32_t)fy1 >> sh;
        box = make_uint2(tx0 | (ty0 << 16), tx1 | (ty1 << 16));
      }
    }
  }
  if (ok) splats[gi] = out;
  keys[gi] = key;
  tbox[gi] = box;
  // per-block visible count and colour max into partials[block] (no atomics: sl_frame_totals reduces them)
  const uint64_t vis = __ballot(ok);
  float wmax = cmax;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, o));
  if (lane_id() == 0) s_part[threadIdx.x >> 6] = make_uint2((uint32_t)__popcll(vis), __float_as_uint(wmax));
  __syncthreads();
  if (threadIdx.x == 0) {
    uint2 t = s_part[0];
#pragma unroll
    for (int k = 1; k < 256 / SL_WAVE; k++) {
      t.x += s_part[k].x;
      t.y = max(t.y, s_part[k].y);   // non-negative floats order as their bits
    }
    partials[blockIdx.x] = t;
  }
}

// The colour words of the visible splats and the per-block colour max, for a batch projected without COLOUR:
// the same slots, colours and block partials (their .y) the full projection writes. Visible: a valid key.
extern "C" __global__ void __launch_bounds__(256)
sl_splat_colour(const ViewDev *__restrict__ views, const uint32_t *__restrict__ slot_base, uint32_t n_views,
                const uint32_t *__restrict__ keys, Splat *__restrict__ splats, uint2 *__restrict__ partials) {
  __shared__ uint32_t s_max[256 / SL_WAVE];
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t v = n_views > 1u ? view_of(slot_base, n_views, blockIdx.x * blockDim.x) : 0u;
  const ViewDev &vd = views[v];
  const FrameDev &f = vd.f;
  const uint32_t i = gi - vd.slot_base;
  float cmax = 0.0f;
  if (keys[gi] != SL_INVALID_KEY) {
    float4 C4;
    uint32_t material = SL_EMITTER_MATERIAL;
    const bool is_node = vd.lod_flag && i >= vd.node_slot0;
    const uint32_t src = is_node ? i - vd.node_slot0 : i;
    if (i > f.count && !is_node) {
      C4 = vd.e3[i - f.count - 1u];
    } else {
      const float4 *__restrict__ s1 = is_node ? vd.n1 : vd.g1, *__restrict__ s2 = is_node ? vd.n2 : vd.g2;
      const float4 *__restrict__ s3 = is_node ? vd.n3 : vd.g3;
      C4 = sl_albedo(s1, s2, s3, vd.storage, src);
      uint32_t part = __float_as_uint(sl_scale(s2, vd.storage, src).w) & 0xffffu;
      part = part < SL_MAX_PARTS - 1u ? part : SL_MAX_PARTS - 1u;
      material = f.part_material[part];
    }
    float col[3];
    splat_colour(f, material, C4, vd.colour, vd.lo_b, i, col, is_node ? vd.node_colour + src : nullptr);
    uint32_t *cw = reinterpret_cast<uint32_t *>(&splats[gi].color_depth);
    cw[0] = __float_as_uint(col[0]);
    cw[1] = __float_as_uint(col[1]);
    cw[2] = __float_as_uint(col[2]);
    cmax = fmaxf(fmaxf(fabsf(col[0]), fabsf(col[1])), fabsf(col[2]));
  }
  float wmax = cmax;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, o));
  if (lane_id() == 0) s_max[threadIdx.x >> 6] = __float_as_uint(wmax);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = s_max[0];
#pragma unroll
    for (int k = 1; k < 256 / SL_WAVE; k++) t = max(t, s_max[k]);   // non-negative floats order as their bits
    partials[blockIdx.x].y = t;
  }
}

// Per-view totals from project's per-block partials (one 256-lane workgroup per view: it starts on a CU with four free wave slots, which the map sums
// running beside it leave sooner than sixteen): visible count and colour max into
// view_counters[v]; the batch's visible total into counters. what: bit 0 the visible counts, bit 1 the colour max.
extern "C" __global__ void __launch_bounds__(256)
sl_frame_totals(const uint2 *__restrict__ partials, const uint32_t *__restrict__ slot_base,
                ViewCounters *__restrict__ view_counters, FrameCounters *__restrict__ counters, uint32_t what) {
  __shared__ uint2 s_w[256 / SL_WAVE];
  const uint32_t v = blockIdx.x;
  const uint32_t b0 = slot_base[v] / 256u, b1 = slot_base[v + 1] / 256u;
  uint2 t = make_uint2(0u, 0u);
  for (uint32_t i = b0 + threadIdx.x; i < b1; i += 256u) {
    const uint2 p = partials[i];
    t.x += p.x;
    t.y = max(t.y, p.y);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    t.x += __shfl_xor(t.x, o);
    t.y = max(t.y, (uint32_t)__shfl_xor(t.y, o));
  }
  if (lane_id() == 0) s_w[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint2 a = s_w[0];
    for (int k = 1; k < 256 / SL_WAVE; k++) {
      a.x += s_w[k].x;
      a.y = max(a.y, s_w[k].y);
    }
    if (what & 1u) {
      view_counters[v].visible = a.x;
      if (a.x) atomicAdd(&counters->visible, a.x);
    }
    if (what & 2u) {
      view_counters[v].color_max_bits = a.y;
      atomicMax(&counters->color_max_bits, a.y);
    }
  }
}
