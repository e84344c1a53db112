// Front-to-back compositing of 16x16 tiles (the blend of shaders/splat.wgsl:22-78 with the "over" state of
// src/renderer.js:140-142, evaluated per tile instead of by the rasteriser).
//
// One 256-lane workgroup per tile, one lane per pixel. The tile's pair range is walked from its end, which is
// the reference's back-to-front sequence for every pixel reversed, in batches of 256 splats: each lane stages
// one splat into LDS already reduced to what a pixel needs (centre, the inverse of the quad's axes, the Gaussian
// falloff and optical depth folded into exp2 arguments, colour), then every lane sweeps the batch. A pixel stops
// once its transmittance is below t_eps = 2^-12 / max(1, max|rgb|) (so the colour it no longer adds is below
// 2^-12 absolute), a wave skips batches when all its pixels stopped, and the workgroup leaves when all did.
//
// Pixel (px, py) has its centre at NDC ((px + 0.5) 2 / W - 1, 1 - (py + 0.5) 2 / H), rows top-down; a pixel is
// covered by a splat when its centre lies in the quad c + u a0 + v a1, |u|, |v| <= 1, and its alpha
// 1 - exp(-tau0 exp(-r^2 (u^2 + v^2) / 2)) reaches 1/255. The fast path evaluates u, v and alpha with folded
// constants and exp2; a fragment within a hair of the quad edge or of the 1/255 threshold (|u| or |v| within
// 1e-4 of 1, alpha within 0.1% of 1/255) is decided again with the exact expression: u, v by the 2x2 solve,
// exp as the correctly rounded float exp, no contraction. Coverage is then decided exactly as the CPU
// restatement decides it, so a fragment never lands on the other side of the discard.
#include "sl_device.h"
#include <hip/hip_fp16.h>

#define CP_THREADS 256
#define CP_BATCH 256
#define LOG2E 1.4426950408889634f

struct __align__(16) StagedSplat {
  float cx, cy, ux, uy;      // centre (NDC); u = dx ux + dy uy
  float vx, vy, kq, kt;      // v = dx vx + dy vy; e = exp2(kq (u^2 + v^2)); alpha = 1 - exp2(kt e)
  float r, g, b, pad;
};

struct __align__(16) RawSplat {  // the splat's own words, for the exact re-evaluation
  float a0x, a0y, a1x, a1y;
  float det, rr, tau0, pad;
};

// The exact coverage of one fragment (the CPU restatement's expression order, no contraction, correctly
// rounded exp): alpha, or -1 when the pixel centre is outside the quad.
__device__ __noinline__ float exact_alpha(float dx, float dy, const RawSplat rs) {
#pragma clang fp contract(off)
  const float u = (dx * rs.a1y - rs.a1x * dy) / rs.det;
  const float v = (rs.a0x * dy - dx * rs.a0y) / rs.det;
  if (fabsf(u) > 1.0f || fabsf(v) > 1.0f) return -1.0f;
  const float lx = u * rs.rr, ly = v * rs.rr;
  const float e = (float)exp((double)(-0.5f * (lx * lx + ly * ly)));
  return 1.0f - (float)exp((double)(-rs.tau0 * e));
}

extern "C" __global__ void __launch_bounds__(CP_THREADS)
sl_composite(FrameDev f, const Splat *__restrict__ splats, const uint32_t *__restrict__ vals,
             const uint2 *__restrict__ ranges, const FrameCounters *__restrict__ counters, void *__restrict__ out,
             uint32_t out_format) {
  __shared__ StagedSplat s_sp[CP_BATCH];
  __shared__ RawSplat s_raw[CP_BATCH];

  // XCD-aware order: consecutive workgroups go round-robin over the 8 XCDs, so give each XCD a contiguous band
  // of tiles (neighbouring tiles share splats, and each XCD has its own L2).
  const uint32_t ntiles = f.tiles_x * f.tiles_y;
  const uint32_t per = (ntiles + 7u) / 8u;
  const uint32_t tile = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  if (tile >= ntiles) return;
  const uint32_t tyi = tile / f.tiles_x, txi = tile - tyi * f.tiles_x;
  const uint32_t px = txi * 16u + (threadIdx.x & 15u), py = tyi * 16u + (threadIdx.x >> 4);
  const bool inside = px < f.width && py < f.height;
  const float nx = __fsub_rn(__fdiv_rn(__fmul_rn(__fadd_rn((float)px, 0.5f), 2.0f), (float)f.width), 1.0f);
  const float ny = __fsub_rn(1.0f, __fdiv_rn(__fmul_rn(__fadd_rn((float)py, 0.5f), 2.0f), (float)f.height));
  const float alpha_min = 1.0f / 255.0f;
  const float cmax = __uint_as_float(counters->color_max_bits);
  const float t_eps = 0x1p-12f / fmaxf(1.0f, cmax);
  const bool overdraw = f.view_mode == 2u;
  const bool depth_view = f.view_mode == 1u;

  const uint2 rg = ranges[tile];
  float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f, nfrag = 0.0f;
  bool done = !inside;
  for (uint32_t end = rg.y; end > rg.x; end -= min(end - rg.x, (uint32_t)CP_BATCH)) {
    const uint32_t nb = min(end - rg.x, (uint32_t)CP_BATCH);
    if (threadIdx.x < nb) {
      const uint32_t g = vals[end - 1u - threadIdx.x];
      const Splat s = splats[g];
      const float cx = s.center_axis0.x, cy = s.center_axis0.y;
      const float a0x = s.center_axis0.z, a0y = s.center_axis0.w;
      const float a1x = s.axis1_r_tau.x, a1y = s.axis1_r_tau.y;
      const float r = s.axis1_r_tau.z, tau0 = s.axis1_r_tau.w;
      const float det = a0x * a1y - a1x * a0y;
      const float inv = det != 0.0f ? 1.0f / det : 0.0f;
      StagedSplat o;
      o.cx = cx; o.cy = cy;
      o.ux = a1y * inv; o.uy = -a1x * inv;
      o.vx = -a0y * inv; o.vy = a0x * inv;
      o.kq = -0.5f * r * r * LOG2E;
      o.kt = det != 0.0f ? -tau0 * LOG2E : 0.0f;   // a degenerate quad covers nothing
      float cr_ = __uint_as_float(s.color_depth.x), cg_ = __uint_as_float(s.color_depth.y),
            cb_ = __uint_as_float(s.color_depth.z);
      if (depth_view) {
        const float depth = __uint_as_float(s.color_depth.w);
        float t = logf(depth / f.near_) / logf(f.far_ / f.near_);
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        cr_ = cg_ = cb_ = 1.0f - t;
      }
      o.r = cr_; o.g = cg_; o.b = cb_; o.pad = 0.0f;
      s_sp[threadIdx.x] = o;
      RawSplat rs;
      rs.a0x = a0x; rs.a0y = a0y; rs.a1x = a1x; rs.a1y = a1y;
      rs.det = det; rs.rr = r; rs.tau0 = tau0; rs.pad = 0.0f;
      s_raw[threadIdx.x] = rs;
    }
    __syncthreads();
    // wave vote: skip the sweep when every pixel of this wave has stopped
    if (__ballot(!done) != 0ull) {
      for (uint32_t k = 0; k < nb; k++) {
        if (done) break;
        const StagedSplat sp = s_sp[k];
        const float dx = __fsub_rn(nx, sp.cx), dy = __fsub_rn(ny, sp.cy);
        const float u = dx * sp.ux + dy * sp.uy;
        const float v = dx * sp.vx + dy * sp.vy;
        const float au = fabsf(u), av = fabsf(v);
        if (au > 1.0001f || av > 1.0001f) continue;
        const float e = __builtin_amdgcn_exp2f(sp.kq * (u * u + v * v));
        float alpha = 1.0f - __builtin_amdgcn_exp2f(sp.kt * e);
        if (au >= 0.9999f || av >= 0.9999f || fabsf(alpha - alpha_min) <= 1e-3f * alpha_min) {
          alpha = exact_alpha(dx, dy, s_raw[k]);
          if (alpha < 0.0f) continue;
        }
        if (alpha < alpha_min) continue;
        if (overdraw) { nfrag += 1.0f; continue; }
        const float wgt = T * alpha;
        cr += wgt * sp.r;
        cg += wgt * sp.g;
        cb += wgt * sp.b;
        T *= 1.0f - alpha;
        done = T < t_eps;
      }
    }
    if (!__syncthreads_or(!done)) break;
  }
  if (!inside) return;
  float4 px4 = overdraw ? make_float4(nfrag, 0.0f, 0.0f, 0.0f) : make_float4(cr, cg, cb, 1.0f - T);
  const size_t o = (size_t)py * f.width + px;
  if (out_format == 0u) {
    reinterpret_cast<float4 *>(out)[o] = px4;
  } else {
    const __half2 lo = __floats2half2_rn(px4.x, px4.y), hi = __floats2half2_rn(px4.z, px4.w);
    uint2 u;
    u.x = *reinterpret_cast<const uint32_t *>(&lo);
    u.y = *reinterpret_cast<const uint32_t *>(&hi);
    reinterpret_cast<uint2 *>(out)[o] = u;
  }
}
