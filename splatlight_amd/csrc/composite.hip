// Front-to-back compositing of 16x16 or 32x32 tiles (the blend of shaders/splat.wgsl:22-78 with the "over" state
// of src/renderer.js:140-142, evaluated per tile instead of by the rasteriser).
//
// Workgroups of 256 lanes, each over a 16x16 pixel block: one per tile for 16 px tiles, four per tile (one per
// quarter, each walking the whole tile's list) for 32 px tiles, whose binning has ~2.3x fewer pairs to emit and
// sort; a 1,024-lane workgroup per 32 px tile, staging each batch once for 4x the pixels, composited 1.3x slower
// (its 16 waves wait at every batch's barrier for the slowest block). Tiles are launched heaviest first; each wave
// owns an 8x8 pixel block, one lane per pixel. The tile's pair range is walked from its end, which is the
// reference's back-to-front sequence for every pixel reversed, in batches of CP_BATCH_T splats (64, 128 or 256,
// tuned): each lane stages one splat into LDS already reduced to what a pixel needs, then each wave culls the batch
// against its 8x8 block (64 splats per ballot, tight ellipse bounds) and sweeps only the survivors, in order. A pixel stops once its transmittance is below t_eps = 2^-12 / max(1, max|rgb|) (the
// colour it no longer adds is below 2^-12 absolute), a wave stops when all its pixels stopped, and the
// workgroup leaves when all waves did.
//
// Coverage. Pixel (px, py) has its centre at NDC ((px + 0.5) 2 / W - 1, 1 - (py + 0.5) 2 / H), rows top-down; it
// is covered by a splat when its centre lies in the quad c + u a0 + v a1, |u|, |v| <= 1, and its alpha
// 1 - exp(-tau0 exp(-r^2 (u^2 + v^2) / 2)) reaches 1/255. Since r^2 = 2 ln(tau0 / tau_min), alpha >= 1/255 is
// q = u^2 + v^2 <= 1 in exact arithmetic, and q <= 1 puts the centre inside the quad. The fast path therefore
// evaluates u and v with the exp2 scale folded into the inverse axes (its rounding error grows only linearly
// with the splat's aspect ratio), and a fragment with q within max(1e-4, 1e-6 aspect) of 1 is decided again with
// the exact expression of the CPU restatement (u, v by the 2x2 solve, exp as the correctly rounded float exp,
// no contraction), so no fragment lands on the other side of the discard.
#include "sl_device.h"
#include <hip/hip_fp16.h>

#define LOG2E 1.4426950408889634f

struct __align__(16) StagedSplat {
  float cx, cy, ux, uy;      // centre (NDC); u' = dx ux + dy uy = u sqrt(r^2 log2(e) / 2)
  float vx, vy, lo, hi;      // v' likewise; s = u'^2 + v'^2: s > hi skip, s >= lo exact, else exp2(-s)
};

struct __align__(16) SplatBounds {   // tight ellipse bounds (NDC), widened by a hair
  float x0, x1, y0, y1;
};

struct __align__(8) SplatExact {      // the splat's own axis, radius and density words, for the exact path
  float a0x, a0y, a1x, a1y, rr, tau0;
};

// The exact coverage of one fragment of a splat from its own words (staged in LDS beside the fast form; the case
// is rare): the CPU restatement's expression order, no contraction, correctly rounded exp; alpha, or -1 when the
// pixel centre is outside the quad.
__device__ __noinline__ float exact_alpha(float dx, float dy, const SplatExact *sp) {
#pragma clang fp contract(off)
  const SplatExact s = *sp;
  const float a0x = s.a0x, a0y = s.a0y, a1x = s.a1x, a1y = s.a1y;
  const float rr = s.rr, tau0 = s.tau0;
  const float det = a0x * a1y - a1x * a0y;
  const float u = (dx * a1y - a1x * dy) / det;
  const float v = (a0x * dy - dx * a0y) / det;
  if (fabsf(u) > 1.0f || fabsf(v) > 1.0f) return -1.0f;
  const float lx = u * rr, ly = v * rr;
  const float e = (float)exp((double)(-0.5f * (lx * lx + ly * ly)));
  return 1.0f - (float)exp((double)(-tau0 * e));
}

// One wave's sweep over a staged batch of nb splats (front to back): cull 64 at a time against the wave's block,
// then two survivors per step. A pixel is live while !(T < t_eps) (pixels outside the view start at T = 0, so they
// never are); a step no live lane takes is skipped, and the exact re-evaluation near the coverage edge is one
// wave-uniform branch for both splats. OVERDRAW counts the fragments instead of blending them.
template <bool OVERDRAW>
__device__ __forceinline__ void sweep_batch(uint32_t nb, uint32_t lane, const StagedSplat *s_sp, const float4 *s_col,
                                            const SplatBounds *s_bd, const SplatExact *s_ex, float nx, float ny, float wx0,
                                            float wx1, float wy0, float wy1, float t_eps, float &T, float &cr,
                                            float &cg, float &cb, float &nfrag) {
  const float alpha_min = 1.0f / 255.0f;
  for (uint32_t k0 = 0; k0 < nb; k0 += 64u) {
    if (__ballot(!(T < t_eps)) == 0ull) return;       // every pixel of this wave has stopped
    const uint32_t kk = k0 + lane;
    bool hit = false;
    if (kk < nb) {
      const SplatBounds b = s_bd[kk];
      hit = b.x0 <= wx1 && b.x1 >= wx0 && b.y0 <= wy1 && b.y1 >= wy0;
    }
    uint64_t mask = __ballot(hit);
    while (mask) {
      const uint32_t ka = k0 + (uint32_t)__ffsll((unsigned long long)mask) - 1u;
      mask &= mask - 1ull;
      const bool two = mask != 0ull;
      const uint32_t kb = two ? k0 + (uint32_t)__ffsll((unsigned long long)mask) - 1u : ka;
      mask = two ? mask & (mask - 1ull) : mask;
      const StagedSplat pa = s_sp[ka], pb = s_sp[kb];
      const float4 ca = s_col[ka], cb4 = s_col[kb];
      const float dxa = __fsub_rn(nx, pa.cx), dya = __fsub_rn(ny, pa.cy);
      const float dxb = __fsub_rn(nx, pb.cx), dyb = __fsub_rn(ny, pb.cy);
      const float ua = dxa * pa.ux + dya * pa.uy, va = dxa * pa.vx + dya * pa.vy;
      const float ub = dxb * pb.ux + dyb * pb.uy, vb = dxb * pb.vx + dyb * pb.vy;
      const float sa = ua * ua + va * va, sb = ub * ub + vb * vb;
      bool ina = sa <= pa.hi, inb = two && sb <= pb.hi;  // within the band (or inside)
      const bool live = !(T < t_eps);
      float alpha_a = 1.0f - __builtin_amdgcn_exp2f(ca.w * __builtin_amdgcn_exp2f(-sa));
      float alpha_b = 1.0f - __builtin_amdgcn_exp2f(cb4.w * __builtin_amdgcn_exp2f(-sb));
      const bool xa = live && ina && sa >= pa.lo, xb = live && inb && sb >= pb.lo;
      if (__ballot(xa || xb) != 0ull) {                  // q in the band around 1: decide exactly (rare)
        if (xa) {
          alpha_a = exact_alpha(dxa, dya, s_ex + ka);
          ina = alpha_a >= alpha_min;
        }
        if (xb) {
          alpha_b = exact_alpha(dxb, dyb, s_ex + kb);
          inb = alpha_b >= alpha_min;
        }
      }
      if (OVERDRAW) {
        nfrag = live && ina ? nfrag + 1.0f : nfrag;
        nfrag = live && inb ? nfrag + 1.0f : nfrag;
        continue;
      }
      bool use = live && ina;
      float wgt = T * alpha_a;
      cr = use ? cr + wgt * ca.x : cr;
      cg = use ? cg + wgt * ca.y : cg;
      cb = use ? cb + wgt * ca.z : cb;
      T = use ? T - wgt : T;
      use = !(T < t_eps) && inb;
      wgt = T * alpha_b;
      cr = use ? cr + wgt * cb4.x : cr;
      cg = use ? cg + wgt * cb4.y : cg;
      cb = use ? cb + wgt * cb4.z : cb;
      T = use ? T - wgt : T;
    }
  }
}

template <uint32_t CP_BATCH_T, uint32_t TSH, uint32_t WSH>
__global__ void __launch_bounds__((1u << WSH) * (1u << WSH), 8)
sl_composite(const ViewDev *__restrict__ views, const uint32_t *__restrict__ tile_base, uint32_t n_views,
             const ViewCounters *__restrict__ view_counters, const Splat *__restrict__ splats,
             const uint32_t *__restrict__ vals, const uint2 *__restrict__ ranges,
             const uint32_t *__restrict__ tile_order) {
  __shared__ StagedSplat s_sp[CP_BATCH_T];
  __shared__ float4 s_col[CP_BATCH_T];
  __shared__ SplatBounds s_bd[CP_BATCH_T];
  __shared__ SplatExact s_ex[CP_BATCH_T];             // the splat's own words, for the exact re-evaluation

  // workgroup b takes the b-th heaviest tile of the batch (sl_tile_order): longest-first over the XCDs
  static_assert(WSH <= TSH, "a workgroup covers at most one tile");
  constexpr uint32_t SUBS = 1u << (TSH - WSH);           // workgroups per tile side
  const uint32_t gtile = tile_order[blockIdx.x / (SUBS * SUBS)];
  const uint32_t sub = blockIdx.x % (SUBS * SUBS);
  const uint32_t v = n_views > 1u ? view_of(tile_base, n_views, gtile) : 0u;
  const FrameDev &f = views[v].f;
  void *__restrict__ out = views[v].out;
  const uint32_t out_format = views[v].out_format;
  const uint32_t tile = gtile - tile_base[v];
  const uint32_t tyi = tile / f.tiles_x, txi = tile - tyi * f.tiles_x;
  constexpr uint32_t TS = 1u << TSH, WS = 1u << WSH, WROW = WS / 8u;   // tile side, workgroup side; 8x8 wave blocks
  const uint32_t gx = txi * TS + (sub % SUBS) * WS, gy = tyi * TS + (sub / SUBS) * WS;   // this workgroup's block
  if (gx >= f.width || gy >= f.height) return;           // a block wholly outside the view (uniform)
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const uint32_t bx = gx + (w % WROW) * 8u, by = gy + (w / WROW) * 8u;   // this wave's 8x8 block
  const uint32_t px = bx + (lane & 7u), py = by + (lane >> 3);
  const bool inside = px < f.width && py < f.height;
  const float fw = (float)f.width, fh = (float)f.height;
  const float nx = __fsub_rn(__fdiv_rn(__fmul_rn(__fadd_rn((float)px, 0.5f), 2.0f), fw), 1.0f);
  const float ny = __fsub_rn(1.0f, __fdiv_rn(__fmul_rn(__fadd_rn((float)py, 0.5f), 2.0f), fh));
  // the wave's block in NDC (pixel centres)
  const float wx0 = ((float)bx + 0.5f) * 2.0f / fw - 1.0f, wx1 = ((float)bx + 7.5f) * 2.0f / fw - 1.0f;
  const float wy1 = 1.0f - ((float)by + 0.5f) * 2.0f / fh, wy0 = 1.0f - ((float)by + 7.5f) * 2.0f / fh;
  const float cmax = __uint_as_float(view_counters[v].color_max_bits);
  const float t_eps = 0x1p-12f / fmaxf(1.0f, cmax);
  const bool overdraw = f.view_mode == 2u;
  const bool depth_view = f.view_mode == 1u;

  const uint2 rg = ranges[gtile];
  float T = inside ? 1.0f : 0.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f, nfrag = 0.0f;   // outside: never live
  for (uint32_t end = rg.y; end > rg.x; end -= min(end - rg.x, CP_BATCH_T)) {
    const uint32_t nb = min(end - rg.x, CP_BATCH_T);
    if (threadIdx.x < nb) {
      const uint32_t g = vals[end - 1u - threadIdx.x];
      const Splat s = splats[g];
      const float cx = s.center_axis0.x, cy = s.center_axis0.y;
      const float a0x = s.center_axis0.z, a0y = s.center_axis0.w;
      const float a1x = s.axis1_r_tau.x, a1y = s.axis1_r_tau.y;
      const float r = s.axis1_r_tau.z, tau0 = s.axis1_r_tau.w;
      const float det = a0x * a1y - a1x * a0y;
      const bool ok = det != 0.0f;                       // a degenerate quad covers nothing
      const float inv = ok ? 1.0f / det : 0.0f;
      const float kq = 0.5f * r * r * LOG2E;            // s = kq q
      const float sk = sqrtf(kq) * inv;
      const float la0 = sqrtf(a0x * a0x + a0y * a0y), la1 = sqrtf(a1x * a1x + a1y * a1y);
      const float aspect = fmaxf(la0, la1) / fmaxf(fminf(la0, la1), 1e-30f);
      const float m = fmaxf(1e-4f, 1e-6f * aspect);     // the band decided exactly
      StagedSplat o;
      o.cx = cx; o.cy = cy;
      o.ux = a1y * sk; o.uy = -a1x * sk;
      o.vx = -a0y * sk; o.vy = a0x * sk;
      o.lo = kq * (1.0f - m);
      o.hi = kq * (1.0f + m);
      s_sp[threadIdx.x] = o;
      float cr_ = __uint_as_float(s.color_depth.x), cg_ = __uint_as_float(s.color_depth.y),
            cb_ = __uint_as_float(s.color_depth.z);
      if (depth_view) {
        const float depth = __uint_as_float(s.color_depth.w);
        float t = logf(depth / f.near_) / logf(f.far_ / f.near_);
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        cr_ = cg_ = cb_ = 1.0f - t;
      }
      s_col[threadIdx.x] = make_float4(cr_, cg_, cb_, -tau0 * LOG2E);
      const float ex = sqrtf(a0x * a0x + a1x * a1x) * (1.0f + 2.0f * m) + 1e-7f;
      const float ey = sqrtf(a0y * a0y + a1y * a1y) * (1.0f + 2.0f * m) + 1e-7f;
      SplatBounds b;
      b.x0 = cx - ex; b.x1 = cx + ex; b.y0 = cy - ey; b.y1 = cy + ey;
      if (!ok) { b.x0 = 2.0f; b.x1 = -2.0f; }            // culled by every wave
      s_bd[threadIdx.x] = b;
      s_ex[threadIdx.x] = SplatExact{a0x, a0y, a1x, a1y, r, tau0};
    }
    __syncthreads();
    if (overdraw)
      sweep_batch<true>(nb, lane, s_sp, s_col, s_bd, s_ex, nx, ny, wx0, wx1, wy0, wy1, t_eps, T, cr, cg, cb,
                        nfrag);
    else
      sweep_batch<false>(nb, lane, s_sp, s_col, s_bd, s_ex, nx, ny, wx0, wx1, wy0, wy1, t_eps, T, cr, cg, cb,
                         nfrag);
    if (!__syncthreads_or(!(T < t_eps))) break;
  }
  if (!inside) return;
  const float4 px4 = overdraw ? make_float4(nfrag, 0.0f, 0.0f, 0.0f) : make_float4(cr, cg, cb, 1.0f - T);
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
