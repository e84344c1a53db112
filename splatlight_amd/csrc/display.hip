// Display pass: the linear HDR target resolved to 8-bit display colour (shaders/display.wgsl:22-189) — in the lit
// view the environment's background sky and the distant lights' discs fill what the splats leave uncovered, then
// exposure, the tone curve and the sRGB encode; the overdraw view maps the count to a heat ramp.
//
// One 256-lane workgroup per 1024 pixels, four per lane (the next one's load in flight while one is evaluated), rows
// top-down. 8 or 16 B read and 4 B written per pixel; where the sky shows, the 81-coefficient background sky makes
// the pass ALU-bound. The uniforms are kernel arguments, read through the scalar cache; the background sky's loops
// are unrolled for the full nine bands, so its coefficient loads are scalar loads at fixed offsets.
//
// Exactness. The bytes are those of the shader's float32 expressions with exp, pow, log2, cos and sin correctly
// rounded (the precision WGSL leaves to the implementation, fixed), stored as floor(255 v + 0.5). A first, fast
// evaluation uses the hardware's exp2 / log2 and multiplies by reciprocals where the shader divides by small
// integers; it lands within 2e-6 of the exact value, so only a channel whose 255 v + 0.5 lies within 2^-9 of an
// integer can round differently: such a pixel goes to a list, and a second kernel evaluates the listed pixels
// (a few percent) densely, in the shader's exact operations with the correctly rounded functions ((float) of the
// double function). SL_DISPLAY_EXACT evaluates every pixel so in the first kernel.
#include "sl_device.h"

#define DP_THREADS 256
#define DP_PER_LANE 4
#define DP_PX (DP_THREADS * DP_PER_LANE)
#define DP_REGIONS 64u

// The uniforms the display pass reads, by value: the frame's camera and display words, Env (416 B) as the host
// wrote it, the SkyBg words after its coefficients (info, disc_std, ring, ground, ring_par, glow, glow_dir) and each
// horizon lobe's direction (cos, sin of 2 pi k / n, rounded once). Every read of them is at a constant offset (the
// loops over lights and lobes are unrolled), so they stay scalar loads from the arguments.
struct DisplayArgs {
  float rot[9];           // view[0..2].xyz (the view matrix's first three columns)
  float vw, vh, fx, fy, exposure;
  uint32_t view_mode, tone;
  uint32_t width, height;
  uint32_t lmax;          // the background sky's last band (SkyBg.info.x > 0)
  float ring_cs[16];
  float env[104];
  float bg_tail[56];      // SkyBg words 324-379
};
#define DP_E(k) a.env[(k)]
#define DP_EU(k) __float_as_uint(a.env[(k)])
#define DP_B(k) a.bg_tail[(k) - 324]
#define DP_BU(k) __float_as_uint(a.bg_tail[(k) - 324])

// The background sky's coefficients as the pass reads them, from a device table staged into LDS by every
// workgroup: for band l and order m >= 0 (entry e = l (l + 1) / 2 + m, 45 entries) {SkyBg.sh[l*l + l + m].rgb, f}
// with f the recurrence's factor k (times 1.41421356 for m > 0; display.wgsl:49-62), and at 45 + e
// {SkyBg.sh[l*l + l - m].rgb, 0}. f depends on (l, m) only; the host forms it with the shader's float operations
// in its order, so it is the value the shader's per-pixel loop forms.
#define DP_SH_ENTRIES 90

// The fast path's exp, pow, log2 and division by a small integer: the hardware's exp2 and log2 (v_exp_f32,
// v_log_f32, about 1 ulp) with exp(x) = exp2(x log2 e) and pow(x, y) = exp2(y log2 x) for x >= 0, and x / d as
// x rcp(d) (v_rcp_f32).
struct FastFn {
  static __device__ __forceinline__ float ex(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
  static __device__ __forceinline__ float pw(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
  }
  static __device__ __forceinline__ float lg2(float x) { return __builtin_amdgcn_logf(x); }
  static __device__ __forceinline__ float div(float x, float d) { return x * __builtin_amdgcn_rcpf(d); }
  static constexpr bool fast = true;
};
// The exact path: correctly rounded functions and the shader's own division.
struct ExactFn {
  static __device__ __forceinline__ float ex(float x) { return (float)exp((double)x); }
  static __device__ __forceinline__ float pw(float x, float y) { return (float)pow((double)x, (double)y); }
  static __device__ __forceinline__ float lg2(float x) { return (float)log2((double)x); }
  static __device__ __forceinline__ float div(float x, float d) { return x / d; }
  static constexpr bool fast = false;
};

__device__ __forceinline__ float dp_clamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// display.wgsl:22-32
__device__ __forceinline__ float3 dp_sky_radiance(const DisplayArgs &a, float3 w) {
#pragma clang fp contract(off)
  const float x = w.x, y = w.y, z = w.z;
  const float b[16] = {0.282095f, 0.488603f * y, 0.488603f * z, 0.488603f * x,
                       1.092548f * x * y, 1.092548f * y * z, 0.315392f * (3.0f * z * z - 1.0f), 1.092548f * x * z,
                       0.546274f * (x * x - y * y),
                       0.590044f * y * (3.0f * x * x - y * y), 2.890611f * x * y * z,
                       0.457046f * y * (5.0f * z * z - 1.0f), 0.373176f * z * (5.0f * z * z - 3.0f),
                       0.457046f * x * (5.0f * z * z - 1.0f), 1.445306f * z * (x * x - y * y),
                       0.590044f * x * (x * x - 3.0f * y * y)};
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    cx = cx + b[k] * DP_E(4 * k + 0);
    cy = cy + b[k] * DP_E(4 * k + 1);
    cz = cz + b[k] * DP_E(4 * k + 2);
  }
  return make_float3(fmaxf(cx, 0.0f), fmaxf(cy, 0.0f), fmaxf(cz, 0.0f));
}

// display.wgsl:36-80: real SH by recurrence (no Condon-Shortley phase), the glow SG, the horizon ring, the ground.
// sh: the coefficient table in LDS. The exact path adds each (l, m) term as the shader does; the fast path sums each
// order's terms in l first (fused multiply-adds, the division by l - m as a reciprocal) and applies cos / sin of
// m phi once per order.
template <class F>
__device__ float3 dp_sky_background(const DisplayArgs &a, const float4 *__restrict__ sh, float3 w) {
#pragma clang fp contract(off)
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
  float cm = 1.0f, sm = 0.0f, pmm = 1.0f;
  const uint32_t lmax = a.lmax;
  for (uint32_t m = 0; m <= lmax; m++) {
    if (m > 0u) {
      const float c2 = cm * w.x - sm * w.y;
      sm = cm * w.y + sm * w.x;
      cm = c2;
      pmm *= (float)(2u * m - 1u);
    }
    float p_prev = 0.0f, p = pmm;
    float ux = 0.0f, uy = 0.0f, uz = 0.0f, vx = 0.0f, vy = 0.0f, vz = 0.0f;
    uint32_t e = m * (m + 3u) / 2u;   // entry (l = m, m)
    for (uint32_t l = m; l <= lmax; e += l + 1u, l++) {
      if (l == m + 1u) {
        p_prev = p;
        p = w.z * (float)(2u * m + 1u) * pmm;
      } else if (l > m + 1u) {
        const float pn = F::div((float)(2u * l - 1u) * w.z * p - (float)(l + m - 1u) * p_prev, (float)(l - m));
        p_prev = p;
        p = pn;
      }
      const float4 A = sh[e];
      if (F::fast) {
        const float kp = A.w * p;
        ux = fmaf(kp, A.x, ux);
        uy = fmaf(kp, A.y, uy);
        uz = fmaf(kp, A.z, uz);
        if (m > 0u) {
          const float4 B = sh[45u + e];
          vx = fmaf(kp, B.x, vx);
          vy = fmaf(kp, B.y, vy);
          vz = fmaf(kp, B.z, vz);
        }
      } else if (m == 0u) {
        cx = cx + A.w * p * A.x;
        cy = cy + A.w * p * A.y;
        cz = cz + A.w * p * A.z;
      } else {
        const float4 B = sh[45u + e];
        const float s = A.w * p;
        cx = cx + s * (cm * A.x + sm * B.x);
        cy = cy + s * (cm * A.y + sm * B.y);
        cz = cz + s * (cm * A.z + sm * B.z);
      }
    }
    if (F::fast) {
      cx = fmaf(cm, ux, fmaf(sm, vx, cx));
      cy = fmaf(cm, uy, fmaf(sm, vy, cy));
      cz = fmaf(cm, uz, fmaf(sm, vz, cz));
    }
  }
  cx = fmaxf(cx, 0.0f);
  cy = fmaxf(cy, 0.0f);
  cz = fmaxf(cz, 0.0f);
  if (DP_B(4 * 93 + 3) > 0.0f) {   // the glow
    const float s2 = DP_B(4 * 93 + 3) * DP_B(4 * 93 + 3);
    const float e = F::ex((w.x * DP_B(4 * 94) + w.y * DP_B(4 * 94 + 1) + w.z * DP_B(4 * 94 + 2) - 1.0f) / s2);
    const float d = (float)(2.0 * 3.14159265) * s2;
    cx = cx + DP_B(4 * 93 + 0) / d * e;
    cy = cy + DP_B(4 * 93 + 1) / d * e;
    cz = cz + DP_B(4 * 93 + 2) / d * e;
  }
  const uint32_t nr = min(DP_BU(4 * 81 + 1), 8u);
  if (nr == 0u) return make_float3(cx, cy, cz);
  const float px = DP_B(4 * 92), py = DP_B(4 * 92 + 1);
#pragma unroll
  for (uint32_t k = 0; k < 8u; k++) {   // the horizon ring
    if (k >= nr) break;
    const float zx = a.ring_cs[2 * k], zy = a.ring_cs[2 * k + 1];
    const float along = w.x * -zy + w.y * zx;
    const float e = F::ex(-px * along * along - py * w.z * w.z);
    const float f = fmaxf(w.x * zx + w.y * zy, 0.0f);
    cx = cx + DP_B(4 * (83 + k) + 0) * f * e;
    cy = cy + DP_B(4 * (83 + k) + 1) * f * e;
    cz = cz + DP_B(4 * (83 + k) + 2) * f * e;
  }
  const float t = dp_clamp(w.z / DP_B(4 * 91 + 3) * 0.5f + 0.5f, 0.0f, 1.0f);   // the ground below the horizon
  return make_float3(DP_B(4 * 91 + 0) * (1.0f - t) + cx * t, DP_B(4 * 91 + 1) * (1.0f - t) + cy * t,
                     DP_B(4 * 91 + 2) * (1.0f - t) + cz * t);
}

// display.wgsl:84-93
template <class F>
__device__ __forceinline__ float3 dp_lights(const DisplayArgs &a, float3 w) {
#pragma clang fp contract(off)
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
  const uint32_t n = min(DP_EU(96), 4u);
#pragma unroll
  for (uint32_t k = 0; k < 4u; k++) {
    if (k >= n) break;
    const float dstd = DP_B(4 * 82 + k);
    const float sd = dstd > 0.0f ? dstd : DP_E(64 + 8 * k + 3);
    const float s2 = fmaxf(sd * sd, 1e-8f);
    const float e = F::ex((w.x * DP_E(64 + 8 * k) + w.y * DP_E(64 + 8 * k + 1) + w.z * DP_E(64 + 8 * k + 2) - 1.0f) /
                          s2);
    const float d = (float)(2.0 * 3.14159265) * s2;
    cx = cx + DP_E(68 + 8 * k + 0) / d * e;
    cy = cy + DP_E(68 + 8 * k + 1) / d * e;
    cz = cz + DP_E(68 + 8 * k + 2) / d * e;
  }
  return make_float3(cx, cy, cz);
}

__device__ __forceinline__ float3 dp_norm(float3 v) {
#pragma clang fp contract(off)
  const float l = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  return make_float3(v.x / l, v.y / l, v.z / l);
}

// display.wgsl:95-99 at the pixel centre (px + 0.5, py + 0.5)
__device__ __forceinline__ float3 dp_view_dir(const DisplayArgs &a, float px, float py) {
#pragma clang fp contract(off)
  const float3 v = dp_norm(make_float3((px - 0.5f * a.vw) / a.fx, -(py - 0.5f * a.vh) / a.fy, -1.0f));
  const float *r = a.rot;
  return dp_norm(make_float3(r[0] * v.x + r[1] * v.y + r[2] * v.z, r[3] * v.x + r[4] * v.y + r[5] * v.z,
                             r[6] * v.x + r[7] * v.y + r[8] * v.z));
}

template <class F>
__device__ __forceinline__ float dp_srgb(float c) {
#pragma clang fp contract(off)
  if (c <= 0.0031308f) return c * 12.92f;
  return 1.055f * F::pw(c, (float)(1.0 / 2.4)) - 0.055f;
}

__device__ __forceinline__ float dp_hable(float x) {
#pragma clang fp contract(off)
  return (x * (0.15f * x + 0.05f) + 0.004f) / (x * (0.15f * x + 0.5f) + 0.06f) - (float)(0.02 / 0.3);
}

__device__ __forceinline__ float dp_aces(float x) {
#pragma clang fp contract(off)
  return dp_clamp(x * (2.51f * x + 0.03f) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f, 1.0f);
}

template <class F>
__device__ __forceinline__ float dp_agx_curve(float xin) {
#pragma clang fp contract(off)
  const float lo = -12.47393f, hi = 4.026069f;
  float x = dp_clamp(F::lg2(fmaxf(xin, 1e-10f)), lo, hi);
  x = (x - lo) / (hi - lo);
  const float x2 = x * x, x4 = x2 * x2;
  return 15.5f * x4 * x2 - 40.14f * x4 * x + 31.96f * x4 - 6.868f * x2 * x + 0.4298f * x2 + 0.1191f * x - 0.00232f;
}

// display.wgsl:140-189
template <class F>
__device__ float3 dp_tone(float3 x, uint32_t mode) {
#pragma clang fp contract(off)
  if (mode == 1u) {   // AgX: inset (columns of m), log2 sigmoid, outset (columns of mi), 2.2 power
    const float ax = 0.842479062253094f * x.x + 0.0784335999999992f * x.y + 0.0792237451477643f * x.z;
    const float ay = 0.0423282422610123f * x.x + 0.878468636469772f * x.y + 0.0791661274605434f * x.z;
    const float az = 0.0423756549057051f * x.x + 0.0784336f * x.y + 0.879142973793104f * x.z;
    const float cx = dp_agx_curve<F>(ax), cy = dp_agx_curve<F>(ay), cz = dp_agx_curve<F>(az);
    const float yx = 1.19687900512017f * cx + -0.0980208811401368f * cy + -0.0990297440797205f * cz;
    const float yy = -0.0528968517574562f * cx + 1.15190312990417f * cy + -0.0989611768448433f * cz;
    const float yz = -0.0529716355144438f * cx + -0.0980434501171241f * cy + 1.15107367264116f * cz;
    return make_float3(dp_clamp(F::pw(fmaxf(yx, 0.0f), 2.2f), 0.0f, 1.0f),
                       dp_clamp(F::pw(fmaxf(yy, 0.0f), 2.2f), 0.0f, 1.0f),
                       dp_clamp(F::pw(fmaxf(yz, 0.0f), 2.2f), 0.0f, 1.0f));
  }
  if (mode == 2u) return make_float3(dp_aces(x.x), dp_aces(x.y), dp_aces(x.z));
  if (mode == 3u) {
    const float l = x.x * 0.2126f + x.y * 0.7152f + x.z * 0.0722f;
    return make_float3(dp_clamp(x.x / (1.0f + l), 0.0f, 1.0f), dp_clamp(x.y / (1.0f + l), 0.0f, 1.0f),
                       dp_clamp(x.z / (1.0f + l), 0.0f, 1.0f));
  }
  if (mode == 4u) {   // Khronos PBR Neutral
    const float start = (float)(0.8 - 0.04), desat = 0.15f;
    const float lo = fminf(x.x, fminf(x.y, x.z));
    const float off = lo < 0.08f ? lo - 6.25f * lo * lo : 0.04f;
    float3 y = make_float3(x.x - off, x.y - off, x.z - off);
    const float peak = fmaxf(y.x, fmaxf(y.y, y.z));
    if (peak < start)
      return make_float3(dp_clamp(y.x, 0.0f, 1.0f), dp_clamp(y.y, 0.0f, 1.0f), dp_clamp(y.z, 0.0f, 1.0f));
    const float d = 1.0f - start;
    const float np = 1.0f - d * d / (peak + d - start);
    const float s = np / peak;
    y = make_float3(y.x * s, y.y * s, y.z * s);
    const float g = 1.0f - 1.0f / (desat * (peak - np) + 1.0f);
    return make_float3(dp_clamp(y.x * (1.0f - g) + np * g, 0.0f, 1.0f), dp_clamp(y.y * (1.0f - g) + np * g, 0.0f, 1.0f),
                       dp_clamp(y.z * (1.0f - g) + np * g, 0.0f, 1.0f));
  }
  if (mode == 5u) {   // Hable, bias 2, white 11.2
    const float wt = dp_hable(11.2f);
    return make_float3(dp_clamp(dp_hable(2.0f * x.x) / wt, 0.0f, 1.0f), dp_clamp(dp_hable(2.0f * x.y) / wt, 0.0f, 1.0f),
                       dp_clamp(dp_hable(2.0f * x.z) / wt, 0.0f, 1.0f));
  }
  return make_float3(dp_clamp(x.x, 0.0f, 1.0f), dp_clamp(x.y, 0.0f, 1.0f), dp_clamp(x.z, 0.0f, 1.0f));
}

// display.wgsl:119-135 for pixel i with HDR value c: the encoded colour before the 8-bit store
template <class F>
__device__ float3 dp_pixel(const DisplayArgs &a, const float4 *__restrict__ sh, uint32_t i, float4 c) {
#pragma clang fp contract(off)
  if (a.view_mode == 2u) {   // heat: black -> blue -> green -> yellow -> red over 0..128 fragments (log)
    const float x = dp_clamp(F::lg2(1.0f + c.x) / 7.0f, 0.0f, 1.0f) * 4.0f;
    return make_float3(dp_clamp(x - 2.0f, 0.0f, 1.0f), dp_clamp(fminf(x - 1.0f, 4.0f - x), 0.0f, 1.0f),
                       dp_clamp(fminf(x, 2.0f - x), 0.0f, 1.0f));
  }
  float3 rgb = make_float3(c.x, c.y, c.z);
  if (a.view_mode == 3u && (DP_EU(96) > 0u || DP_EU(97) != 0u)) {
    const float bare = 1.0f - dp_clamp(c.w, 0.0f, 1.0f);
    if (bare > 1.0f / 1024.0f) {
      const float3 w = dp_view_dir(a, (float)(i % a.width) + 0.5f, (float)(i / a.width) + 0.5f);
      float3 s = make_float3(0.0f, 0.0f, 0.0f);
      if (DP_EU(97) != 0u) {
        s = DP_BU(4 * 81) == 0u ? dp_sky_radiance(a, w) : dp_sky_background<F>(a, sh, w);
      }
      const float3 l = dp_lights<F>(a, w);
      rgb = make_float3(rgb.x + bare * (s.x + l.x), rgb.y + bare * (s.y + l.y), rgb.z + bare * (s.z + l.z));
    }
  }
  const float3 t = dp_tone<F>(make_float3(fmaxf(rgb.x * a.exposure, 0.0f), fmaxf(rgb.y * a.exposure, 0.0f),
                                          fmaxf(rgb.z * a.exposure, 0.0f)), a.tone);
  return make_float3(dp_srgb<F>(t.x), dp_srgb<F>(t.y), dp_srgb<F>(t.z));
}

__device__ __forceinline__ float dp_scaled(float v) {   // 255 v + 0.5 of the clamped value: the byte is its floor
#pragma clang fp contract(off)
  return dp_clamp(v, 0.0f, 1.0f) * 255.0f + 0.5f;
}

__device__ __forceinline__ bool dp_near_boundary(float3 o) {
  const float e = 1.0f / 512.0f;
  const float s[3] = {dp_scaled(o.x), dp_scaled(o.y), dp_scaled(o.z)};
  bool near = false;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float f = s[k] - floorf(s[k]);
    near |= f < e || f > 1.0f - e;
  }
  return near;
}

template <bool BGRA>
__device__ __forceinline__ void dp_store(uchar4 *out, uint32_t i, float3 o) {
  const unsigned char r = (unsigned char)floorf(dp_scaled(o.x)), g = (unsigned char)floorf(dp_scaled(o.y)),
                      b = (unsigned char)floorf(dp_scaled(o.z));
  out[i] = BGRA ? make_uchar4(b, g, r, 255) : make_uchar4(r, g, b, 255);
}

template <bool HALF_IN>
__device__ __forceinline__ float4 dp_load(const void *hdr, uint32_t i) {
  if (HALF_IN) {
    const uint2 h = reinterpret_cast<const uint2 *>(hdr)[i];
    return make_float4(__half2float(__ushort_as_half((unsigned short)(h.x & 0xffffu))),
                       __half2float(__ushort_as_half((unsigned short)(h.x >> 16))),
                       __half2float(__ushort_as_half((unsigned short)(h.y & 0xffffu))),
                       __half2float(__ushort_as_half((unsigned short)(h.y >> 16))));
  }
  return reinterpret_cast<const float4 *>(hdr)[i];
}

// The pass. EXACT: every pixel exactly. Else the fast evaluation, with the pixels it cannot settle collected per
// workgroup in LDS and appended to one of DP_REGIONS lists (workgroup b to region b % DP_REGIONS: list r holds
// q_cap words at q_list + r q_cap, its count in q_count[r], cleared before the launch; spreading the appends over
// the regions keeps the workgroups' atomics off one address), which sl_display_settle evaluates exactly
// afterwards, densely packed. sh_table: the DP_SH_ENTRIES coefficient entries.
template <bool HALF_IN, bool BGRA, bool EXACT>
__global__ void __launch_bounds__(DP_THREADS) sl_display(const void *__restrict__ hdr, uchar4 *__restrict__ out,
                                                         const float4 *__restrict__ sh_table,
                                                         uint32_t *__restrict__ q_count,
                                                         uint32_t *__restrict__ q_list, uint32_t q_cap,
                                                         DisplayArgs a) {
  __shared__ float4 s_sh[DP_SH_ENTRIES];
  __shared__ uint32_t s_q[DP_PX];
  __shared__ uint32_t s_n, s_base;
  const uint32_t npx = a.width * a.height;
  const uint32_t base = blockIdx.x * DP_PX + threadIdx.x;
  // this lane's four pixels, all loads issued before any is evaluated (rotated through c0 as they are)
  const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 c0 = base < npx ? dp_load<HALF_IN>(hdr, base) : zero;
  float4 c1 = base + DP_THREADS < npx ? dp_load<HALF_IN>(hdr, base + DP_THREADS) : zero;
  float4 c2 = base + 2u * DP_THREADS < npx ? dp_load<HALF_IN>(hdr, base + 2u * DP_THREADS) : zero;
  float4 c3 = base + 3u * DP_THREADS < npx ? dp_load<HALF_IN>(hdr, base + 3u * DP_THREADS) : zero;
  if (threadIdx.x < DP_SH_ENTRIES) s_sh[threadIdx.x] = sh_table[threadIdx.x];
  if (threadIdx.x == 0) s_n = 0u;
  __syncthreads();
#pragma unroll 1
  for (uint32_t k = 0; k < DP_PER_LANE; k++) {
    const uint32_t i = base + k * DP_THREADS;
    if (i >= npx) break;
    const float4 ck = c0;
    c0 = c1;
    c1 = c2;
    c2 = c3;
    if (EXACT) {
      dp_store<BGRA>(out, i, dp_pixel<ExactFn>(a, s_sh, i, ck));
    } else {
      const float3 o = dp_pixel<FastFn>(a, s_sh, i, ck);
      if (dp_near_boundary(o)) s_q[atomicAdd(&s_n, 1u)] = i;   // at most DP_PX entries: one per pixel
      else dp_store<BGRA>(out, i, o);
    }
  }
  if (EXACT) return;
  __syncthreads();
  const uint32_t n = s_n;
  if (n == 0u) return;
  // a region's list holds q_cap = DP_PX ceil(blocks / DP_REGIONS) words: every pixel of its workgroups at most once
  const uint32_t r = blockIdx.x % DP_REGIONS;
  if (threadIdx.x == 0) s_base = r * q_cap + atomicAdd(q_count + r, n);
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < n; j += DP_THREADS) q_list[s_base + j] = s_q[j];
}

// The exact evaluation of the listed pixels, one lane each: grid (x, DP_REGIONS), workgroup (c, r) takes entries
// c * DP_THREADS + lane of list r, striding by gridDim.x * DP_THREADS.
template <bool HALF_IN, bool BGRA>
__global__ void __launch_bounds__(DP_THREADS) sl_display_settle(const void *__restrict__ hdr,
                                                                uchar4 *__restrict__ out,
                                                                const float4 *__restrict__ sh_table,
                                                                const uint32_t *__restrict__ q_count,
                                                                const uint32_t *__restrict__ q_list, uint32_t q_cap,
                                                                DisplayArgs a) {
  __shared__ float4 s_sh[DP_SH_ENTRIES];
  const uint32_t r = blockIdx.y, n = q_count[r];
  if (blockIdx.x * DP_THREADS >= n) return;   // the whole workgroup leaves together: no barrier is pending
  if (threadIdx.x < DP_SH_ENTRIES) s_sh[threadIdx.x] = sh_table[threadIdx.x];
  __syncthreads();
  const uint32_t *list = q_list + r * q_cap;
  for (uint32_t j = blockIdx.x * DP_THREADS + threadIdx.x; j < n; j += gridDim.x * DP_THREADS) {
    const uint32_t i = list[j];
    dp_store<BGRA>(out, i, dp_pixel<ExactFn>(a, s_sh, i, dp_load<HALF_IN>(hdr, i)));
  }
}
