// Device-side declarations shared by the splat-path kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#define SL_INVALID_KEY 0xffffffffu
#define SL_MAX_PARTS 32u
#define SL_PART_HIDDEN 4u
#define SL_EMITTER_MATERIAL 254u
#define SL_WAVE 64

// Per-frame parameters decoded from the 1200 B RenderFrame (shaders/common.wgsl:44-77) on the host and passed
// to the kernels by value.
struct FrameDev {
  float view[16];            // column-major
  float palette[8][4];       // palette slots 0-7 (slot 3: bounce gain, emitter count, shutter, blur coverage)
  uint32_t part_material[SL_MAX_PARTS];
  float part_keep[SL_MAX_PARTS];   // 1 - clamp(part_fade)
  float vw, vh, fx, fy;
  float near_, far_;
  float log_far_near;        // (float)log((double)(far / near)): the depth key's denominator
  float albedo_scale, lowpass, gain;
  uint32_t count;            // scene Gaussians
  uint32_t n_emit;           // emitter slots after count + 1
  uint32_t nslots;           // count + 1 + n_emit
  uint32_t view_mode;
  uint32_t glass_mask;       // bit m: material m's mat_view flags bit 0 (coverage set: SPLAT_GLASS)
  uint32_t width, height, tiles_x, tiles_y;
  float shutter;             // motion blur: shutter fraction (palette[3].z); moving parts blur when > 0
  uint32_t blur_cover;       // the blurred splats' coverage cap (palette[3].w > 0.5)
  float lod_threshold;       // screen-space LOD: the node footprint (px) to open at (frame.lod.x); 0 = off
  uint32_t n_nodes;          // its internal nodes drawn after the emitters (frame.lod.z), 0 when off
  uint32_t tile_shift;       // the composite tile's side is 1 << tile_shift pixels (4: 16, 5: 32)
};

// Gaussian attribute storage (SL_STORAGE_*). F32: four float4 streams g0 (mu, kappa), g1 quat, g2 (scale,
// part_surface bits), g3 (albedo, flags bits). F16 / BF16: g0 unchanged (means and opacity stay f32), g1 = uint4
// {quat.x | quat.y << 16, quat.z | quat.w << 16, albedo.r | albedo.g << 16, albedo.b}, g2 = uint4 {scale.x |
// scale.y << 16, scale.z, part_surface, flags}, halves rounded to nearest even on the host; g3 unused. Decoding is
// exact, so every kernel computes in f32 on the stored values.
#define SL_ST_F32 0u
#define SL_ST_BF16 2u
__device__ __forceinline__ float sl_h2f(uint32_t b, uint32_t st) {
  return st == SL_ST_BF16 ? __uint_as_float((b & 0xffffu) << 16)
                          : __half2float(__ushort_as_half((unsigned short)(b & 0xffffu)));
}
__device__ __forceinline__ float4 sl_quat(const float4 *g1, uint32_t st, uint32_t i) {
  if (st == SL_ST_F32) return g1[i];
  const uint4 h = reinterpret_cast<const uint4 *>(g1)[i];
  return make_float4(sl_h2f(h.x, st), sl_h2f(h.x >> 16, st), sl_h2f(h.y, st), sl_h2f(h.y >> 16, st));
}
__device__ __forceinline__ float4 sl_scale(const float4 *g2, uint32_t st, uint32_t i) {   // .w: part_surface bits
  if (st == SL_ST_F32) return g2[i];
  const uint4 h = reinterpret_cast<const uint4 *>(g2)[i];
  return make_float4(sl_h2f(h.x, st), sl_h2f(h.x >> 16, st), sl_h2f(h.y, st), __uint_as_float(h.z));
}
__device__ __forceinline__ float4 sl_albedo(const float4 *g1, const float4 *g2, const float4 *g3, uint32_t st,
                                            uint32_t i) {                                // .w: flags bits
  if (st == SL_ST_F32) return g3[i];
  const uint4 h1 = reinterpret_cast<const uint4 *>(g1)[i];
  const uint32_t fl = reinterpret_cast<const uint4 *>(g2)[i].w;
  return make_float4(sl_h2f(h1.z, st), sl_h2f(h1.z >> 16, st), sl_h2f(h1.w, st), __uint_as_float(fl));
}

// One view of a batch: its frame, its scene's streams and its target. Views lie back to back in the global slot
// space (each view's slots start at a multiple of 256, so a 256-slot block belongs to one view) and in the global
// tile space; a single frame is a batch of one view.
struct ViewDev {
  FrameDev f;
  const float4 *g0, *g1, *g2, *g3;          // the scene's SoA Gaussian streams
  const float4 *e0, *e1, *e2, *e3;          // its emitter streams
  const float4 *parts, *colour, *lo_b;
  void *out;                                // width * height pixels of out_format
  uint32_t out_format;
  uint32_t slot_base, tile_base;
  uint32_t storage;                         // SL_STORAGE_* of the scene's attribute streams
  // screen-space LOD (lod_flag non-null): slots from node_slot0 on are the tree's internal nodes k < n_nodes, their
  // records in the n streams (the scene's storage), their lit colour node_colour[k]; only nodes and leaves whose
  // lod_flag is SL_LOD_DRAW are drawn (node count + k is node k's flag)
  const uint32_t *lod_flag;
  const float4 *n0, *n1, *n2, *n3;
  const float4 *node_colour;
  uint32_t node_slot0, n_nodes;
};

#define SL_LOD_NONE 0u
#define SL_LOD_OPEN 1u
#define SL_LOD_DRAW 2u
#define SL_LOD_UNDER 4u

// Per-view results of a batch.
struct ViewCounters {
  uint32_t visible;        // valid depth keys of the view
  uint32_t pairs;          // its (splat, tile) pairs
  uint32_t color_max_bits; // max |rgb| over its visible splats (float bits)
  uint32_t pad;
};

// the view owning global index x: the last v with base[v] <= x (base ascending, base[0] = 0)
__device__ __forceinline__ uint32_t view_of(const uint32_t *__restrict__ base, uint32_t n_views, uint32_t x) {
  uint32_t lo = 0, hi = n_views - 1u;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1u) >> 1;
    if (base[mid] <= x) lo = mid; else hi = mid - 1u;
  }
  return lo;
}

// Splat record, 48 B: the layout project.wgsl writes (shaders/common.wgsl:29-35).
struct __align__(16) Splat {
  float4 center_axis0;   // centre.xy (NDC), axis0.xy (NDC)
  float4 axis1_r_tau;    // axis1.xy (NDC), r, tau0
  uint4 color_depth;     // rgb bits, depth bits & ~3
};

// Per-frame device counters, one 256 B block.
struct FrameCounters {
  uint32_t visible;        // valid depth keys
  uint32_t pairs;          // (splat, tile) pairs emitted
  uint32_t overflow;       // pairs that did not fit the pair buffers
  uint32_t color_max_bits; // max |rgb| over visible splats (float bits, non-negative)
  uint32_t pad[60];
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t l = __lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}
