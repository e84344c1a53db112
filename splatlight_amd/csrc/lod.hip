// Screen-space LOD (runtime/wgsl/lod.wgsl:61-91, 141-224): the cut through the asset's tree the frame draws, and
// the lit colour of its nodes.
//
// The reference runs lod_select once per tree depth top-down and lod_colour once per depth bottom-up (~25 levels
// each, one dispatch per level). Here the cut is two launches and the colours one:
//  * sl_lod_open: one lane per internal node, whether it opens (its on-screen size, lod.wgsl:188-215, exceeds the
//    threshold, or it belongs to the emitters' part): the per-node half of lod_select, independent of the others;
//  * sl_lod_flags: one lane per node walks its ancestors up to the nearest part root. A node is a candidate when it
//    is a part root or every ancestor up to that root opens; a candidate draws when it is a leaf or does not open
//    (else it is OPEN); a non-candidate below a part root lies under a drawn node (UNDER), above every part root
//    it is NONE. These are exactly the flags the level-by-level passes leave;
//  * sl_lod_colour: one lane per UNDER leaf climbs its drawn subtree: at each parent it counts itself in with an
//    atomic, and the last child to arrive computes the parent's colour (lod.wgsl:72-91: the mass-weighted mean of its
//    children's shown colours, in child order) and carries on up while the parent is UNDER. Every drawn or UNDER
//    node's colour is computed once, from final child values, with the same operations in the same order as the
//    per-level passes, so the colours are the same.
#include "sl_device.h"

// What the LOD kernels read: the tree (scene-resident) and this frame's selection parameters.
struct LodDev {
  const float4 *__restrict__ na;     // internal node k: mean.xyz, radius
  const float4 *__restrict__ nb;     // cinv_diag.xyz, kappa
  const float4 *__restrict__ nc;     // cinv_off.xyz, part (bits)
  const uint32_t *__restrict__ up;   // parent of each node (0xFFFFFFFF: none)
  const uint32_t *__restrict__ part; // each node's part
  const uint2 *__restrict__ child;   // internal node k: [child_lo, child_hi) into kids
  const uint32_t *__restrict__ kids;
  const float *__restrict__ mass;    // node_mass of each node (lod.wgsl:61-65)
  const float4 *__restrict__ parts;  // the scene's part table (8 float4 per part)
  uint32_t N, G;
  float view[16];
  float threshold, focal, near_;
  uint32_t emit_part, open_parts, size_rule;
};

#define LOD_NO_PARENT 0xFFFFFFFFu
#define LOD_MAX_DEPTH 64u

__device__ __forceinline__ bool lod_is_part_root(const LodDev &L, uint32_t n) {
  const uint32_t p = L.part[n];
  return p < 32u && __float_as_uint(L.parts[p * 8u + 7u].x) == n;
}

// lod.wgsl:141-163: the node's covariance projected to the screen at its world mean w; S[col][row]
__device__ __forceinline__ void lod_screen_cov(const LodDev &L, float4 b, float4 c, uint32_t P, const float w[3],
                                               float S[2][2]) {
#pragma clang fp contract(off)
  const float c0[3] = {b.x, c.x, c.y}, c1[3] = {c.x, b.y, c.z}, c2[3] = {c.y, c.z, b.z};
  float r[3][3];
  r[0][0] = c1[1] * c2[2] - c1[2] * c2[1]; r[0][1] = c1[2] * c2[0] - c1[0] * c2[2]; r[0][2] = c1[0] * c2[1] - c1[1] * c2[0];
  r[1][0] = c2[1] * c0[2] - c2[2] * c0[1]; r[1][1] = c2[2] * c0[0] - c2[0] * c0[2]; r[1][2] = c2[0] * c0[1] - c2[1] * c0[0];
  r[2][0] = c0[1] * c1[2] - c0[2] * c1[1]; r[2][1] = c0[2] * c1[0] - c0[0] * c1[2]; r[2][2] = c0[0] * c1[1] - c0[1] * c1[0];
  const float inv = 1.0f / ((c0[0] * r[0][0] + c0[1] * r[0][1]) + c0[2] * r[0][2]);
  const float4 x0 = L.parts[P * 8u], x1 = L.parts[P * 8u + 1u], x2 = L.parts[P * 8u + 2u];
  const float Rp[3][3] = {{x0.x, x0.y, x0.z}, {x1.x, x1.y, x1.z}, {x2.x, x2.y, x2.z}};   // [row][col]
  float C[3][3], M[3][3], T[3][3], Cv[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) C[i][j] = r[i][j] * inv;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      M[i][j] = (L.view[i] * Rp[0][j] + L.view[4 + i] * Rp[1][j]) + L.view[8 + i] * Rp[2][j];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) T[i][j] = (M[i][0] * C[0][j] + M[i][1] * C[1][j]) + M[i][2] * C[2][j];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Cv[i][j] = (T[i][0] * M[j][0] + T[i][1] * M[j][1]) + T[i][2] * M[j][2];
  float pv[3];
#pragma unroll
  for (int q = 0; q < 3; q++)
    pv[q] = ((L.view[q] * w[0] + L.view[4 + q] * w[1]) + L.view[8 + q] * w[2]) + L.view[12 + q] * 1.0f;
  const float z = -pv[2], f = L.focal;
  const float j0[3] = {f / z, 0.0f, f * pv[0] / (z * z)};
  const float j1[3] = {0.0f, f / z, f * pv[1] / (z * z)};
  float a0[3], a1[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    a0[i] = (Cv[i][0] * j0[0] + Cv[i][1] * j0[1]) + Cv[i][2] * j0[2];
    a1[i] = (Cv[i][0] * j1[0] + Cv[i][1] * j1[1]) + Cv[i][2] * j1[2];
  }
  S[0][0] = (j0[0] * a0[0] + j0[1] * a0[1]) + j0[2] * a0[2];
  S[0][1] = (j1[0] * a0[0] + j1[1] * a0[1]) + j1[2] * a0[2];
  S[1][0] = (j0[0] * a1[0] + j0[1] * a1[1]) + j0[2] * a1[2];
  S[1][1] = (j1[0] * a1[0] + j1[1] * a1[1]) + j1[2] * a1[2];
}

// Whether each internal node opens (lod.wgsl:188-215); open[k] for node G + k. Nodes above every part root
// (part >= 32) are never candidates, so their bit is left 0.
extern "C" __global__ void __launch_bounds__(256) sl_lod_open(LodDev L, uint8_t *__restrict__ open) {
#pragma clang fp contract(off)
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= L.N - L.G) return;
  const float4 a = L.na[k], c = L.nc[k];
  const uint32_t P = __float_as_uint(c.w);
  bool op = P == L.emit_part || (P < 32u && ((L.open_parts >> P) & 1u) != 0u);
  if (!op && P < 32u) {
    const float4 x0 = L.parts[P * 8u], x1 = L.parts[P * 8u + 1u], x2 = L.parts[P * 8u + 2u];
    const float w[3] = {((x0.x * a.x + x0.y * a.y) + x0.z * a.z) + x0.w * 1.0f,
                        ((x1.x * a.x + x1.y * a.y) + x1.z * a.z) + x1.w * 1.0f,
                        ((x2.x * a.x + x2.y * a.y) + x2.z * a.z) + x2.w * 1.0f};
    const float z = -(((L.view[2] * w[0] + L.view[6] * w[1]) + L.view[10] * w[2]) + L.view[14] * 1.0f);
    const float R = a.w;
    float size = (float)(2.0 / 3.0) * L.focal * R / fmaxf(z - R, L.near_);
    if ((L.size_rule & 1u) == 1u && z - R > L.near_) {
      float S[2][2];
      lod_screen_cov(L, L.nb[k], c, P, w, S);
      const float tr = S[0][0] + S[1][1];
      const float dt = fmaxf(S[0][0] * S[1][1] - S[1][0] * S[0][1], 0.0f);
      const float disc = sqrtf(fmaxf(0.25f * (S[0][0] - S[1][1]) * (S[0][0] - S[1][1]) + S[0][1] * S[0][1], 0.0f));
      const float l1 = fmaxf(0.5f * tr + disc, 1e-12f);
      size *= sqrtf(sqrtf(fmaxf(dt / l1, 0.0f) / l1));
    }
    op = size > L.threshold;
  }
  open[k] = op ? 1u : 0u;
}

// lod_select's flags (NONE, OPEN, DRAW, UNDER) of every node, from the open bits by an ancestor walk. The walk is
// bounded by LOD_MAX_DEPTH steps; the host checked that every child sits one level below its parent, so a walk
// never cycles and ends within the tree's depth.
extern "C" __global__ void __launch_bounds__(256) sl_lod_flags(LodDev L, const uint8_t *__restrict__ open,
                                                             uint32_t *__restrict__ flag) {
  const uint32_t n = blockIdx.x * 256u + threadIdx.x;
  if (n >= L.N) return;
  bool cand = lod_is_part_root(L, n), under = false;
  if (!cand) {
    bool all_open = true, found = false;
    uint32_t a = L.up[n];
    for (uint32_t step = 0; step < LOD_MAX_DEPTH && a != LOD_NO_PARENT; step++) {
      all_open = all_open && open[a - L.G] != 0u;
      if (lod_is_part_root(L, a)) {
        found = true;
        break;
      }
      a = L.up[a];
    }
    cand = found && all_open;
    under = found && !all_open;
  }
  uint32_t fl = SL_LOD_NONE;
  if (cand) fl = n < L.G ? SL_LOD_DRAW : (open[n - L.G] ? SL_LOD_OPEN : SL_LOD_DRAW);
  else if (under) fl = SL_LOD_UNDER;
  flag[n] = fl;
}

// The colour of every drawn node and every node under one (lod.wgsl:72-91), by the climb described above. cnt:
// one zeroed word per internal node; node_colour: (rgb, 1) per internal node; lo_b null: no bounce.
extern "C" __global__ void __launch_bounds__(256)
sl_lod_colour(LodDev L, const uint32_t *__restrict__ flag, const float4 *__restrict__ colour,
              const float4 *__restrict__ lo_b, float gain, uint32_t *__restrict__ cnt, float4 *node_colour) {
#pragma clang fp contract(off)
  const uint32_t g = blockIdx.x * 256u + threadIdx.x;
  if (g >= L.G || flag[g] != SL_LOD_UNDER) return;
  uint32_t n = g;
  for (uint32_t step = 0; step < LOD_MAX_DEPTH; step++) {
    const uint32_t p = L.up[n];   // an UNDER node lies below a drawn node: it has a parent
    if (p == LOD_NO_PARENT) return;
    const uint2 ch = L.child[p - L.G];
    __threadfence();              // this lane's colour of n (when n is a node) before the count that publishes it
    if (atomicAdd(&cnt[p - L.G], 1u) + 1u != ch.y - ch.x) return;   // not the last child in
    __threadfence();              // the other children's colours, published before their counts
    float cx = 0.0f, cy = 0.0f, cz = 0.0f, w = 0.0f;
    for (uint32_t k = ch.x; k < ch.y; k++) {
      const uint32_t c = L.kids[k];
      const float m = L.mass[c];
      float sx, sy, sz;
      if (c < L.G) {
        const float4 cv = colour[c], lb = lo_b ? lo_b[3u * c] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        sx = cv.x + gain * lb.x;
        sy = cv.y + gain * lb.y;
        sz = cv.z + gain * lb.z;
      } else {
        const float4 nc = node_colour[c - L.G];
        sx = nc.x;
        sy = nc.y;
        sz = nc.z;
      }
      cx = cx + m * sx;
      cy = cy + m * sy;
      cz = cz + m * sz;
      w = w + m;
    }
    const float d = fmaxf(w, 1e-30f);
    node_colour[p - L.G] = make_float4(cx / d, cy / d, cz / d, 1.0f);
    if (flag[p] != SL_LOD_UNDER) return;   // p is the drawn node: its subtree is done
    n = p;
  }
}

// Display fill (runtime/wgsl/fill.wgsl:15-33): a fill target's colour and bounce toward the eye become the weighted
// sum of its sources' (targets and sources are disjoint, so the sources are read unfilled). fill: offsets [G + 1],
// then (source, f32 weight bits) pairs; lo_b null: no bounce.
extern "C" __global__ void __launch_bounds__(256)
sl_display_fill(const uint32_t *__restrict__ fill, uint32_t G, float4 *__restrict__ colour, float4 *__restrict__ lo_b) {
#pragma clang fp contract(off)
  const uint32_t g = blockIdx.x * 256u + threadIdx.x;
  if (g >= G) return;
  const uint32_t lo = fill[g], hi = fill[g + 1u];
  if (hi <= lo) return;
  float cx = 0.0f, cy = 0.0f, cz = 0.0f, bx = 0.0f, by = 0.0f, bz = 0.0f;
  for (uint32_t k = lo; k < hi; k++) {
    const uint32_t j = fill[G + 1u + 2u * k];
    const float w = __uint_as_float(fill[G + 2u + 2u * k]);
    const float4 c = colour[j], b = lo_b ? lo_b[3u * j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    cx = cx + w * c.x;
    cy = cy + w * c.y;
    cz = cz + w * c.z;
    bx = bx + w * b.x;
    by = by + w * b.y;
    bz = bz + w * b.z;
  }
  colour[g] = make_float4(cx, cy, cz, colour[g].w);
  if (lo_b) lo_b[3u * g] = make_float4(bx, by, bz, 0.0f);
}
