/* CPU oracle, screen-space LOD and display fill (TEST INFRASTRUCTURE ONLY): runtime/wgsl/lod.wgsl:166-224
 * (lod_select, screen_cov, is_part_root), :72-91 (lod_colour, node_mass, shown_colour) and runtime/wgsl/fill.wgsl:15-33
 * (display_fill), float32 with no contraction, in the shaders' own per-depth passes: lod_select top-down one depth
 * at a time (part roots are candidates; a candidate opens when its on-screen size exceeds the threshold, its children
 * become candidates; a drawn node's subtree is marked UNDER), lod_colour bottom-up over the drawn nodes and their
 * subtrees. Matrix and vector products are summed in index order.
 *
 * Pinned by tests/test_lod_oracle.py: every leaf of a part covered exactly once (drawn, or under exactly one drawn
 * node), the cut monotone in the threshold, threshold 0 drawing every leaf and a huge one drawing the part roots, the
 * emitters' part always opened to its leaves, screen_cov against the float64 projection J V R C R^T V^T J^T, the node
 * colour against the mass-weighted mean written out by hand on small trees, and the fill against a dense matrix
 * product.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define LOD_NONE 0u
#define LOD_OPEN 1u
#define LOD_DRAW 2u
#define LOD_CAND 3u
#define LOD_UNDER 4u

static float nf(const void *nodes, uint32_t n, int k) {
  float v;
  memcpy(&v, (const char *)nodes + 96 * (size_t)n + 4 * k, 4);
  return v;
}
static uint32_t nu(const void *nodes, uint32_t n, int k) {
  uint32_t v;
  memcpy(&v, (const char *)nodes + 96 * (size_t)n + 4 * k, 4);
  return v;
}
static float pf(const void *parts, uint32_t p, int k) {
  float v;
  memcpy(&v, (const char *)parts + 128 * (size_t)p + 4 * k, 4);
  return v;
}
static uint32_t pu(const void *parts, uint32_t p, int k) {
  uint32_t v;
  memcpy(&v, (const char *)parts + 128 * (size_t)p + 4 * k, 4);
  return v;
}

/* lod.wgsl:168-171 */
static int is_part_root(const void *nodes, const void *parts, uint32_t n) {
  const uint32_t p = nu(nodes, n, 15);
  return p < 32u && pu(parts, p, 28) == n;
}

/* lod.wgsl:141-163: node n's covariance projected to the screen (px^2) at its world mean w; S[col][row] */
static void screen_cov(const float *view, float focal, const void *nodes, const void *parts, uint32_t n, uint32_t P,
                       const float w[3], float S[2][2]) {
  const float ax = nf(nodes, n, 4), ay = nf(nodes, n, 5), az = nf(nodes, n, 6);
  const float ox = nf(nodes, n, 8), oy = nf(nodes, n, 9), oz = nf(nodes, n, 10);
  const float c0[3] = {ax, ox, oy}, c1[3] = {ox, ay, oz}, c2[3] = {oy, oz, az};
  float r[3][3];
  r[0][0] = c1[1] * c2[2] - c1[2] * c2[1]; r[0][1] = c1[2] * c2[0] - c1[0] * c2[2]; r[0][2] = c1[0] * c2[1] - c1[1] * c2[0];
  r[1][0] = c2[1] * c0[2] - c2[2] * c0[1]; r[1][1] = c2[2] * c0[0] - c2[0] * c0[2]; r[1][2] = c2[0] * c0[1] - c2[1] * c0[0];
  r[2][0] = c0[1] * c1[2] - c0[2] * c1[1]; r[2][1] = c0[2] * c1[0] - c0[0] * c1[2]; r[2][2] = c0[0] * c1[1] - c0[1] * c1[0];
  const float inv = 1.0f / ((c0[0] * r[0][0] + c0[1] * r[0][1]) + c0[2] * r[0][2]);
  float C[3][3], Rp[3][3], V[3][3], M[3][3], T[3][3], Cv[3][3];   /* [row][col] */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      C[i][j] = r[i][j] * inv;
      Rp[i][j] = pf(parts, P, 4 * i + j);
      V[i][j] = view[4 * j + i];
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) M[i][j] = (V[i][0] * Rp[0][j] + V[i][1] * Rp[1][j]) + V[i][2] * Rp[2][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) T[i][j] = (M[i][0] * C[0][j] + M[i][1] * C[1][j]) + M[i][2] * C[2][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Cv[i][j] = (T[i][0] * M[j][0] + T[i][1] * M[j][1]) + T[i][2] * M[j][2];
  float pv[3];
  for (int r2 = 0; r2 < 3; r2++) pv[r2] = ((view[r2] * w[0] + view[4 + r2] * w[1]) + view[8 + r2] * w[2]) + view[12 + r2] * 1.0f;
  const float z = -pv[2];
  const float j0[3] = {focal / z, 0.0f, focal * pv[0] / (z * z)};
  const float j1[3] = {0.0f, focal / z, focal * pv[1] / (z * z)};
  float a0[3], a1[3];
  for (int i = 0; i < 3; i++) {
    a0[i] = (Cv[i][0] * j0[0] + Cv[i][1] * j0[1]) + Cv[i][2] * j0[2];
    a1[i] = (Cv[i][0] * j1[0] + Cv[i][1] * j1[1]) + Cv[i][2] * j1[2];
  }
  S[0][0] = (j0[0] * a0[0] + j0[1] * a0[1]) + j0[2] * a0[2];
  S[0][1] = (j1[0] * a0[0] + j1[1] * a0[1]) + j1[2] * a0[2];
  S[1][0] = (j0[0] * a1[0] + j0[1] * a1[1]) + j0[2] * a1[2];
  S[1][1] = (j1[0] * a1[0] + j1[1] * a1[1]) + j1[2] * a1[2];
}

/* lod.wgsl:191-208: internal node n's on-screen size (px) */
float or_lod_node_size(const or_lod_params *L, const void *nodes, const void *parts, uint32_t n) {
  const uint32_t P = nu(nodes, n, 15);
  const float h[3] = {nf(nodes, n, 0), nf(nodes, n, 1), nf(nodes, n, 2)};
  float w[3];
  for (int r = 0; r < 3; r++)
    w[r] = ((pf(parts, P, 4 * r) * h[0] + pf(parts, P, 4 * r + 1) * h[1]) + pf(parts, P, 4 * r + 2) * h[2]) +
           pf(parts, P, 4 * r + 3) * 1.0f;
  const float *v = L->view;
  const float z = -(((v[2] * w[0] + v[6] * w[1]) + v[10] * w[2]) + v[14] * 1.0f);
  const float R = nf(nodes, n, 3);
  float size = (float)(2.0 / 3.0) * L->focal * R / fmaxf(z - R, L->near_);
  if ((L->size_rule & 1u) == 1u && z - R > L->near_) {
    float S[2][2];
    screen_cov(v, L->focal, nodes, parts, n, P, w, S);
    const float tr = S[0][0] + S[1][1];
    const float dt = fmaxf(S[0][0] * S[1][1] - S[1][0] * S[0][1], 0.0f);
    const float disc = sqrtf(fmaxf(0.25f * (S[0][0] - S[1][1]) * (S[0][0] - S[1][1]) + S[0][1] * S[0][1], 0.0f));
    const float l1 = fmaxf(0.5f * tr + disc, 1e-12f);
    size *= sqrtf(sqrtf(fmaxf(dt / l1, 0.0f) / l1));
  }
  return size;
}

/* lod.wgsl:188-215: does internal node n open (its children become candidates)? */
static int node_opens(const or_lod_params *L, const void *nodes, const void *parts, uint32_t n) {
  const uint32_t P = nu(nodes, n, 15);
  if (P == L->emit_part || (P < 32u && ((L->open_parts >> P) & 1u) != 0u)) return 1;
  return or_lod_node_size(L, nodes, parts, n) > L->threshold;
}

void or_lod_select(const or_lod_params *L, const void *nodes, const uint32_t *kids, const void *parts, uint32_t N,
                   uint32_t G, uint32_t *flag) {
  uint32_t maxd = 0;
  for (uint32_t n = 0; n < N; n++) {
    const uint32_t d = nu(nodes, n, 21);
    if (d > maxd) maxd = d;
    flag[n] = LOD_NONE;
  }
  /* nodes by depth, as lod.js levels: one pass per depth, top-down */
  uint32_t *cnt = (uint32_t *)calloc(maxd + 2, 4), *order = (uint32_t *)malloc(4 * (size_t)(N ? N : 1));
  for (uint32_t n = 0; n < N; n++) cnt[nu(nodes, n, 21) + 1]++;
  for (uint32_t d = 0; d <= maxd; d++) cnt[d + 1] += cnt[d];
  uint32_t *pos = (uint32_t *)malloc(4 * (size_t)(maxd + 1));
  memcpy(pos, cnt, 4 * (size_t)(maxd + 1));
  for (uint32_t n = 0; n < N; n++) order[pos[nu(nodes, n, 21)]++] = n;
  for (uint32_t d = 0; d <= maxd; d++) {
    for (uint32_t q = cnt[d]; q < cnt[d + 1]; q++) {
      const uint32_t n = order[q];
      const uint32_t f = flag[n];
      const int cand = is_part_root(nodes, parts, n) || f == LOD_CAND;
      const uint32_t lo = nu(nodes, n, 19), hi = nu(nodes, n, 20);
      if (!cand) {
        const int under = f == LOD_UNDER;
        if (!under) flag[n] = LOD_NONE;
        if (n >= G)
          for (uint32_t k = lo; k < hi; k++) flag[kids[k]] = under ? LOD_UNDER : LOD_NONE;
        continue;
      }
      if (n < G) {
        flag[n] = LOD_DRAW;
        continue;
      }
      if (!node_opens(L, nodes, parts, n)) {
        flag[n] = LOD_DRAW;
        for (uint32_t k = lo; k < hi; k++) flag[kids[k]] = LOD_UNDER;
        continue;
      }
      flag[n] = LOD_OPEN;
      for (uint32_t k = lo; k < hi; k++) flag[kids[k]] = LOD_CAND;
    }
  }
  free(cnt);
  free(order);
  free(pos);
}

/* lod.wgsl:61-65 */
float or_node_mass(const void *nodes, uint32_t n) {
  const float ax = nf(nodes, n, 4), ay = nf(nodes, n, 5), az = nf(nodes, n, 6);
  const float ox = nf(nodes, n, 8), oy = nf(nodes, n, 9), oz = nf(nodes, n, 10);
  const float det_inv = ax * (ay * az - oz * oz) - ox * (ox * az - oz * oy) + oy * (ox * oz - ay * oy);
  return nf(nodes, n, 7) / sqrtf(fmaxf(det_inv, 1e-30f));
}

void or_lod_colour(const void *nodes, const uint32_t *kids, uint32_t N, uint32_t G, const uint32_t *flag,
                   const float *colour, const float *lo_b, float gain, float *node_colour) {
  uint32_t maxd = 0;
  for (uint32_t n = G; n < N; n++) {
    const uint32_t d = nu(nodes, n, 21);
    if (d > maxd) maxd = d;
  }
  for (int64_t d = maxd; d >= 0; d--) {   /* deepest first: children before parents */
    for (uint32_t n = G; n < N; n++) {
      if (nu(nodes, n, 21) != (uint32_t)d) continue;
      if (flag[n] != LOD_DRAW && flag[n] != LOD_UNDER) continue;
      float c[3] = {0.0f, 0.0f, 0.0f}, w = 0.0f;
      for (uint32_t k = nu(nodes, n, 19); k < nu(nodes, n, 20); k++) {
        const uint32_t ch = kids[k];
        const float m = or_node_mass(nodes, ch);
        float s[3];
        for (int j = 0; j < 3; j++)
          s[j] = ch < G ? colour[4 * (size_t)ch + j] + gain * lo_b[12 * (size_t)ch + j]
                        : node_colour[4 * (size_t)(ch - G) + j];
        for (int j = 0; j < 3; j++) c[j] = c[j] + m * s[j];
        w = w + m;
      }
      for (int j = 0; j < 3; j++) node_colour[4 * (size_t)(n - G) + j] = c[j] / fmaxf(w, 1e-30f);
      node_colour[4 * (size_t)(n - G) + 3] = 1.0f;
    }
  }
}

void or_display_fill(const uint32_t *fill, uint32_t G, float *colour, float *lo_b) {
  /* targets and sources are disjoint (fill.wgsl:1-6), so every target reads its sources' unfilled values */
  for (uint32_t g = 0; g < G; g++) {
    const uint32_t lo = fill[g], hi = fill[g + 1];
    if (hi <= lo) continue;
    float c[3] = {0.0f, 0.0f, 0.0f}, b[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t k = lo; k < hi; k++) {
      const uint32_t j = fill[G + 1 + 2 * k];
      float w;
      memcpy(&w, &fill[G + 1 + 2 * k + 1], 4);
      for (int q = 0; q < 3; q++) {
        c[q] = c[q] + w * colour[4 * (size_t)j + q];
        b[q] = b[q] + w * lo_b[12 * (size_t)j + q];
      }
    }
    for (int q = 0; q < 3; q++) {
      colour[4 * (size_t)g + q] = c[q];
      lo_b[12 * (size_t)g + q] = b[q];
    }
    lo_b[12 * (size_t)g + 3] = 0.0f;
  }
}
