/* CPU oracle for the splat frame path: TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's WGSL/JS for the path, used by tests/, the smoke check and the
 * CPU-baseline leg of bench.py as the checker. The shipped library never links, loads or calls it.
 * Each function cites the reference file:line it follows. Each is pinned by the CPU tests in tests/test_oracle.py,
 * tests/test_lighting_oracle.py and tests/test_bounce_oracle.py: the reference's own fixtures where it ships them
 * (the hill_lite tree), closed forms (a lone splat's coverage, a sky seen through nothing), invariants (sums of
 * leaves up a tree, order independence of integer sums) and brute force on small inputs, chosen so that a dropped
 * term or a wrong index fails one of them.
 */
#ifndef SL_ORACLE_H
#define SL_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* threads of the parallel loops (OMP_NUM_THREADS, else the host cores) */
uint32_t or_threads(void);

/* The screen-space LOD's cut as project draws it: flag[N] (lod_select's flags, LOD_DRAW = 2), the N - G internal
 * nodes' 64 B Gaussian records (lod.js nodeGaussians) and their lit colours (vec4), in splat slots count + 1 +
 * n_emit + k. */
typedef struct or_lod_view {
  const uint32_t *flag;
  const void *records;
  const float *colour;
  uint32_t n_nodes;
} or_lod_view;

/* lod_select's uniforms (lod.wgsl LodU): the view matrix (column-major), threshold px, focal px, near plane, the
 * emitters' part (always opened; 0xFFFFFFFF none), parts always opened (bit per part), size_rule bit 0 the
 * projected-area measure */
typedef struct or_lod_params {
  float view[16];
  float threshold, focal, near_;
  uint32_t emit_part, open_parts, size_rule;
} or_lod_params;

/* 16-bit log-depth key (shaders/project.wgsl:366-367) with log fixed as the correctly rounded float log. */
uint32_t or_depth_key(float depth, float near_, float far_);

/* project (shaders/project.wgsl:84-370), float32 evaluated in the WGSL expression order, no contraction.
 * frame: 1200 B RenderFrame; gaussians/emitters: 64 B records; parts: 128 B records; colour: vec4 per Gaussian
 * (may be NULL: zero); lo_b: 3 vec4 per Gaussian (may be NULL: zero). Writes splats[12*nslots] (the 48 B Splat
 * words, untouched where the key is invalid) and keys[nslots]. nslots = count + 1 + n_emit. Returns visible. */
uint32_t or_project_f32(const void *frame, const void *gaussians, const void *parts, const float *colour,
                        const float *lo_b, const void *emitters, uint32_t n_emit, uint32_t *splats,
                        uint32_t *keys);

/* or_project_f32 with the LOD's cut (lod may be NULL); nslots = count + 1 + n_emit + lod->n_nodes. */
uint32_t or_project_lod_f32(const void *frame, const void *gaussians, const void *parts, const float *colour,
                            const float *lo_b, const void *emitters, uint32_t n_emit, const or_lod_view *lod,
                            uint32_t *splats, uint32_t *keys);

/* The same in float64: out[12*slot] = centre.xy, axis0.xy, axis1.xy, r, tau0, rgb, depth; valid[slot]. */
void or_project_f64(const void *frame, const void *gaussians, const void *parts, const float *colour,
                    const float *lo_b, const void *emitters, uint32_t n_emit, double *out, uint8_t *valid);

/* Counting-sort artefacts of the 16-bit keys: histogram (shaders/project.wgsl:369), exclusive offsets and the
 * visible count (shaders/scan.wgsl:16-43), and the stable order (shaders/radix.wgsl:1-139 == a stable argsort,
 * INVALID last). hist/offsets: 65536 entries each (may be NULL); order: n entries. Returns visible. */
uint32_t or_sort_keys(const uint32_t *keys, uint32_t n, uint32_t *hist, uint32_t *offsets, uint32_t *order);

/* Tile overlap of one splat's extent box (shaders/project.wgsl:314) in 16x16 (tile_px) tiles: returns the
 * number of tiles and writes the inclusive tile box (tx0, ty0, tx1, ty1). */
uint32_t or_tile_box(const uint32_t *splat, uint32_t W, uint32_t H, uint32_t tile_px, uint32_t box[4]);

/* Per-tile lists in draw order: tile_start[ntiles + 1] (CSR), vals[P] = Gaussian slot index ascending in
 * (key, index) within each tile. vals may be NULL to count only. Returns P. */
uint64_t or_tile_lists(const uint32_t *splats, const uint32_t *order, uint32_t visible, uint32_t W, uint32_t H,
                       uint32_t tile_px, uint64_t *tile_start, uint32_t *vals);

/* Composite (shaders/splat.wgsl:22-78, src/renderer.js:133-142): back-to-front premultiplied "over" of the
 * quads in `order`, rows top-down. mode 0: float32 accumulation; 2: float16 stored after every blend (the
 * rgba16float target, src/renderer.js:13). view_mode from the frame (0 flat, 1 depth, 2 overdraw, 3 lit).
 * out: W*H*4 floats. */
void or_composite(const void *frame, const uint32_t *splats, const uint32_t *order, uint32_t visible,
                  uint32_t W, uint32_t H, int mode, float *out);
void or_composite_rows(const void *frame, const uint32_t *splats, const uint32_t *order, uint32_t visible,
                       uint32_t W, uint32_t H, int mode, uint32_t row_step, float *out);

/* The same blend in float64 over or_project_f64's splats (out: W*H*4 doubles). */
void or_composite_f64(const void *frame, const double *splats, const uint32_t *order, uint32_t visible,
                      uint32_t W, uint32_t H, double *out);

/* ---- relighting (lighting.c): float32 restatements of runtime/wgsl for the default M39 matte path ---------- */

void or_exp_map_n(const float *x, float *y, uint32_t n);                        /* the map splats' e^x */
float or_erfc_pos(float z);                                                      /* common.wgsl:113-118 */
float or_ein(float x);                                                           /* common.wgsl:226-239 */
float or_footprint(const void *node, float detC, float wx, float wy, float wz);  /* common.wgsl:242-246 */
float or_vgrid_lookup(const uint8_t *vgrid128, float wx, float wy, float wz);    /* common.wgsl:162-177 */
/* iso_norm per leaf (iso_init.wgsl:29-56) */
void or_iso_init(const void *gaussians, const void *nodes, const uint8_t *vgrid, uint32_t n, float *iso_norm);
/* colour per Gaussian, vec4 (emissive.wgsl:853-980); lamp_T: emitter-light transmittance per Gaussian or NULL */
void or_shade_emit(const void *U, const void *EP, const void *emit_nodes, const void *gaussians, const void *nodes,
                   const uint8_t *vgrid, const void *parts, const float *lamp_T, const void *materials,
                   const float *iso_norm, float *colour);
/* emitter-light opacity map A (light_shadow.wgsl:191-392): acc [parts][res^2][3] u32 (zeroed by the caller),
 * mip [parts][layer_texels] vec4, T per Gaussian */
void or_lsm_splat(const void *U, const void *LS, const void *gaussians, const void *nodes, const void *parts,
                  uint32_t *acc);
void or_lsm_resolve(const void *LS, const uint32_t *acc, float *mip);
void or_lsm_lookup(const void *U, const void *LS, const void *gaussians, const void *parts, const float *mip,
                   float *T_out);
/* sun maps (sun_map.wgsl:91-324): layers 12 floats each; acc [layers][res^2][3]; env_T vec4 per Gaussian */
void or_sun_frame(const void *SM, const void *desc, const void *parts, float *layers);
void or_sun_splat(const void *U, const void *SM, const void *desc, const float *layers, const void *gaussians,
                  const void *nodes, const void *parts, uint32_t *acc);
void or_sun_resolve(const void *SM, const void *desc, const uint32_t *acc, float *mip);
void or_sun_lookup(const void *U, const void *SM, const void *desc, const float *layers, const void *gaussians,
                   const void *parts, const float *mip, float *env_T);
/* colour += the distant lights' direct light (environment.wgsl:207-229, 266-277) */
void or_env_direct(const void *U, const void *ENV, const void *gaussians, const void *nodes, const uint8_t *vgrid,
                   const void *parts, const void *materials, const float *iso_norm, const float *env_T,
                   float *colour);

/* ---- sky and bounce through probes (bounce.wgsl, default path; bounce.c) ------------------------------------
 * cells: 128 x (direction xyz, solid angle) floats; probe_data: 2 vec4 per (probe, cell) ray; node_acc: 14 vec4 per
 * node; node_sh: 14 u32 (27 halves) per node; probe_sh: 9 vec4 per probe; xfer: 3 vec4 per Gaussian; lo_dir: vec4
 * per Gaussian; lo_b: 3 vec4 per Gaussian (eye, send, lanes). */
uint32_t or_f16_bits(float f);
void or_bounce_lo_dir(const void *U, const void *ENV, const void *gaussians, const void *nodes, const uint8_t *vgrid,
                      const void *parts, const void *materials, const float *iso_norm, const float *env_T,
                      const float *c0, float *lo_dir);
void or_node_leaf(const void *U, const void *PP, const void *gaussians, const void *nodes, const uint8_t *vgrid,
                  const void *parts, const float *lo_dir, const float *lo_b, float *node_acc);
void or_node_up(const void *U, const void *PP, const void *nodes, const uint32_t *kids, const uint32_t *up_nodes,
                uint32_t depth, uint32_t r0, uint32_t count, float *node_acc);
void or_node_sh2(const void *PP, const float *node_acc, uint32_t *node_sh);
void or_probe_gather(const void *U, const void *PP, const void *ENV, const void *nodes, const uint8_t *vgrid,
                     const void *parts, const uint32_t *kids, const uint32_t *node_sh, const float *cells,
                     float *probe_data, uint32_t id0, uint32_t id1);
void or_probe_valid(const void *U, const void *PP, const void *parts, float *probe_data);
void or_probe_sh2(const void *PP, const float *probe_data, const float *cells, float *probe_sh);
void or_iso_xfer(const void *U, const void *PP, const void *gaussians, const void *nodes, const uint8_t *vgrid,
                 const void *parts, const float *iso_norm, const float *cells, float *xfer);
void or_probe_weights(const void *U, const void *PP, const void *nodes, const void *parts, const float *cells,
                      const float *probe_data, float px, float py, float pz, uint32_t own, uint32_t *pid, float *w);
void or_bounce_sh(const void *U, const void *PP, const void *gaussians, const void *nodes, const void *parts,
                  const float *cells, const float *probe_data, const float *probe_sh, const float *xfer, float *lo_b);
void or_probe_pos(const void *U, const void *PP, const void *nodes, const void *parts, uint32_t count, float *out);

/* shiny receivers (runtime/wgsl/material_ext.wgsl): the specular flake phase, the Duff et al. tangent pair, the
 * specular phase over a light ASG (lb: 2 x 3 floats, sl: (xx, xy, yy)) */
float or_phase_specular(const float *Sd, const float *So, const float *wi, const float *wo);
void or_phase_specular_n(const float *Sd, const float *So, const float *wi, const float *wo, uint32_t n, float *out);
void or_onb(const float *n, float *out6);
float or_asg_phase(const float *wi, const float *lb, const float *sl, const float *n0, const float *t1,
                   const float *t2, const float *flake, const float *Sd, const float *So, const float *wo);

/* screen-space LOD and display fill (lod.c): lod_select's flags per node (N, leaves [0, G) first; 0 none, 1 open,
 * 2 draw, 4 under), lod_colour's node colours (vec4 per internal node) from the leaves' colour (vec4) and bounce (3
 * vec4 per leaf, lane 0 toward the eye), node_mass, and display_fill in place on colour and lo_b's lane 0
 * (fill: offsets [G + 1], then (source index, f32 weight bits) pairs) */
void or_lod_select(const or_lod_params *L, const void *nodes, const uint32_t *kids, const void *parts, uint32_t N,
                   uint32_t G, uint32_t *flag);
float or_node_mass(const void *nodes, uint32_t n);
float or_lod_node_size(const or_lod_params *L, const void *nodes, const void *parts, uint32_t n);
void or_lod_colour(const void *nodes, const uint32_t *kids, uint32_t N, uint32_t G, const uint32_t *flag,
                   const float *colour, const float *lo_b, float gain, float *node_colour);
void or_display_fill(const uint32_t *fill, uint32_t G, float *colour, float *lo_b);

/* display pass (display.c; shaders/display.wgsl:22-189): hdr W*H*4 floats rows top-down; env: Env (416 B) or
 * NULL; bg: SkyBg (1520 B) or NULL; out: W*H*4 floats (the encoded colour before the 8-bit store, alpha 1) or
 * NULL; bytes: W*H*4 RGBA8 (round half up of 255 v) or NULL */
void or_display(const void *frame, const float *hdr, uint32_t W, uint32_t H, const void *env, const void *bg,
                float *out, uint8_t *bytes);
void or_tone_map_n(const float *x, uint32_t n, uint32_t mode, float *y);
void or_srgb_encode_n(const float *x, uint32_t n, float *y);
void or_heat_n(const float *t, uint32_t n, float *y);
void or_sky_n(const void *env, const void *bg, const float *w, uint32_t n, uint32_t which, float *y);
void or_view_dir_n(const void *frame, const float *pxy, uint32_t n, float *w);

#ifdef __cplusplus
}
#endif
#endif
