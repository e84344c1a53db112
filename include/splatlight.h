/* splatlight_hip: the C ABI of the MI355X splat frame path.
 *
 * The reference reaches the GPU through WebGPU from its JavaScript host; this ABI is what a thin N-API addon
 * (or ctypes) binds in place of those WebGPU calls for the splat frame path. All entry points are extern "C",
 * take plain pointers and sizes, and return an sl_status. Host byte buffers are copied before the call returns
 * (the caller keeps ownership, as queue.writeBuffer does). Device work is enqueued on the context's HIP stream.
 *
 * Errors: a call that fails returns a non-zero status and sets sl_last_error(ctx). A HIP (device) error
 * poisons the context: every later call on it returns SL_ERR_POISONED, with no retry (the containment of
 * src/main.js:74-77).
 */
#ifndef SPLATLIGHT_H
#define SPLATLIGHT_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int sl_status;
#define SL_OK 0
#define SL_ERR_INVALID 1       /* bad argument */
#define SL_ERR_LAYOUT 2        /* record stride / offset / count / byte-size mismatch (src/asset.js:18-38) */
#define SL_ERR_DEVICE 3        /* a HIP call failed; the context is now poisoned */
#define SL_ERR_POISONED 4      /* an earlier device error poisoned this context */
#define SL_ERR_UNSUPPORTED 5   /* a feature outside this path (explode, particles, unified lights, ...) */
#define SL_ERR_BUSY 6          /* the GPU session is held elsewhere */
#define SL_ERR_CAPACITY 7      /* a budget or size limit would be exceeded */

typedef struct sl_context sl_context;
typedef struct sl_scene sl_scene;

/* ---- GPU session: replaces acquireGpuSession (src/gpu-session.js:3-14) ----------------------------------------
 * An exclusive, non-blocking lock named `name` (NULL: "gs-game-engine-gpu"). Fails at once with SL_ERR_BUSY and
 * the reference's message when another holder has it. */
sl_status sl_session_acquire(const char *name, int64_t *token);
sl_status sl_session_release(int64_t token);
const char *sl_session_error(void);

/* ---- context: one device, one HIP stream (the reference's device + queue, src/main.js:56-62) ----------------- */
sl_status sl_ctx_create(int device_ordinal, sl_context **out);
void sl_ctx_destroy(sl_context *ctx);
const char *sl_last_error(const sl_context *ctx);
sl_status sl_ctx_status(const sl_context *ctx);
sl_status sl_sync(sl_context *ctx);      /* wait for the context's stream (onSubmittedWorkDone) */

/* ---- device buffers (createBuffer / writeBuffer / mapAsync read) ------------------------------------------- */
sl_status sl_buf_alloc(sl_context *ctx, size_t bytes, void **dev);
sl_status sl_buf_free(sl_context *ctx, void *dev);
sl_status sl_buf_write(sl_context *ctx, void *dev, size_t offset, const void *host, size_t bytes);
sl_status sl_buf_read(sl_context *ctx, const void *dev, size_t offset, void *host, size_t bytes);

/* ---- scene: the asset records the renderer and lighting upload (src/renderer.js:54-62, src/asset.js:8-16) --- */
/* Device storage of the Gaussian attributes. F32: the records' values. F16 / BF16: quaternion, scale and albedo
 * rounded to half / bfloat16 (nearest even) at sl_scene_create, means, opacity (kappa), part and flags kept exact:
 * 48 B per Gaussian on the device instead of 64. All arithmetic stays f32 on the stored values, and depth keys
 * depend on the means only, so they are the f32 scene's. */
#define SL_STORAGE_F32 0u
#define SL_STORAGE_F16 1u
#define SL_STORAGE_BF16 2u
typedef struct sl_scene_desc {
  const void *gaussians; uint32_t gaussian_count; uint32_t gaussian_stride;   /* 64 B records, <= 2^24 */
  const void *parts; uint32_t part_count; uint32_t part_stride;               /* 128 B records, <= 32 */
  const void *materials; uint32_t material_count; uint32_t material_stride;   /* 32 B records (may be 0) */
  const void *nodes; uint32_t node_count; uint32_t node_stride;               /* 96 B, leaves first (relighting) */
  const void *vgrid; uint32_t vgrid_count; uint32_t vgrid_stride;             /* 128 B grids (NULL: all 255) */
  const void *emitters; uint32_t emitter_count;                               /* world-space emissive records */
  uint32_t storage;                                                           /* SL_STORAGE_* */
} sl_scene_desc;
sl_status sl_scene_create(sl_context *ctx, const sl_scene_desc *desc, sl_scene **out);
void sl_scene_destroy(sl_context *ctx, sl_scene *scene);
/* SplatRenderer.writeParts (src/renderer.js:407-409): the whole part table, part_count * 128 B */
sl_status sl_scene_write_parts(sl_context *ctx, sl_scene *scene, const void *part_bytes, size_t bytes);
/* the lit colour per Gaussian (vec4) the renderer reads (bindColour, src/renderer.js:264-268); host copy */
sl_status sl_scene_write_colour(sl_context *ctx, sl_scene *scene, const float *rgba, size_t count);
/* the bounce per Gaussian (3 vec4 each, lo_b); NULL clears it (bounce gain then adds nothing) */
sl_status sl_scene_write_bounce(sl_context *ctx, sl_scene *scene, const float *lo_b, size_t count);
/* The Gaussian records rewritten in place, count * 64 B (the Lighting constructor's re-upload after pointing the box
 * parts' Gaussians at their own materials, src/lighting.js:467). Waits for the context's streams first; iso_norm and
 * the bounce's leaf weights are formed again at the next encode that reads them. */
sl_status sl_scene_write_gaussians(sl_context *ctx, sl_scene *scene, const void *gaussian_bytes, size_t bytes);
/* The material table rewritten, 1 to SL_MAX_MATERIALS records of 32 B (Lighting's writes of its materials buffer:
 * setBoxMaterial, setMaterialShine, setRoughness, src/lighting.js:1152-1227). Waits for the context's streams. */
#define SL_MAX_MATERIALS 256u
sl_status sl_scene_write_materials(sl_context *ctx, sl_scene *scene, const void *material_bytes, size_t bytes);
size_t sl_scene_device_bytes(const sl_scene *scene);

/* ---- the real-asset path: screen-space LOD and display fill ------------------------------------------------
 * The asset's tree as the LOD (runtime/wgsl/lod.wgsl, src/lod.js; on by default at 2 px, src/main.js:113): every
 * internal node is drawn as its moment-matched Gaussian in place of its subtree when it is small on screen. Once the
 * scene holds its tree, a frame whose RenderFrame.lod (words 180-182: threshold px as f32 bits, first node slot,
 * internal node count; src/main.js:548) is non-zero selects the cut (lod_select: part roots are candidates, a
 * candidate opens when its on-screen size, by the projected-area rule, exceeds the threshold, and the emitters' part
 * always opens), colours the drawn nodes from their leaves' lit colour plus bounce (lod_colour) and draws the cut
 * in place of the leaves. The internal nodes take the splat slots after the emitters, in node order (the reference
 * puts them after its 16,384 emitter slots: the same relative order, so the same depth order); any first-node-slot
 * word past this scene's emitters is accepted. A LOD frame renders alone (sl_frame_encode, not a batch, the
 * server or the no-alloc entry point); its cut lives in the context. */
typedef struct sl_lod_tree {
  const void *nodes; uint32_t node_count; uint32_t node_stride;   /* all N nodes, 96 B (nodes.bin): leaves [0, G) */
  const uint32_t *kids; uint32_t kid_count;                        /* kids.bin */
  const void *node_gaussians; uint32_t node_gaussian_count;        /* the N - G internal nodes as 64 B Gaussian
                                                                      records (lod.js nodeGaussians) */
  uint32_t emit_part;                                              /* the emitters' part, or 0xFFFFFFFF */
  uint32_t size_rule;                                              /* 1 projected area (default), 0 radius */
} sl_lod_tree;
/* Lod's constructor (src/lod.js:114-140) with nodeGaussians' records: validates the tree (leaves first, every
 * internal node's children one depth below it, each node at most one parent, depth <= 64) and keeps it on the
 * device. Once per scene. */
sl_status sl_scene_set_lod(sl_context *ctx, sl_scene *scene, const sl_lod_tree *tree);
/* Lighting.loadFill (src/lighting.js:961-985): the asset's display fill table (onepath.bin `fill`: u32 offsets
 * [count + 1], then (source index, f32 weight bits) pairs; targets and sources disjoint). */
sl_status sl_scene_set_fill(sl_context *ctx, sl_scene *scene, const uint32_t *fill, size_t words);
/* Lighting.encodeFill (runtime/wgsl/fill.wgsl:15-33), after the lighting and the bounce: each fill target's colour
 * and bounce toward the eye become the weighted sum of its sources'. Enqueued on the context's stream. */
sl_status sl_fill_encode(sl_context *ctx, sl_scene *scene);

/* ---- frame: SplatRenderer.encode (src/renderer.js:411-524) up to the linear HDR target ---------------------- */
#define SL_OUT_RGBA32F 0u
#define SL_OUT_RGBA16F 1u      /* the reference's rgba16float target (src/renderer.js:13) */
#define SL_FRAME_TIMERS 1u     /* record hipEvent pairs per pass (sl_timers_read) */
/* The depth order by the reference's counting scatter (shaders/scatter.wgsl:9-18, ?sort=counting, src/main.js:204)
 * instead of the stable radix sort: keys ascending, equal keys in the order their atomics land (not deterministic,
 * as in the reference); one view only (sl_batch_encode refuses it for several). */
#define SL_FRAME_COUNTING_SORT 4u
typedef struct sl_frame_opts {
  uint32_t width, height;      /* must equal the RenderFrame viewport */
  uint32_t out_format;         /* SL_OUT_* */
  uint32_t flags;              /* SL_FRAME_* */
} sl_frame_opts;
/* Enqueue one frame: project, depth order, pair emission, tile sort, tile ranges, composite. out_device holds
 * width * height pixels of out_format, rows top-down. frame_bytes: the 1200 B RenderFrame. Encoded after
 * sl_lighting_encode, the frame's geometry stages run beside that lighting (they read nothing it writes) and only
 * the colours and the composite wait for it, the composite on a stream of its own, so that the next lighting
 * encode may run beside it; the results are the same. */
sl_status sl_frame_encode(sl_context *ctx, sl_scene *scene, const void *frame_bytes, size_t frame_len,
                          const sl_frame_opts *opts, void *out_device);
typedef struct sl_frame_stats {
  uint32_t slots;        /* splat slots projected: count + 1 + emitters */
  uint32_t visible;      /* valid depth keys (draw_args[1], src/main.js:723) */
  uint32_t pairs;        /* (splat, tile) pairs */
  uint32_t overflow;     /* pairs that did not fit (the frame was re-rendered with larger buffers if > 0) */
  uint32_t tiles;        /* 16x16 tiles */
  uint32_t pad[3];
} sl_frame_stats;
/* Wait for the last frame and report its counts. Re-runs the frame once with grown pair buffers when the
 * pair count exceeded them; when a lighting was encoded after the frame (its colours are then gone) it grows the
 * buffers for the next frame and returns SL_ERR_CAPACITY instead (the image is incomplete, stats filled). */
sl_status sl_frame_finish(sl_context *ctx, sl_frame_stats *stats);

/* sl_frame_encode for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference. It runs the same kernels with the same
 * results, but its workspace (splats, keys, order, tile boxes, pair lists, counters) lies in a fixed device array of
 * SL_NOALLOC_SCRATCH_BYTES per device, held from library load to process end, instead of the context's growable
 * workspace: the call allocates and frees nothing. The frame's slots (count + 1 + emitters) and tiles must fit
 * beside at least one pair per slot (about 80 B per slot: 200k Gaussians at 640x480 fit), or it returns
 * SL_ERR_CAPACITY at once; the rest of the array holds the pair lists, and sl_frame_finish returns SL_ERR_CAPACITY
 * (stats still filled) instead of re-running when the pairs overflow it. Every such call on one device shares that
 * device's array, which keeps the last call's contents: calls on one device that overlap in time (from different
 * contexts or threads, or before the previous frame's sl_frame_finish / sl_frame_read) race. */
#define SL_NOALLOC_SCRATCH_BYTES (64u << 20)
sl_status sl_frame_encode_noalloc(sl_context *ctx, sl_scene *scene, const void *frame_bytes, size_t frame_len,
                                  const sl_frame_opts *opts, void *out_device);

/* ---- display: the pass after the frame (shaders/display.wgsl:22-189, src/renderer.js:475-485) ----------------
 * The linear HDR target (what sl_frame_encode wrote, rows top-down) resolved to 8-bit display colour: in the lit
 * view (RenderFrame.view_mode 3) the environment's background sky (SkyBg) and its lights' discs (Env) fill what the
 * splats leave uncovered (1 - alpha > 1/1024), then RenderFrame.exposure, the tone curve RenderFrame.tone.x (0 clamp,
 * 1 AgX, 2 ACES fit, 3 Reinhard on luminance, 4 Khronos PBR Neutral, 5 Hable filmic) and the sRGB encode; view_mode 2
 * maps the overdraw count to the heat ramp. Each channel is the byte floor(255 v + 0.5) of the shader's expression
 * with exp, pow, log2, cos and sin correctly rounded (the precision WGSL leaves open, fixed). Enqueued on the
 * context's stream after whatever wrote hdr_device; sl_sync waits for it. */
#define SL_DISPLAY_RGBA8 0u    /* rgba8unorm bytes r, g, b, 255 */
#define SL_DISPLAY_BGRA8 1u    /* bgra8unorm (navigator.gpu.getPreferredCanvasFormat() on most systems) */
#define SL_DISPLAY_EXACT 1u    /* flag: evaluate every pixel with the correctly rounded functions (same bytes) */
typedef struct sl_display_opts {
  uint32_t width, height;      /* must equal the RenderFrame viewport */
  uint32_t hdr_format;         /* SL_OUT_RGBA32F or SL_OUT_RGBA16F: the target's format */
  uint32_t out_format;         /* SL_DISPLAY_* */
  uint32_t flags;              /* SL_DISPLAY_EXACT */
} sl_display_opts;
/* env: the 416 B Env uniform (src/environment.js:111-134; NULL: no environment), sky_bg: the 1520 B SkyBg uniform
 * (src/environment.js:135-150; NULL: none, the lighting sky's 16 coefficients are drawn instead);
 * SplatRenderer.bindEnv (src/renderer.js:400-404) and the display pass of SplatRenderer.encode. out_device holds
 * width * height * 4 bytes. */
sl_status sl_display_encode(sl_context *ctx, const void *frame_bytes, size_t frame_len, const void *env,
                            size_t env_len, const void *sky_bg, size_t sky_bg_len, const sl_display_opts *opts,
                            const void *hdr_device, void *out_device);

/* ---- batches: many views over resident scenes in one launch per stage -------------------------------------
 * The views (each: a scene, its 1200 B RenderFrame, its own target size and format) are laid out back to back in
 * one slot space and one tile space; project, the depth sort (keyed by view, then depth), pair emission, the tile
 * sort, the ranges, the heaviest-first order and the composite each run once over all of them. Every view's
 * image, keys order and pairs are those it gets alone. With SL_BATCH_GRAPH the stage sequence is captured as a
 * hipGraph the first time and replayed while the layout (view count, sizes, scenes' slot counts) is unchanged;
 * camera and frame changes only re-upload the view table. No counterpart in the reference (one frame per
 * requestAnimationFrame). */
typedef struct sl_view {
  sl_scene *scene;
  const void *frame_bytes;     /* 1200 B RenderFrame */
  size_t frame_len;
  uint32_t width, height;      /* must equal the RenderFrame viewport */
  uint32_t out_format;         /* SL_OUT_* */
  void *out_device;            /* width * height pixels, rows top-down */
} sl_view;
/* Launch shape of the frame kernels, picked per workload by an autotuner (splatlight_amd/autotune.py); every
 * setting gives the same results. composite_batch: splats staged in LDS per round (64, 128, 256); radix_items: keys
 * per lane of an onesweep tile (8: 2048-key tiles, 16: 4096); grid_mul: persistent radix workgroups per CU (1-16).
 * Defaults 256, 16, 4. No counterpart in the reference. */
typedef struct sl_tuning {
  uint32_t composite_batch;    /* splats staged in LDS per composite round: 64, 128 or 256 */
  uint32_t radix_items;        /* keys per lane of an onesweep tile: 8 or 16 */
  uint32_t grid_mul;           /* persistent sort workgroups per CU: 1-16 */
  uint32_t radix_bits;         /* the tile sort's digit width: 8 or 11 (0: 8) */
  uint32_t tile_px;            /* the binning tile: 16 or 32 pixels square (0: 16); composited in 16x16 blocks */
  uint32_t pad;
} sl_tuning;
sl_status sl_ctx_set_tuning(sl_context *ctx, const sl_tuning *tuning);
sl_status sl_ctx_get_tuning(sl_context *ctx, sl_tuning *tuning);
#define SL_BATCH_GRAPH 2u      /* capture / replay the batch as a hipGraph (timers are not recorded) */
sl_status sl_batch_encode(sl_context *ctx, const sl_view *views, uint32_t n_views, uint32_t flags);
/* wait for the last batch (re-running it once with grown pair buffers on overflow); per-view stats */
sl_status sl_batch_finish(sl_context *ctx, sl_frame_stats *stats, uint32_t n_stats);

/* ---- serving: continuous batching of render requests over resident scenes --------------------------------
 * A worker thread owns `lanes` execution lanes (each its own stream and workspace on the context's device). Whenever
 * a lane is free it admits waiting requests, in arrival order, into one batch while the batch has at most
 * max_batch_views views and the workspace the lane needs for it fits beside the other lanes' workspaces within
 * hbm_budget, and launches it; up to `lanes` batches run at once, so requests arriving while others render are
 * admitted as lanes and memory free up. The budget bounds the device memory the lanes allocate (their workspaces:
 * splat slots, sort buffers, pair lists, tile ranges; the targets are the caller's): a lane grows to at most what
 * the others leave (a pair-list overflow that would pass it fails its requests with SL_ERR_CAPACITY), idle lanes
 * give their workspaces back when a batch cannot fit otherwise, and a request whose workspace alone exceeds the
 * budget is refused by sl_server_submit with SL_ERR_CAPACITY.
 * Each result is the view rendered alone (same kernels, per-view state). The scenes and targets must stay alive
 * until their results are polled. No counterpart in the reference (one frame per requestAnimationFrame). */
typedef struct sl_server sl_server;
typedef struct sl_server_opts {
  uint64_t hbm_budget;         /* bytes the lanes' workspaces may hold together */
  uint32_t lanes;              /* 1-8 concurrent batches */
  uint32_t max_batch_views;    /* 1-256 */
} sl_server_opts;
typedef struct sl_request_result {
  uint64_t id;
  sl_status status;
  uint32_t batch_views;        /* size of the batch it ran in */
  sl_frame_stats stats;
  double submit_s, start_s, done_s;   /* steady clock, seconds since sl_server_create */
} sl_request_result;
sl_status sl_server_create(sl_context *ctx, const sl_server_opts *opts, sl_server **out);
/* copies the frame bytes; the view's target receives the image. The request's lane waits for the work queued on
 * the server's context when it is submitted (a lighting encode of the scene there needs no sl_sync first). */
sl_status sl_server_submit(sl_server *server, const sl_view *view, uint64_t *request_id);
/* completed requests (up to cap), waiting up to wait_us for the first when none is ready */
sl_status sl_server_poll(sl_server *server, sl_request_result *out, uint32_t cap, uint32_t *n, uint32_t wait_us);
/* in_flight_bytes: the lanes' workspaces (held, or reserved for their running batches) */
sl_status sl_server_info(sl_server *server, uint32_t *pending, uint32_t *running, uint64_t *in_flight_bytes);
const char *sl_server_error(sl_server *server);   /* the last lane error (a copy, valid until this thread calls again) */
/* Failure-handling check from the host: the worker's next completion check of a busy lane reports a device error
 * instead of querying the lane's event, as a faulted batch would; its requests complete with SL_ERR_DEVICE, the lane
 * is poisoned (later batches on it complete with SL_ERR_POISONED) and sl_server_error names it. The device is not
 * touched. No counterpart in the reference. */
sl_status sl_server_fail_next(sl_server *server);
/* finishes every submitted request, then stops the worker */
void sl_server_destroy(sl_server *server);

/* Artefacts of the last frame, for parity checks (a batch: the global slot space; view v's slots start at the
 * sum of the earlier views' slot counts rounded up to 256, its keys carry v in bits 16-23, its tile ids start at
 * the sum of the earlier views' tile counts): */
#define SL_ART_KEYS 1u         /* u32 per slot: 16-bit depth key or 0xFFFFFFFF (project.wgsl:366-368) */
#define SL_ART_ORDER 2u        /* u32 per slot: stable order by key, INVALID last (radix.wgsl) */
#define SL_ART_SPLATS 3u       /* 48 B per slot (shaders/common.wgsl:29-35); invalid slots undefined */
#define SL_ART_PAIR_TILES 4u   /* u32 per pair: tile id, ascending */
#define SL_ART_PAIR_SLOTS 5u   /* u32 per pair: splat slot, each tile's list in back-to-front order */
#define SL_ART_RANGES 6u       /* 2 u32 per tile: [start, end) into the pair lists */
#define SL_ART_LOD_FLAGS 7u    /* a LOD frame: u32 per tree node, lod_select's flag (0 none, 1 open, 2 draw, 4 under) */
#define SL_ART_LOD_COLOUR 8u   /* a lit LOD frame: vec4 per internal node, lod_colour's colour (drawn nodes and the
                                  nodes under them; others undefined) */
sl_status sl_frame_read(sl_context *ctx, uint32_t which, void *host, size_t bytes, size_t *written);

/* ---- relighting: Lighting.encode (src/lighting.js:1258-1274) for the default static M39 matte path ----------------
 * The host keeps the reference's parameter logic and hands over the uniform bytes its writers produce. The scene
 * must carry nodes (leaf Node records), materials and, optionally, vgrid (absent: every cell 255, V = 1). */
typedef struct sl_lighting_desc {
  const void *frame_uniforms;   /* 64 B FrameUniforms (runtime/wgsl/common.wgsl:39-47, src/lighting.js:1229-1238) */
  const void *emit_params;      /* 64 B EmitParams (runtime/wgsl/emissive.wgsl:34-53, src/lighting.js:657-673) */
  const void *emit_nodes;       /* EmitNode records, 96 B (src/lighting.js:61-65) */
  uint32_t emit_node_count;
  const void *ls_params;        /* 336 B LSParams (src/light_shadow.js:113-140); NULL: no emitter shadows (T = 1) */
  const void *env;              /* 416 B Env (src/environment.js:111-153); NULL: no distant lights */
  const void *sun_params;       /* 432 B SunParams (src/sun_map.js:133-150); NULL: distant lights unshadowed */
  const void *sun_layers;       /* LayerDesc records, 32 B (src/sun_map.js:127-132) */
  uint32_t sun_layer_count;
  uint32_t flags;               /* SL_LIGHT_TIMERS */
} sl_lighting_desc;
#define SL_LIGHT_TIMERS 1u
/* iso_norm of every leaf from its V grid (runIsoInit, src/lighting.js:546-555); once per asset */
sl_status sl_lighting_iso_init(sl_context *ctx, sl_scene *scene);
/* the emitter and sun opacity maps of dirty parts / layers (LSParams.dirty, LayerDesc.dirty), their lookups,
 * shade_emit and env_direct, leaving the lit colour the next sl_frame_encode draws. It may be encoded before the
 * last frame's sl_frame_finish (as a render loop encodes the next frame): it then runs beside that frame's
 * composite, which reads nothing it writes. */
sl_status sl_lighting_encode(sl_context *ctx, sl_scene *scene, const sl_lighting_desc *desc);

/* Scene-resident lighting state, for parity checks: */
#define SL_SCENE_COLOUR 1u     /* vec4 per Gaussian */
#define SL_SCENE_LAMP_T 2u     /* float per Gaussian: emitter-light transmittance (lamp_state slot 0) */
#define SL_SCENE_ENV_T 3u      /* vec4 per Gaussian: per distant light transmittance */
#define SL_SCENE_ISO_NORM 4u   /* float per Gaussian (Node.pad0) */
#define SL_SCENE_LSM_ACC 5u    /* u32 [parts][res^2][3]: the emitter map's fixed-point sums */
#define SL_SCENE_LSM_MIP 6u    /* vec4 [parts][layer texels] */
#define SL_SCENE_SUN_ACC 7u    /* u32 [layers][res^2][3] */
#define SL_SCENE_SUN_MIP 8u    /* vec4 [layers][layer texels] */
sl_status sl_scene_read(sl_context *ctx, sl_scene *scene, uint32_t which, void *host, size_t bytes, size_t *written);

/* Sizes of the last relighting's opacity-map splats on this context (binned records and pairs), for throughput
 * figures. No counterpart in the reference (its splats are one atomic loop per Gaussian). */
typedef struct sl_lighting_stats {
  uint32_t lsm_records, lsm_pairs, lsm_tiles, lsm_tiles_used;   /* emitter map: record slots, (record, tile) */
  uint32_t sun_records, sun_pairs, sun_tiles, sun_tiles_used;   /* sun maps; tiles: all, with records */
} sl_lighting_stats;
sl_status sl_lighting_stats_read(sl_context *ctx, sl_scene *scene, sl_lighting_stats *out);

/* ---- sky and bounce through probes: Lighting.initBounce / encodeBounce (src/lighting.js:723-840, 1020-1103) ----
 * The default path of runtime/wgsl/bounce.wgsl: no sender LOD, no near field, M39 matte receivers with SH2 probes
 * (FrameUniforms.iso bits 0 and 1). The scene's tree (the reference's nodes.bin + kids.bin): 96 B Node records,
 * the scene's leaves first (nodes [0, gaussian_count)), internal nodes after, children kids[child_lo, child_hi);
 * each part's tree starts at Part.root (written with the part table). max_depth: the asset's walk.max_depth (node
 * sums run for depths max_depth - 1 .. 1). The tree's records replace the scene's leaf array. */
typedef struct sl_bounce_tree {
  const void *nodes; uint32_t node_count; uint32_t node_stride;   /* 96 B */
  const uint32_t *kids; uint32_t kid_count;
  uint32_t max_depth;
} sl_bounce_tree;
/* probe_params: the 9632 B ProbeParams initBounce writes (runtime/wgsl/bounce.wgsl:23-56, src/lighting.js:776-797);
 * its layout fixes the probe count. Allocates the bounce state (zeros) and makes the scene's bounce (lo_b, which
 * sl_frame_encode adds with the RenderFrame's bounce gain) the device-computed one. */
sl_status sl_bounce_init(sl_context *ctx, sl_scene *scene, const sl_bounce_tree *tree, const void *probe_params,
                         size_t probe_params_len);
#define SL_BOUNCE_TIMERS 1u       /* "bounce" and "probe_gather" in sl_timers_read */
#define SL_BOUNCE_XFER 2u         /* run iso_xfer (the reference's isoStale: a forced refresh or a part moved) */
#define SL_BOUNCE_LIGHT_STILL 4u  /* lightStill: only the eye-dependent read (bounce_sh) */
/* One encodeBounce after the scene's sl_lighting_encode of the same frame (it reads that encode's FrameUniforms,
 * Env, shadow T and the shading's emitter light): the senders' light toward the probes, the node sums and SH2, the
 * probe gather (sky at escape), probe validity and SH2, the transfer when asked, and the bounce into lo_b. On the
 * context's stream, so the next sl_frame_encode's colours see it. probe_params: this frame's ProbeParams (alpha,
 * force_near, subset, phase as encodeBounce sets them; alpha must lie in (0, 1]). */
sl_status sl_bounce_encode(sl_context *ctx, sl_scene *scene, const void *probe_params, size_t probe_params_len,
                           uint32_t flags);
/* Bounce state, for parity checks (after the context's stream finishes): */
#define SL_BOUNCE_LO_DIR 1u       /* vec4 per Gaussian: light toward the probes */
#define SL_BOUNCE_LO_B 2u         /* 3 vec4 per Gaussian: bounce toward the eye, the probes, the lanes */
#define SL_BOUNCE_NODE_ACC 3u     /* 14 vec4 per node: (sum w L, sum w); internal nodes only (leaves' rows zero:
                                     they are formed from LEAF_W and the light where they are read) */
#define SL_BOUNCE_NODE_SH 4u      /* 14 u32 per node: SH2 rgb as 27 halves */
#define SL_BOUNCE_PROBE_DATA 5u   /* 2 vec4 per (probe, cell): (rgb, first-hit part), (m1, m2, valid, near) */
#define SL_BOUNCE_PROBE_SH 6u     /* 9 vec4 per probe */
#define SL_BOUNCE_XFER_SH 7u      /* 3 vec4 per Gaussian: the transfer's SH2 */
#define SL_BOUNCE_C0 8u           /* vec4 per Gaussian: the shading's emitter light before own emission */
#define SL_BOUNCE_LEAF_W 9u       /* 14 floats per Gaussian: its weights A V toward the 14 directions */
sl_status sl_bounce_read(sl_context *ctx, sl_scene *scene, uint32_t which, void *host, size_t bytes, size_t *written);

/* Per-pass GPU milliseconds of the last frame encoded with SL_FRAME_TIMERS; names as in the reference's PASSES
 * (src/renderer.js:7-8) where one exists, then, after sl_lighting_encode with SL_LIGHT_TIMERS, its passes
 * "shadow", "shade", "sun_wait", "env_direct", "sunmap" (the sun-map chain, which runs on a second stream during
 * shadow and shade; sun_wait is what remains of it after shade), "lighting" (the whole encode) and the map-sum
 * kernels alone, "lsm_sum" and "sun_sum" (of the last such encode, or of the one before it while the last still
 * runs); after sl_bounce_encode with SL_BOUNCE_TIMERS, "bounce" (the whole
 * encode) and "probe_gather". Fills up to `cap` entries; returns the count in *n. */
sl_status sl_timers_read(sl_context *ctx, const char **names, float *ms, uint32_t cap, uint32_t *n);

#ifdef __cplusplus
}
#endif
#endif
