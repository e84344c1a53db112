// libsplatlight_hip: host side of the C ABI (include/splatlight.h) and the frame pipeline.
//
// One translation unit: the kernels are included here so every launch resolves in the same code object.
#include <hip/hip_runtime.h>

#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <string.h>
#include <sys/file.h>
#include <unistd.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/splatlight.h"
#include "sl_device.h"

#include "project.hip"
#pragma clang fp contract(fast)
#include "composite.hip"
#include "sort.hip"
#include "tiles.hip"
#include "lighting.hip"
#include "bounce.hip"
#include "display.hip"
#include "lod.hip"
#pragma clang fp contract(fast)

namespace {

// AoS 64 B records -> four float4 streams
__global__ void __launch_bounds__(256) sl_aos_to_soa(const float4 *__restrict__ aos, float4 *__restrict__ s0,
                                                     float4 *__restrict__ s1, float4 *__restrict__ s2,
                                                     float4 *__restrict__ s3, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  s0[i] = aos[4 * (size_t)i + 0];
  s1[i] = aos[4 * (size_t)i + 1];
  s2[i] = aos[4 * (size_t)i + 2];
  s3[i] = aos[4 * (size_t)i + 3];
}

constexpr uint32_t kGaussianStride = 64, kPartStride = 128, kMaterialStride = 32, kNodeStride = 96,
                   kVgridStride = 128, kFrameBytes = 1200, kMaxCount = 1u << 24;

const char *const kPassNames[] = {"clear", "project", "radix", "pairs", "tilesort", "ranges", "splat"};
constexpr int kPasses = 7;
constexpr uint32_t kMaxViews = 256;   // the view id sits in bits 16-23 of the depth sort key
constexpr size_t kViewBlock = sizeof(ViewDev) * kMaxViews + 4 * (3 * kMaxViews + 2);

// A batch as encoded: the views' device descriptors and the global slot / tile layout.
struct Batch {
  std::vector<ViewDev> views;
  std::vector<sl_scene *> scenes;
  std::vector<uint32_t> slot_base, tile_base, tiles_x;   // V + 1, V + 1, V
  uint32_t total_slots = 0, total_tiles = 0;
  uint32_t depth_passes = 2, tile_passes = 2;
  uint32_t tile_bits = 8;  // the tile sort's digit width (8 or 11)
  uint32_t tile_shift = 4; // log2 of the composite tile's side (4: 16 px, 5: 32 px)
  uint32_t flags = 0;
  bool graph = false;
  bool fixed = false;   // laid over the device's no-alloc scratch instead of the context's workspace
  bool counting = false;   // depth order by the counting scatter (SL_FRAME_COUNTING_SORT) instead of the radix sort
  bool blur = false;       // some view blurs its moving parts (shutter > 0): the projection's motion-blur variant
  bool lod = false;        // the (single) view draws a screen-space LOD cut: lod_select / lod_colour first
  LodDev lod_dev{};
};

uint32_t rd_u32(const void *p, int word) { uint32_t v; memcpy(&v, (const char *)p + 4 * word, 4); return v; }
float rd_f32(const void *p, int word) { float v; memcpy(&v, (const char *)p + 4 * word, 4); return v; }

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

// A device buffer that grows on demand and is owned by a scene.
struct Grow {
  void *p = nullptr;
  size_t bytes = 0;
};

struct sl_scene {
  uint32_t count = 0, part_count = 0, n_emit = 0, material_count = 0, storage = SL_STORAGE_F32;
  uint32_t parts_version = 0;       // counts sl_scene_write_parts calls (what depends on the poses rebuilds)
  float4 *g[4] = {nullptr, nullptr, nullptr, nullptr};
  float4 *e[4] = {nullptr, nullptr, nullptr, nullptr};
  float4 *parts = nullptr;
  float4 *colour = nullptr;
  float4 *lo_b = nullptr;
  std::vector<uint8_t> parts_host;
  std::vector<uint8_t> materials_host;
  size_t materials_cap = 0;         // bytes of the device materials buffer
  std::vector<void *> allocs;
  size_t bytes = 0;
  // relighting state
  float4 *nodes = nullptr;          // leaf Node records (AoS, 96 B)
  uint8_t *vgrid = nullptr;         // 128 B per Gaussian
  uint32_t *materials = nullptr;    // 32 B records
  float *iso_norm = nullptr;
  bool iso_done = false;
  uint32_t *order = nullptr;      // per-Gaussian kernels visit the Gaussians in this order (spatially coherent waves)
  // the shading's and map records' per-Gaussian inputs copied into that order (row t = Gaussian order[t]): the
  // record streams as stored, the leaf nodes, the V-grids and iso_norm; formed at the lighting encode after any of
  // them changed (valid false)
  struct Visit {
    float4 *g[4] = {nullptr, nullptr, nullptr, nullptr};
    float4 *nodes = nullptr;
    uint8_t *vgrid = nullptr;
    float *iso = nullptr;
    bool valid = false;
  } visit;
  float *lamp_T = nullptr;
  float4 *env_T = nullptr;
  Grow emit_nodes, lsm_acc, lsm_mip, lsm_dirty, sun_acc, sun_mip, sun_desc, sun_layers, sun_dirty;

  uint32_t lsm_parts = 0, lsm_res = 0, lsm_layer_texels = 0;
  uint32_t sun_n = 0, sun_res = 0, sun_layer_texels = 0;
  // screen-space LOD (sl_scene_set_lod): the tree as the LOD kernels read it and its internal nodes as Gaussians
  struct Lod {
    bool on = false;
    uint32_t N = 0, emit_part = 0xFFFFFFFFu, size_rule = 1u;
    float4 *na = nullptr, *nb = nullptr, *nc = nullptr;
    uint32_t *up = nullptr, *part = nullptr, *kids = nullptr;
    uint2 *child = nullptr;
    float *mass = nullptr;
    float4 *ng[4] = {nullptr, nullptr, nullptr, nullptr};
  } lod;
  // display fill (sl_scene_set_fill): offsets [count + 1], then (source, weight bits) pairs
  uint32_t *fill = nullptr;
  // sky and bounce through probes (sl_bounce_init): the whole tree's nodes (leaves first) replace the leaf array
  // above; the state the bounce passes carry from frame to frame; the last lighting encode's uniforms
  struct Bounce {
    bool on = false, lit = false;
    uint32_t node_count = 0, kid_count = 0, max_depth = 0, n_probes = 0;
    uint32_t *kids = nullptr, *node_sh = nullptr;
    uint4 *up_ent = nullptr;                  // node_up's entries, depth by depth: (node, kid count, kid, kid)
    std::vector<uint32_t> lv_off, lv_cnt;     // node_up's lists per depth
    UpTail tail{};                            // depths 1..tail_hi (0: none), each within UP_TAIL_MAX nodes
    uint32_t tail_hi = 0;
    float4 *lo_dir = nullptr, *lo_b = nullptr, *node_acc = nullptr, *probe_data = nullptr, *probe_sh = nullptr;
    float4 *xfer = nullptr, *c0 = nullptr, *cells = nullptr;
    void *pp = nullptr;                        // ProbeParams
    // the leaves' weights toward the 14 directions and the part-table version they were built for; node_acc
    // holds the internal nodes' sums direction-major ([14][node_count])
    float *wcache = nullptr;
    uint32_t w_version = 0xFFFFFFFFu;
    uint32_t walk_stack = 0;   // the most entries the gather's depth-first walk can hold on this tree
    float4 *walk = nullptr;    // the gather's compact tree: 2 float4 per walk node (breadth first per part)
    uint32_t *walk_root = nullptr;
    FrameUniformsDev U{};
    EnvDev ENV{};
  } bounce;
};

struct Workspace {
  uint32_t slots_cap = 0, pairs_cap = 0, tiles_cap = 0;
  uint32_t tile_bits = 8;         // the tile sort's digit width its look-back state is laid out for
  void *base = nullptr;           // one allocation for everything below
  size_t bytes = 0;
  Splat *splats = nullptr;
  uint32_t *keys = nullptr, *keys_alt = nullptr, *vals_alt = nullptr, *order = nullptr;
  uint2 *tbox = nullptr;
  uint2 *partials = nullptr;      // project's per-block (visible, colour max)
  uint32_t *pk0 = nullptr, *pv0 = nullptr, *pk1 = nullptr, *pv1 = nullptr;
  // the batch's view table, uploaded before every batch: ViewDev[kMaxViews], then slot_base, tile_base (kMaxViews + 1
  // each) and tiles_x (kMaxViews)
  ViewDev *views = nullptr;
  uint32_t *slot_base = nullptr, *tile_base = nullptr, *tiles_x = nullptr;
  ViewCounters *view_counters = nullptr;
  // zeroed at the start of every frame:
  uint8_t *scratch = nullptr;
  size_t scratch_words = 0;
  FrameCounters *counters = nullptr;
  uint32_t *tickets = nullptr, *depth_hist = nullptr, *tile_hist = nullptr;
  uint32_t *count_hist = nullptr;  // the counting scatter's 65,537 buckets
  uint32_t *dstatus[3] = {nullptr, nullptr, nullptr}, *tstatus[3] = {nullptr, nullptr, nullptr};
  uint64_t *chunk_status = nullptr;
  uint2 *ranges = nullptr;
  uint32_t *tile_order = nullptr;  // composite launch order (written every frame)
  // 11-bit tile digits: the tile sort's two look-back arrays (2,048 words per onesweep tile) lie past the zeroed
  // block, each cleared just before its pass over the tiles the pair count fills
  size_t wide_words = 0;           // words of one such array
};

struct sl_context {
  int device = 0;
  int cus = 256;
  hipStream_t stream = nullptr;
  sl_status poisoned = SL_OK;
  std::string err;
  Workspace ws;
  FrameCounters *counters_host = nullptr;   // pinned
  hipEvent_t ev[kPasses + 1] = {};
  hipEvent_t ev_comp = nullptr;   // before the composite when the geometry ran beside the lighting
  // a lighting encode's timer events, two sets used in turn so that the set of the last completed encode can be
  // read while the next one runs: lev the context stream's (shadow, shade, sun wait, env), mev the two map-sum
  // launches (emitter map, sun maps: begin, end), sev the side stream's begin and end
  struct LightTimers {
    hipEvent_t lev[5] = {}, mev[4] = {}, sev[2] = {};
    bool valid = false, map_timed[2] = {false, false}, sun_timed = false;
  } lt[2];
  uint32_t lt_cur = 0;            // the set of the last lighting encode
  hipEvent_t bev[4] = {};         // bounce: begin, before probe_gather, after it, end
  bool bounce_timed = false;
  bool timers_valid = false;
  float ms[kPasses] = {};
  // the last batch (a frame is a batch of one), for sl_*_finish (re-run on overflow) and sl_frame_read
  bool has_last = false;
  Batch last;
  uint32_t *final_keys = nullptr, *final_vals = nullptr;   // the tile-sorted pair lists of the last batch
  ViewCounters *view_counters_host = nullptr;               // pinned, kMaxViews
  // the last captured batch graph and what it was captured for
  hipGraphExec_t graph = nullptr;
  std::vector<uint64_t> graph_key;
  sl_tuning tuning = {256u, 16u, 4u, 8u, 16u, 0u};
  size_t ws_limit = 0;   // the workspace may grow to at most this many bytes (0: no limit; a server lane's budget)
  // the no-alloc frame (sl_frame_encode_noalloc): the device's scratch array's address, looked up once per device
  // at context creation, and the workspace laid over it for the last such frame
  uint8_t *fixed_base = nullptr;
  std::string fixed_err;
  Workspace fws;
  // opacity-map splat binning, one set per map kind (0 emitter map, 1 sun maps) so the two chains can run at once,
  // shared by every scene lit on this context: records, their tile boxes, the (tile, record) pairs and their sorted
  // copies, a cleared scratch block; per-map stats (two words per kind)
  struct MapBufs {
    Grow rec, tbox, tmask, k0, v0, k1, v1, scratch;
  } map[2];
  Grow map_stats;
  Grow disp_q;        // the display pass's list of pixels to settle exactly: a count word, then one word per pixel
  Grow lod_flag, lod_open, lod_cnt, lod_col;   // the LOD cut of the last frame: flags, open bits, counts, colours
  uint32_t map_records[2] = {0, 0}, map_tiles[2] = {0, 0};
  const uint32_t *map_touched[2] = {nullptr, nullptr};   // the last sums' tile bits (map scratch)
  // the sun-map chain's stream (created on first use, default priority), the fork/join events that order it with
  // the main stream, and its begin/end timer events
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, cleared = nullptr;
  // the frame geometry's stream (created with the side stream), the event before the last lighting encode that the
  // geometry waits for, the event after it that the context's stream waits for; light_pending: a lighting encode
  // since the last batch, so the next batch's geometry may run beside it
  hipStream_t geo = nullptr;
  hipEvent_t pre_light = nullptr, geo_done = nullptr;
  bool light_pending = false, timed_overlap = false;
  // the composite of a frame encoded after a lighting runs on its own stream once the splats' colours are written
  // (colours_done), so the next lighting encode, which writes nothing the composite reads, runs beside it;
  // comp_done follows the composite and the counters' read-back. comp_pending: the context's and the geometry's
  // streams have not waited for comp_done yet (every entry point but the lighting encode makes them wait first).
  // light_seq counts lighting encodes; last_light_seq is the count when the last batch was encoded.
  hipStream_t comp = nullptr;
  hipEvent_t colours_done = nullptr, comp_done = nullptr;
  bool comp_pending = false, comp_last = false;
  uint64_t light_seq = 0, last_light_seq = 0;
};

namespace {

sl_status fail(sl_context *ctx, sl_status s, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return s;
}

sl_status hip_fail(sl_context *ctx, hipError_t e, const char *what) {
  ctx->poisoned = SL_ERR_DEVICE;
  ctx->err = std::string(what) + ": " + hipGetErrorString(e);
  return SL_ERR_DEVICE;
}

#define SL_HIP(ctx, call)                                   \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return hip_fail((ctx), e_, #call); \
  } while (0)

// The context's and the geometry's streams wait for the last composite that ran on its own stream (a failed wait
// poisons the context: later work would not be ordered after that composite).
sl_status join_composite(sl_context *ctx) {
  if (!ctx->comp_pending) return SL_OK;
  hipError_t e = hipStreamWaitEvent(ctx->stream, ctx->comp_done, 0);
  if (e == hipSuccess && ctx->geo) e = hipStreamWaitEvent(ctx->geo, ctx->comp_done, 0);
  if (e != hipSuccess) {
    ctx->poisoned = SL_ERR_DEVICE;
    ctx->err = std::string("joining the composite's stream: ") + hipGetErrorString(e);
    return SL_ERR_DEVICE;
  }
  ctx->comp_pending = false;
  return SL_OK;
}

// Every entry point: a live context, then the last composite joined (SL_LIVE_BESIDE for the lighting encode only,
// which may run beside that composite).
#define SL_LIVE_BESIDE(ctx)                                                                       \
  do {                                                                                            \
    if (!(ctx)) return SL_ERR_INVALID;                                                            \
    if ((ctx)->poisoned != SL_OK) return SL_ERR_POISONED;                                         \
  } while (0)
#define SL_LIVE(ctx)                                                                              \
  do {                                                                                            \
    SL_LIVE_BESIDE(ctx);                                                                          \
    if (join_composite(ctx) != SL_OK) return SL_ERR_DEVICE;                                       \
  } while (0)

sl_status check_launch(sl_context *ctx, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(ctx, e, what);
  return SL_OK;
}

sl_status grow_ctx(sl_context *ctx, Grow *gb, size_t bytes);

sl_status dev_alloc(sl_context *ctx, size_t bytes, void **p) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  SL_HIP(ctx, hipMalloc(p, bytes));
  return SL_OK;
}

// The no-alloc frame's workspace (sl_frame_encode_noalloc): one array per device, held until the process ends. In
// host code `scratch` names a host-side stand-in, never the device memory: its device address comes from
// hipGetSymbolAddress, per device (each device holds its own copy at its own address).
static __device__ __align__(256) uint8_t scratch[SL_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current; looked up once per device.
uint8_t *scratch_address(int device, std::string *err) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) {
    *err = "device ordinal beyond the no-alloc scratch table";
    return nullptr;
  }
  if (!g_scratch_dev[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();   // leave no sticky error for the context's later launch checks
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_dev[device] = p;
  }
  return (uint8_t *)g_scratch_dev[device];
}

// Lay the frame workspace for `s_cap` splat slots, `p_cap` pairs and `t_cap` tiles, its tile sort in `tile_bits`-bit
// digits, over `base` (nullptr: only measure); returns the bytes it spans.
size_t layout_workspace(uint8_t *b, uint32_t s_cap, uint32_t p_cap, uint32_t t_cap, uint32_t tile_bits, Workspace &w) {
  const uint32_t d_tiles = (s_cap + RS_TILE_MIN - 1) / RS_TILE_MIN;
  const uint32_t p_tiles = (p_cap + RS_TILE_MIN - 1) / RS_TILE_MIN;
  const uint32_t chunks = (s_cap + PE_CHUNK - 1) / PE_CHUNK;
  const bool wide = tile_bits > 8u;
  size_t off = 0;
  auto take = [&](size_t bytes) { const size_t o = off; off = align_up(off + bytes, 256); return o; };
  const size_t o_splats = take((size_t)s_cap * sizeof(Splat));
  const size_t o_keys = take((size_t)s_cap * 4), o_keys_alt = take((size_t)s_cap * 4);
  const size_t o_vals_alt = take((size_t)s_cap * 4), o_order = take((size_t)s_cap * 4);
  const size_t o_tbox = take((size_t)s_cap * 8);
  const size_t o_partials = take((size_t)((s_cap + 255) / 256) * 8);
  const size_t o_pk0 = take((size_t)p_cap * 4), o_pv0 = take((size_t)p_cap * 4);
  const size_t o_pk1 = take((size_t)p_cap * 4), o_pv1 = take((size_t)p_cap * 4);
  const size_t o_views = take(kViewBlock);
  const size_t o_vcount = take(sizeof(ViewCounters) * kMaxViews);
  const size_t wide_words = (size_t)p_tiles * 2048;
  const size_t o_wide0 = wide ? take(wide_words * 4) : 0, o_wide1 = wide ? take(wide_words * 4) : 0;
  const size_t o_scratch = off;
  const size_t o_counters = take(sizeof(FrameCounters));
  const size_t o_tickets = take(16 * 4);
  const size_t o_dhist = take(3 * 256 * 4), o_thist = take(3 * 2048 * 4);
  const size_t o_chist = take(65537 * 4);
  const size_t o_ds0 = take((size_t)d_tiles * 256 * 4), o_ds1 = take((size_t)d_tiles * 256 * 4);
  const size_t o_ds2 = take((size_t)d_tiles * 256 * 4);
  const size_t t_words = wide ? 0 : (size_t)p_tiles * 256;
  const size_t o_ts0 = take(t_words * 4), o_ts1 = take(t_words * 4);
  const size_t o_ts2 = take(t_words * 4);
  const size_t o_chunk = take(((size_t)chunks + 1) * 8);   // the pair emission's chunk first pairs and total
  const size_t o_ranges = take((size_t)t_cap * 8);
  const size_t o_torder = take((size_t)t_cap * 4);
  const size_t total = off;
  if (!b) return total;
  w.base = b;
  w.bytes = total;
  w.slots_cap = s_cap;
  w.pairs_cap = p_cap;
  w.tiles_cap = t_cap;
  w.tile_bits = tile_bits;
  w.wide_words = wide ? wide_words : 0;
  w.splats = (Splat *)(b + o_splats);
  w.keys = (uint32_t *)(b + o_keys);
  w.keys_alt = (uint32_t *)(b + o_keys_alt);
  w.vals_alt = (uint32_t *)(b + o_vals_alt);
  w.order = (uint32_t *)(b + o_order);
  w.tbox = (uint2 *)(b + o_tbox);
  w.partials = (uint2 *)(b + o_partials);
  w.pk0 = (uint32_t *)(b + o_pk0);
  w.pv0 = (uint32_t *)(b + o_pv0);
  w.pk1 = (uint32_t *)(b + o_pk1);
  w.pv1 = (uint32_t *)(b + o_pv1);
  w.views = (ViewDev *)(b + o_views);
  w.slot_base = (uint32_t *)(b + o_views + sizeof(ViewDev) * kMaxViews);
  w.tile_base = w.slot_base + kMaxViews + 1;
  w.tiles_x = w.tile_base + kMaxViews + 1;
  w.view_counters = (ViewCounters *)(b + o_vcount);
  w.scratch = b + o_scratch;
  w.scratch_words = (total - o_scratch) / 4;
  w.counters = (FrameCounters *)(b + o_counters);
  w.tickets = (uint32_t *)(b + o_tickets);
  w.depth_hist = (uint32_t *)(b + o_dhist);
  w.tile_hist = (uint32_t *)(b + o_thist);
  w.count_hist = (uint32_t *)(b + o_chist);
  w.dstatus[0] = (uint32_t *)(b + o_ds0);
  w.dstatus[1] = (uint32_t *)(b + o_ds1);
  w.dstatus[2] = (uint32_t *)(b + o_ds2);
  w.tstatus[0] = (uint32_t *)(b + (wide ? o_wide0 : o_ts0));
  w.tstatus[1] = (uint32_t *)(b + (wide ? o_wide1 : o_ts1));
  w.tstatus[2] = wide ? nullptr : (uint32_t *)(b + o_ts2);
  w.chunk_status = (uint64_t *)(b + o_chunk);
  w.ranges = (uint2 *)(b + o_ranges);
  w.tile_order = (uint32_t *)(b + o_torder);
  return total;
}

void drop_graph(sl_context *ctx) {
  if (ctx->graph) {   // captured with the old buffers
    (void)hipGraphExecDestroy(ctx->graph);
    ctx->graph = nullptr;
    ctx->graph_key.clear();
  }
}

// Grow the context's own frame workspace to hold `slots` splat slots, `pairs` pairs and `tiles` tiles.
sl_status ensure_workspace(sl_context *ctx, uint32_t slots, uint32_t pairs, uint32_t tiles, uint32_t tile_bits) {
  Workspace &w = ctx->ws;
  if (w.base && slots <= w.slots_cap && pairs <= w.pairs_cap && tiles <= w.tiles_cap && tile_bits == w.tile_bits)
    return SL_OK;
  const uint32_t s_cap = std::max(slots, w.slots_cap);
  const uint32_t p_cap = std::max(pairs, w.pairs_cap);
  const uint32_t t_cap = std::max(tiles, w.tiles_cap);
  const size_t total = layout_workspace(nullptr, s_cap, p_cap, t_cap, tile_bits, w);
  if (ctx->ws_limit && total > ctx->ws_limit)
    return fail(ctx, SL_ERR_CAPACITY, "workspace of " + std::to_string(total) + " B over the " +
                                          std::to_string(ctx->ws_limit) + " B the server's HBM budget leaves this lane");
  if (w.base) {
    SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->geo) SL_HIP(ctx, hipStreamSynchronize(ctx->geo));
    if (ctx->comp) SL_HIP(ctx, hipStreamSynchronize(ctx->comp));
    SL_HIP(ctx, hipFree(w.base));
    w.base = nullptr;
  }
  drop_graph(ctx);
  void *base = nullptr;
  sl_status s = dev_alloc(ctx, total, &base);
  if (s != SL_OK) return s;
  layout_workspace((uint8_t *)base, s_cap, p_cap, t_cap, tile_bits, w);
  return SL_OK;
}

// RenderFrame (shaders/common.wgsl:44-77) -> FrameDev, refusing frames that turn on features outside this path.
sl_status decode_frame(sl_context *ctx, const sl_scene *sc, const void *fb, const sl_frame_opts *o, FrameDev *f,
                       uint32_t tile_shift) {
  memset(f, 0, sizeof(*f));
  for (int k = 0; k < 16; k++) f->view[k] = rd_f32(fb, k);
  for (int m = 0; m < 3; m++)
    for (int k = 0; k < 4; k++) f->palette[m][k] = rd_f32(fb, 32 + 4 * m + k);
  f->palette[3][0] = f->palette[3][1] = f->palette[3][2] = 1.0f;   // slot 3 is not a colour (project.wgsl:27-31)
  f->palette[3][3] = 0.0f;
  for (int m = 4; m < 8; m++)
    for (int k = 0; k < 4; k++) f->palette[m][k] = rd_f32(fb, 84 + 4 * (m - 4) + k);
  for (uint32_t p = 0; p < SL_MAX_PARTS; p++) {
    f->part_material[p] = p < 16 ? rd_u32(fb, 48 + (int)p) : rd_u32(fb, 164 + (int)(p - 16));
    const float fade = rd_f32(fb, 264 + (int)p);
    f->part_keep[p] = 1.0f - std::min(std::max(fade, 0.0f), 1.0f);
  }
  for (uint32_t m = 0; m < 32; m++) {
    const float wv = m < 16 ? rd_f32(fb, 100 + 4 * (int)m + 3) : rd_f32(fb, 196 + 4 * (int)(m - 16) + 3);
    if (((uint32_t)(wv + 0.5f)) & 1u) f->glass_mask |= 1u << m;
  }
  f->vw = rd_f32(fb, 72);
  f->vh = rd_f32(fb, 73);
  f->fx = rd_f32(fb, 74);
  f->fy = rd_f32(fb, 75);
  f->near_ = rd_f32(fb, 76);
  f->far_ = rd_f32(fb, 77);
  f->log_far_near = (float)log((double)(f->far_ / f->near_));
  f->count = rd_u32(fb, 78);
  f->view_mode = rd_u32(fb, 79);
  f->albedo_scale = rd_f32(fb, 80);
  f->lowpass = rd_f32(fb, 83);
  f->gain = rd_f32(fb, 67) > 0.5f ? 0.0f : rd_f32(fb, 44);       // lamp.w: flake colour already holds bounce
  const float n_emit_f = rd_f32(fb, 45);
  const float shutter = rd_f32(fb, 46);
  const uint32_t refl = (uint32_t)(rd_f32(fb, 71) + 0.5f);
  if (f->count != sc->count)
    return fail(ctx, SL_ERR_INVALID, "RenderFrame.count " + std::to_string(f->count) + " != scene count " +
                                         std::to_string(sc->count));
  if ((uint32_t)n_emit_f != sc->n_emit)
    return fail(ctx, SL_ERR_INVALID, "RenderFrame emitter count (palette[3].y) != scene emitters");
  if (f->view_mode > 3u) return fail(ctx, SL_ERR_INVALID, "view_mode must be 0-3");
  if (refl != 0u && f->view_mode == 3u && rd_f32(fb, 67) < 0.5f)
    return fail(ctx, SL_ERR_UNSUPPORTED, "reflection coverage / Fresnel split (lamp_color.w) is not on this path");
  if (rd_u32(fb, 180) != 0u) {   // screen-space LOD (project.wgsl:90-109): frame.lod = (threshold px, base, nodes)
    if (!sc->lod.on)
      return fail(ctx, SL_ERR_UNSUPPORTED, "frame.lod draws the tree's cut: give the scene its tree (sl_scene_set_lod)");
    const uint32_t n_nodes = rd_u32(fb, 182);
    if (n_nodes != sc->lod.N - sc->count)
      return fail(ctx, SL_ERR_INVALID, "frame.lod.z " + std::to_string(n_nodes) + " != the tree's internal nodes " +
                                           std::to_string(sc->lod.N - sc->count));
    if (rd_u32(fb, 181) < sc->count + 1u + sc->n_emit)
      return fail(ctx, SL_ERR_INVALID, "frame.lod.y (the first node slot) lies among the Gaussian or emitter slots");
    f->lod_threshold = rd_f32(fb, 180);
    f->n_nodes = n_nodes;
  }
  if (rd_f32(fb, 194) > 0.5f) return fail(ctx, SL_ERR_UNSUPPORTED, "explode is not on this path");
  if (rd_u32(fb, 296) != 0u) return fail(ctx, SL_ERR_UNSUPPORTED, "particles are not on this path");
  f->shutter = shutter > 0.0f ? shutter : 0.0f;   // moving parts (Part.flags bit 1) blur over it (project.wgsl:138)
  f->blur_cover = rd_f32(fb, 47) > 0.5f ? 1u : 0u;
  if (!(f->near_ > 0.0f) || !(f->far_ > f->near_)) return fail(ctx, SL_ERR_INVALID, "bad near / far");
  if ((uint32_t)f->vw != o->width || (uint32_t)f->vh != o->height || o->width == 0 || o->height == 0)
    return fail(ctx, SL_ERR_INVALID, "RenderFrame viewport does not match the output size");
  f->n_emit = sc->n_emit;
  f->nslots = sc->count + 1u + sc->n_emit + f->n_nodes;
  f->width = o->width;
  f->height = o->height;
  f->tile_shift = tile_shift;
  f->tiles_x = (o->width + (1u << tile_shift) - 1u) >> tile_shift;
  f->tiles_y = (o->height + (1u << tile_shift) - 1u) >> tile_shift;
  return SL_OK;
}

uint32_t grid_for(uint64_t n, uint32_t per_block, uint32_t cap) {
  uint64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return (uint32_t)std::min<uint64_t>(g, cap);
}

// The stage sequence of a batch: one launch per stage over every view. overlap (a batch encoded after a lighting
// encode): the geometry stages (projection without colour, sorts, pair emission, ranges) run on the geometry
// stream after the work queued before that lighting, beside it; the context's stream waits for them and for the
// lighting, then writes the splats' colours and colour max and composites. Otherwise every stage runs on the
// context's stream.
sl_status enqueue_batch(sl_context *ctx, const Batch &b, bool timers, bool overlap) {
  Workspace &w = b.fixed ? ctx->fws : ctx->ws;
  const hipStream_t ms = ctx->stream, st = overlap ? ctx->geo : ctx->stream;
  const uint32_t V = (uint32_t)b.views.size();
  auto mark = [&](int k, hipStream_t on) -> sl_status {
    if (timers) SL_HIP(ctx, hipEventRecord(ctx->ev[k], on));
    return SL_OK;
  };
  sl_status s;
  if ((s = mark(0, st)) != SL_OK) return s;
  sl_fill_u32<<<grid_for(w.scratch_words, 256, 2048), 256, 0, st>>>((uint32_t *)w.scratch, 0u,
                                                                    (uint32_t)w.scratch_words);
  // the LOD cut (lod_select) before the projection draws it; the node colours (lod_colour) once the leaves' lit
  // colours are final: here without overlap, after the lighting's wait with it
  const LodDev &L = b.lod_dev;
  uint32_t *lod_flag = (uint32_t *)ctx->lod_flag.p;
  auto lod_colour = [&](hipStream_t on) {
    const uint32_t ni = L.N - L.G;
    sl_fill_u32<<<grid_for(ni, 256, 2048), 256, 0, on>>>((uint32_t *)ctx->lod_cnt.p, 0u, ni);
    sl_lod_colour<<<(L.G + 255u) / 256u, 256, 0, on>>>(L, lod_flag, b.scenes[0]->colour, b.scenes[0]->lo_b,
                                                        b.views[0].f.gain, (uint32_t *)ctx->lod_cnt.p,
                                                        (float4 *)ctx->lod_col.p);
  };
  const bool lod_lit = b.lod && b.views[0].f.view_mode == 3u;
  if (b.lod) {
    sl_lod_open<<<(L.N - L.G + 255u) / 256u, 256, 0, st>>>(L, (uint8_t *)ctx->lod_open.p);
    sl_lod_flags<<<(L.N + 255u) / 256u, 256, 0, st>>>(L, (const uint8_t *)ctx->lod_open.p, lod_flag);
    if (lod_lit && !overlap) lod_colour(st);
  }
  if ((s = mark(1, st)) != SL_OK) return s;
  auto project = [&](auto kernel) {
    kernel<<<b.total_slots / 256u, 256, 0, st>>>(w.views, w.slot_base, V, w.splats, w.keys, w.tbox, w.partials);
  };
  if (b.blur)
    overlap ? project(sl_project_kernel<false, true>) : project(sl_project_kernel<true, true>);
  else
    overlap ? project(sl_project_kernel<false, false>) : project(sl_project_kernel<true, false>);
  sl_frame_totals<<<V, 256, 0, st>>>(w.partials, w.slot_base, w.view_counters, w.counters, overlap ? 1u : 3u);
  if ((s = mark(2, st)) != SL_OK) return s;
  // depth order: stable sort of (view, 16-bit depth) keys, values = global slot
  const uint32_t n = b.total_slots;
  const sl_tuning &tu = ctx->tuning;
  const uint32_t rs_tile = RS_THREADS * tu.radix_items;
  const uint32_t d_grid = std::min<uint32_t>((n + rs_tile - 1) / rs_tile, tu.grid_mul * ctx->cus);
  // one onesweep pass in `bits`-bit digits (8: the depth sort and 8-bit tile sorts; 11: the tile sort when tuned so)
  auto radix_pass = [&](const uint32_t *ki, const uint32_t *vi, uint32_t *ko, uint32_t *vo, const uint32_t *n_ptr,
                        uint32_t n_static, uint32_t shift, const uint32_t *bases, uint32_t *status, uint32_t *ticket,
                        uint32_t grid, uint32_t bits) {
    if (bits == 11u) {
      if (tu.radix_items == 8u)
        sl_radix_pass<8, 11><<<grid, RS_THREADS, 0, st>>>(ki, vi, ko, vo, n_ptr, n_static, shift, bases, status, ticket);
      else
        sl_radix_pass<16, 11><<<grid, RS_THREADS, 0, st>>>(ki, vi, ko, vo, n_ptr, n_static, shift, bases, status, ticket);
    } else if (tu.radix_items == 8u) {
      sl_radix_pass<8, 8><<<grid, RS_THREADS, 0, st>>>(ki, vi, ko, vo, n_ptr, n_static, shift, bases, status, ticket);
    } else {
      sl_radix_pass<16, 8><<<grid, RS_THREADS, 0, st>>>(ki, vi, ko, vo, n_ptr, n_static, shift, bases, status, ticket);
    }
  };
  if (b.counting) {     // one view: keys are 16-bit depth keys or INVALID
    sl_count_hist<<<(n + 255u) / 256u, 256, 0, st>>>(w.keys, n, w.count_hist);
    sl_count_scan<<<1, 1024, 0, st>>>(w.count_hist);
    sl_count_scatter<<<(n + 255u) / 256u, 256, 0, st>>>(w.keys, n, w.count_hist, w.order);
  } else {
  sl_radix_hist<8><<<grid_for(n, 256 * 16, 1024), RS_THREADS, 0, st>>>(w.keys, nullptr, n, 0u, b.depth_passes,
                                                                      w.depth_hist);
  sl_radix_bases<8><<<b.depth_passes, RS_THREADS, 0, st>>>(w.depth_hist);
  radix_pass(w.keys, nullptr, w.keys_alt, w.vals_alt, nullptr, n, 0u, w.depth_hist, w.dstatus[0], w.tickets + 0, d_grid,
             8u);
  if (b.depth_passes == 2) {
    radix_pass(w.keys_alt, w.vals_alt, nullptr, w.order, nullptr, n, 8u, w.depth_hist + 256, w.dstatus[1],
               w.tickets + 1, d_grid, 8u);
  } else {
    radix_pass(w.keys_alt, w.vals_alt, w.pk1, w.pv1, nullptr, n, 8u, w.depth_hist + 256, w.dstatus[1], w.tickets + 1,
               d_grid, 8u);
    radix_pass(w.pk1, w.pv1, nullptr, w.order, nullptr, n, 16u, w.depth_hist + 512, w.dstatus[2], w.tickets + 2,
               d_grid, 8u);
  }
  }
  if ((s = mark(3, st)) != SL_OK) return s;
  const uint32_t tb = b.tile_bits, tdig = 1u << tb;
  // pair counts per 1,024-rank chunk, their scan (the chunks' first pairs and the pair total), the pairs
  const uint32_t pe_chunks = std::max((n + PE_CHUNK - 1) / PE_CHUNK, 1u);
  auto pairs_emit = [&](auto kernel) {
    sl_pairs_count<false><<<pe_chunks, PE_THREADS, 0, st>>>(w.order, w.tbox, nullptr, w.counters, w.chunk_status);
    sl_pairs_scan<<<1, 256, 0, st>>>(w.counters, w.chunk_status, w.pairs_cap);
    kernel<<<pe_chunks, PE_THREADS, 0, st>>>(w.order, w.tbox, nullptr, w.slot_base, w.tile_base, w.tiles_x, V,
                                             b.tile_passes, w.pairs_cap, w.counters, w.pk0, w.pv0, w.chunk_status,
                                             w.tile_hist);
  };
  if (tb == 11u) pairs_emit(sl_pairs_emit<false, 11>); else pairs_emit(sl_pairs_emit<false, 8>);
  if ((s = mark(4, st)) != SL_OK) return s;
  const uint32_t p_grid = std::max(std::min<uint32_t>((w.pairs_cap + rs_tile - 1) / rs_tile, tu.grid_mul * ctx->cus), 1u);
  if (tb == 11u) sl_radix_bases<11><<<b.tile_passes, RS_THREADS, 0, st>>>(w.tile_hist);
  else sl_radix_bases<8><<<b.tile_passes, RS_THREADS, 0, st>>>(w.tile_hist);
  uint32_t *kin = w.pk0, *vin = w.pv0, *kout = w.pk1, *vout = w.pv1;
  for (uint32_t pass = 0; pass < b.tile_passes; pass++) {
    if (tb == 11u)   // the wide look-back array: only the onesweep tiles the pair count fills, zeroed here
      sl_fill_pairs_status<<<grid_for(w.wide_words, 256, 2048), 256, 0, st>>>(w.tstatus[pass], &w.counters->pairs,
                                                                              rs_tile, tdig, (uint32_t)w.wide_words);
    radix_pass(kin, vin, kout, vout, &w.counters->pairs, 0u, tb * pass, w.tile_hist + tdig * pass, w.tstatus[pass],
               w.tickets + 4 + pass, p_grid, tb);
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  ctx->final_keys = kin;
  ctx->final_vals = vin;
  if ((s = mark(5, st)) != SL_OK) return s;
  sl_tile_ranges<<<grid_for(w.pairs_cap, 256, 4096), 256, 0, st>>>(kin, w.counters, w.ranges);
  sl_tile_order<<<1, 256, 0, st>>>(w.ranges, b.total_tiles, w.tile_order);
  if (V > 1) sl_view_pairs<<<V, 256, 0, st>>>(w.ranges, w.tile_base, w.view_counters);   // one view: the total
  if ((s = mark(6, st)) != SL_OK) return s;
  // overlapped: the composite on its own stream once the colours are in the splats (the context's stream is then
  // free for the next lighting encode, which writes nothing the composite reads)
  hipStream_t cs = ms;
  if (overlap) {
    SL_HIP(ctx, hipEventRecord(ctx->geo_done, st));
    SL_HIP(ctx, hipStreamWaitEvent(ms, ctx->geo_done, 0));
    if (lod_lit) lod_colour(ms);
    sl_splat_colour<<<b.total_slots / 256u, 256, 0, ms>>>(w.views, w.slot_base, V, w.keys, w.splats, w.partials);
    sl_frame_totals<<<V, 256, 0, ms>>>(w.partials, w.slot_base, w.view_counters, w.counters, 2u);
    if (timers) SL_HIP(ctx, hipEventRecord(ctx->ev_comp, ms));
    cs = ctx->comp;
    SL_HIP(ctx, hipEventRecord(ctx->colours_done, ms));
    SL_HIP(ctx, hipStreamWaitEvent(cs, ctx->colours_done, 0));
  }
  ctx->timed_overlap = overlap;
  auto composite = [&](auto kernel, uint32_t threads) {
    kernel<<<b.total_tiles * (b.tile_shift == 5u ? 4u : 1u), threads, 0, cs>>>(w.views, w.tile_base, V, w.view_counters, w.splats, vin, w.ranges,
                                              w.tile_order);
  };
  const bool big = b.tile_shift == 5u;
  switch (tu.composite_batch) {
    case 64u:
      big ? composite(sl_composite<64, 5, 4>, 256u) : composite(sl_composite<64, 4, 4>, 256u);
      break;
    case 128u:
      big ? composite(sl_composite<128, 5, 4>, 256u) : composite(sl_composite<128, 4, 4>, 256u);
      break;
    default:
      big ? composite(sl_composite<256, 5, 4>, 256u) : composite(sl_composite<256, 4, 4>, 256u);
  }
  if ((s = mark(7, cs)) != SL_OK) return s;
  return check_launch(ctx, "frame kernels");
}

// Upload the view table, run the batch (replaying its graph when one was captured for the same layout), and
// queue the counters' read-back. After a lighting encode (and without a graph) the geometry stages run beside the
// lighting: the view table then goes up on the geometry stream, after the work queued before that lighting.
sl_status launch_batch(sl_context *ctx, const Batch &b) {
  Workspace &w = b.fixed ? ctx->fws : ctx->ws;
  hipStream_t st = ctx->stream;
  const uint32_t V = (uint32_t)b.views.size();
  const bool overlap = !b.graph && ctx->light_pending;
  ctx->light_pending = false;
  if (overlap) SL_HIP(ctx, hipStreamWaitEvent(ctx->geo, ctx->pre_light, 0));
  std::vector<uint8_t> blk(kViewBlock, 0);
  memcpy(blk.data(), b.views.data(), sizeof(ViewDev) * V);
  uint32_t *sb = (uint32_t *)(blk.data() + sizeof(ViewDev) * kMaxViews);
  memcpy(sb, b.slot_base.data(), 4 * (V + 1));
  memcpy(sb + kMaxViews + 1, b.tile_base.data(), 4 * (V + 1));
  memcpy(sb + 2 * (kMaxViews + 1), b.tiles_x.data(), 4 * V);
  SL_HIP(ctx, hipMemcpyAsync(w.views, blk.data(), blk.size(), hipMemcpyHostToDevice, overlap ? ctx->geo : st));
  const bool timers = (b.flags & SL_FRAME_TIMERS) != 0 && !b.graph;
  sl_status s;
  if (b.graph) {
    // what the captured launches depend on: the layout and the workspace (the view table is read at run time)
    std::vector<uint64_t> key = {(uint64_t)(uintptr_t)w.base, V, b.counting, b.blur, b.total_slots, b.total_tiles, w.pairs_cap,
                                 b.depth_passes, b.tile_passes, b.tile_bits, b.tile_shift,
                                 ctx->tuning.composite_batch, ctx->tuning.radix_items, ctx->tuning.grid_mul};
    if (!ctx->graph || key != ctx->graph_key) {
      if (ctx->graph) {
        (void)hipGraphExecDestroy(ctx->graph);
        ctx->graph = nullptr;
      }
      hipGraph_t g = nullptr;
      SL_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      s = enqueue_batch(ctx, b, false, false);
      hipError_t e = hipStreamEndCapture(st, &g);
      if (s != SL_OK) {
        if (g) (void)hipGraphDestroy(g);
        return s;
      }
      if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
      e = hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return hip_fail(ctx, e, "hipGraphInstantiate");
      ctx->graph_key = key;
    } else {
      // the same launches: the pair lists end in the buffers the capture left them in
      Workspace &ww = b.fixed ? ctx->fws : ctx->ws;
      ctx->final_keys = b.tile_passes % 2 ? ww.pk1 : ww.pk0;
      ctx->final_vals = b.tile_passes % 2 ? ww.pv1 : ww.pv0;
    }
    SL_HIP(ctx, hipGraphLaunch(ctx->graph, st));
  } else if ((s = enqueue_batch(ctx, b, timers, overlap)) != SL_OK) {
    return s;
  }
  const hipStream_t fin = overlap ? ctx->comp : st;   // the stream the batch ends on
  SL_HIP(ctx, hipMemcpyAsync(ctx->counters_host, w.counters, sizeof(FrameCounters), hipMemcpyDeviceToHost, fin));
  SL_HIP(ctx, hipMemcpyAsync(ctx->view_counters_host, w.view_counters, sizeof(ViewCounters) * V,
                             hipMemcpyDeviceToHost, fin));
  if (overlap) {
    SL_HIP(ctx, hipEventRecord(ctx->comp_done, fin));
    ctx->comp_pending = true;
  }
  ctx->comp_last = overlap;
  ctx->last_light_seq = ctx->light_seq;
  ctx->timers_valid = timers;
  return SL_OK;
}

std::mutex g_session_mu;
std::string g_session_err;

}  // namespace

// ---- session ------------------------------------------------------------------------------------------------

extern "C" sl_status sl_session_acquire(const char *name, int64_t *token) {
  std::lock_guard<std::mutex> lk(g_session_mu);
  if (!token) return SL_ERR_INVALID;
  *token = -1;
  const char *dir = getenv("XDG_RUNTIME_DIR");
  std::string path = std::string(dir && *dir ? dir : "/tmp") + "/" + (name && *name ? name : "gs-game-engine-gpu") +
                     ".lock";
  const int fd = open(path.c_str(), O_CREAT | O_RDWR | O_CLOEXEC, 0600);
  if (fd < 0) {
    g_session_err = "GPU session lock " + path + ": " + strerror(errno);
    return SL_ERR_INVALID;
  }
  if (flock(fd, LOCK_EX | LOCK_NB) != 0) {
    close(fd);
    g_session_err = "Another tab on this origin owns the GPU session. Close that experiment tab first.";
    return SL_ERR_BUSY;
  }
  *token = fd;
  g_session_err.clear();
  return SL_OK;
}

extern "C" sl_status sl_session_release(int64_t token) {
  if (token < 0) return SL_ERR_INVALID;
  flock((int)token, LOCK_UN);
  close((int)token);
  return SL_OK;
}

extern "C" const char *sl_session_error(void) { return g_session_err.c_str(); }

// ---- context ------------------------------------------------------------------------------------------------

extern "C" sl_status sl_ctx_create(int device_ordinal, sl_context **out) {
  if (!out) return SL_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SL_ERR_DEVICE;
  if (device_ordinal < 0 || device_ordinal >= n) return SL_ERR_INVALID;
  sl_context *ctx = new sl_context();
  ctx->device = device_ordinal;
  auto bail = [&](hipError_t e) { delete ctx; (void)e; return SL_ERR_DEVICE; };
  hipError_t e;
  if ((e = hipSetDevice(device_ordinal)) != hipSuccess) return bail(e);
  if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess) return bail(e);
  (void)hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, device_ordinal);
  if (ctx->cus <= 0) ctx->cus = 256;
  ctx->fixed_base = scratch_address(device_ordinal, &ctx->fixed_err);   // the device is current
  if ((e = hipHostMalloc((void **)&ctx->counters_host, sizeof(FrameCounters), hipHostMallocDefault)) != hipSuccess)
    return bail(e);
  memset(ctx->counters_host, 0, sizeof(FrameCounters));
  if ((e = hipHostMalloc((void **)&ctx->view_counters_host, sizeof(ViewCounters) * kMaxViews, hipHostMallocDefault)) !=
      hipSuccess)
    return bail(e);
  for (int k = 0; k <= kPasses; k++)
    if ((e = hipEventCreate(&ctx->ev[k])) != hipSuccess) return bail(e);
  if ((e = hipEventCreate(&ctx->ev_comp)) != hipSuccess) return bail(e);
  for (sl_context::LightTimers &t : ctx->lt) {
    for (hipEvent_t &x : t.lev)
      if ((e = hipEventCreate(&x)) != hipSuccess) return bail(e);
    for (hipEvent_t &x : t.mev)
      if ((e = hipEventCreate(&x)) != hipSuccess) return bail(e);
    for (hipEvent_t &x : t.sev)
      if ((e = hipEventCreate(&x)) != hipSuccess) return bail(e);
  }
  for (int k = 0; k < 4; k++)
    if ((e = hipEventCreate(&ctx->bev[k])) != hipSuccess) return bail(e);
  *out = ctx;
  return SL_OK;
}

extern "C" void sl_ctx_destroy(sl_context *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->side) (void)hipStreamSynchronize(ctx->side);
  if (ctx->geo) (void)hipStreamSynchronize(ctx->geo);
  if (ctx->comp) (void)hipStreamSynchronize(ctx->comp);
  if (ctx->ws.base) (void)hipFree(ctx->ws.base);
  for (int k = 0; k <= kPasses; k++)
    if (ctx->ev[k]) (void)hipEventDestroy(ctx->ev[k]);
  if (ctx->ev_comp) (void)hipEventDestroy(ctx->ev_comp);
  for (sl_context::LightTimers &t : ctx->lt)
    for (hipEvent_t x : {t.lev[0], t.lev[1], t.lev[2], t.lev[3], t.lev[4], t.mev[0], t.mev[1], t.mev[2], t.mev[3],
                         t.sev[0], t.sev[1]})
      if (x) (void)hipEventDestroy(x);
  for (int k = 0; k < 4; k++)
    if (ctx->bev[k]) (void)hipEventDestroy(ctx->bev[k]);
  if (ctx->counters_host) (void)hipHostFree(ctx->counters_host);
  if (ctx->view_counters_host) (void)hipHostFree(ctx->view_counters_host);
  if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
  if (ctx->side) (void)hipStreamSynchronize(ctx->side);
  for (sl_context::MapBufs &m : ctx->map)
    for (Grow *gb : {&m.rec, &m.tbox, &m.tmask, &m.k0, &m.v0, &m.k1, &m.v1, &m.scratch})
      if (gb->p) (void)hipFree(gb->p);
  if (ctx->map_stats.p) (void)hipFree(ctx->map_stats.p);
  for (Grow *gb : {&ctx->disp_q, &ctx->lod_flag, &ctx->lod_open, &ctx->lod_cnt, &ctx->lod_col})
    if (gb->p) (void)hipFree(gb->p);
  for (hipEvent_t e : {ctx->fork, ctx->join, ctx->cleared, ctx->pre_light, ctx->geo_done, ctx->colours_done,
                       ctx->comp_done})
    if (e) (void)hipEventDestroy(e);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->geo) (void)hipStreamDestroy(ctx->geo);
  if (ctx->comp) (void)hipStreamDestroy(ctx->comp);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char *sl_last_error(const sl_context *ctx) { return ctx ? ctx->err.c_str() : "no context"; }
extern "C" sl_status sl_ctx_status(const sl_context *ctx) { return ctx ? ctx->poisoned : SL_ERR_INVALID; }

extern "C" sl_status sl_sync(sl_context *ctx) {
  SL_LIVE(ctx);
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SL_OK;
}

// ---- buffers ------------------------------------------------------------------------------------------------

extern "C" sl_status sl_buf_alloc(sl_context *ctx, size_t bytes, void **dev) {
  SL_LIVE(ctx);
  if (!dev) return SL_ERR_INVALID;
  return dev_alloc(ctx, bytes, dev);
}

extern "C" sl_status sl_buf_free(sl_context *ctx, void *dev) {
  SL_LIVE(ctx);
  if (!dev) return SL_OK;
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  SL_HIP(ctx, hipFree(dev));
  return SL_OK;
}

extern "C" sl_status sl_buf_write(sl_context *ctx, void *dev, size_t offset, const void *host, size_t bytes) {
  SL_LIVE(ctx);
  if (!dev || (!host && bytes)) return SL_ERR_INVALID;
  SL_HIP(ctx, hipMemcpyAsync((char *)dev + offset, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));   // the caller may reuse `host` at once
  return SL_OK;
}

extern "C" sl_status sl_buf_read(sl_context *ctx, const void *dev, size_t offset, void *host, size_t bytes) {
  SL_LIVE(ctx);
  if (!dev || (!host && bytes)) return SL_ERR_INVALID;
  SL_HIP(ctx, hipMemcpyAsync(host, (const char *)dev + offset, bytes, hipMemcpyDeviceToHost, ctx->stream));
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SL_OK;
}

// ---- scene --------------------------------------------------------------------------------------------------

namespace {

// A Morton order of the Gaussians within each part (rest-frame means quantised to 16 bits per axis over the
// part's box, part index above them), for the per-Gaussian lighting kernels: neighbouring lanes then light
// neighbouring Gaussians and walk similar emitter cuts. Only the order of the work changes, never a result.
std::vector<uint32_t> spatial_order(const void *gaussians, uint32_t n) {
  float lo[SL_MAX_PARTS][3], hi[SL_MAX_PARTS][3];
  for (uint32_t p = 0; p < SL_MAX_PARTS; p++)
    for (int a = 0; a < 3; a++) { lo[p][a] = 3.4e38f; hi[p][a] = -3.4e38f; }
  auto part_of = [&](const uint8_t *r) { return std::min(rd_u32(r, 11) & 0xffffu, SL_MAX_PARTS - 1u); };
  for (uint32_t i = 0; i < n; i++) {
    const uint8_t *r = (const uint8_t *)gaussians + 64 * (size_t)i;
    const uint32_t p = part_of(r);
    for (int a = 0; a < 3; a++) {
      const float v = rd_f32(r, a);
      if (v == v) { lo[p][a] = std::min(lo[p][a], v); hi[p][a] = std::max(hi[p][a], v); }
    }
  }
  auto spread = [](uint64_t x) {   // 16 bits -> every third of 48
    x &= 0xffffull;
    x = (x | (x << 16)) & 0x0000ff0000ffull;
    x = (x | (x << 8)) & 0x00f00f00f00full;
    x = (x | (x << 4)) & 0x0c30c30c30c3ull;
    x = (x | (x << 2)) & 0x249249249249ull;
    return x;
  };
  std::vector<std::pair<uint64_t, uint32_t>> key(n);
  for (uint32_t i = 0; i < n; i++) {
    const uint8_t *r = (const uint8_t *)gaussians + 64 * (size_t)i;
    const uint32_t p = part_of(r);
    uint64_t m = 0;
    for (int a = 0; a < 3; a++) {
      const float ext = hi[p][a] - lo[p][a];
      float t = ext > 0.0f ? (rd_f32(r, a) - lo[p][a]) / ext : 0.0f;
      t = t == t ? std::min(std::max(t, 0.0f), 1.0f) : 0.0f;
      m |= spread((uint64_t)(t * 65535.0f)) << a;
    }
    key[i] = {((uint64_t)p << 48) | m, i};
  }
  std::sort(key.begin(), key.end());
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; i++) order[i] = key[i].second;
  return order;
}

// f32 -> 16-bit storage, round to nearest even (NaN stays NaN, overflow goes to infinity)
uint16_t to_half_bits(float f) {
  const _Float16 h = (_Float16)f;
  uint16_t b;
  memcpy(&b, &h, 2);
  return b;
}
uint16_t to_bf16_bits(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  return (uint16_t)((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
}

// F16 / BF16 scenes: g0 = (mu, kappa) f32; g1 = {quat halves, albedo halves}; g2 = {scale halves, part_surface,
// flags}; packed on the host from the 64 B records (words: mu 0-2, kappa 3, quat 4-7, scale 8-10, part_surface 11,
// albedo 12-14, flags 15).
// The packed streams' host words of n records (the layout above).
void pack_streams(const void *aos, uint32_t n, uint32_t storage, std::vector<float4> &g0, std::vector<uint4> &h1,
                  std::vector<uint4> &h2) {
  auto cv = [&](const uint8_t *r, int word) -> uint32_t {
    const float f = rd_f32(r, word);
    return storage == SL_STORAGE_BF16 ? to_bf16_bits(f) : to_half_bits(f);
  };
  g0.resize(n);
  h1.resize(n);
  h2.resize(n);
  for (uint32_t i = 0; i < n; i++) {
    const uint8_t *r = (const uint8_t *)aos + 64 * (size_t)i;
    g0[i] = make_float4(rd_f32(r, 0), rd_f32(r, 1), rd_f32(r, 2), rd_f32(r, 3));
    h1[i] = make_uint4(cv(r, 4) | cv(r, 5) << 16, cv(r, 6) | cv(r, 7) << 16, cv(r, 12) | cv(r, 13) << 16, cv(r, 14));
    h2[i] = make_uint4(cv(r, 8) | cv(r, 9) << 16, cv(r, 10), rd_u32(r, 11), rd_u32(r, 15));
  }
}

// n 64 B records into existing streams: the three packed ones, or the four float4 ones through a staging copy and
// sl_aos_to_soa. Synchronous.
sl_status fill_streams(sl_context *ctx, const void *aos, uint32_t n, uint32_t storage, float4 *const dst[4]) {
  if (n == 0) return SL_OK;
  if (storage != SL_STORAGE_F32) {
    std::vector<float4> g0;
    std::vector<uint4> h1, h2;
    pack_streams(aos, n, storage, g0, h1, h2);
    const void *src[3] = {g0.data(), h1.data(), h2.data()};
    for (int k = 0; k < 3; k++) SL_HIP(ctx, hipMemcpy(dst[k], src[k], (size_t)n * 16, hipMemcpyHostToDevice));
    return SL_OK;
  }
  void *staging = nullptr;
  sl_status s = dev_alloc(ctx, (size_t)n * 64, &staging);
  if (s != SL_OK) return s;
  hipError_t e = hipMemcpyAsync(staging, aos, (size_t)n * 64, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    sl_aos_to_soa<<<(n + 255) / 256, 256, 0, ctx->stream>>>((const float4 *)staging, dst[0], dst[1], dst[2], dst[3],
                                                            n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(staging);
  if (e != hipSuccess) return hip_fail(ctx, e, "scene upload");
  return SL_OK;
}

// Allocates the record streams of n records (three for F16 / BF16 storage, four for F32) and fills them.
sl_status upload_streams(sl_context *ctx, sl_scene *sc, const void *aos, uint32_t n, uint32_t storage, float4 *dst[4]) {
  const int k_n = storage == SL_STORAGE_F32 ? 4 : 3;
  for (int k = 0; k < 4; k++) {
    dst[k] = nullptr;
    if (k >= k_n) continue;
    void *p = nullptr;
    sl_status s = dev_alloc(ctx, (size_t)std::max(n, 1u) * 16, &p);
    if (s != SL_OK) return s;
    sc->allocs.push_back(p);
    sc->bytes += (size_t)std::max(n, 1u) * 16;
    dst[k] = (float4 *)p;
  }
  return fill_streams(ctx, aos, n, storage, dst);
}

sl_status scene_alloc(sl_context *ctx, sl_scene *sc, size_t bytes, void **p) {
  sl_status s = dev_alloc(ctx, bytes, p);
  if (s != SL_OK) return s;
  sc->allocs.push_back(*p);
  sc->bytes += bytes;
  return SL_OK;
}

void scene_free(sl_scene *sc) {
  for (void *p : sc->allocs) (void)hipFree(p);
  for (Grow *gb : {&sc->emit_nodes, &sc->lsm_acc, &sc->lsm_mip, &sc->lsm_dirty, &sc->sun_acc, &sc->sun_mip,
                   &sc->sun_desc, &sc->sun_layers, &sc->sun_dirty})
    if (gb->p) (void)hipFree(gb->p);
  delete sc;
}

sl_status grow(sl_context *ctx, sl_scene *sc, Grow *gb, size_t bytes) {
  if (gb->p && gb->bytes >= bytes) return SL_OK;
  if (gb->p) {
    SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->side) SL_HIP(ctx, hipStreamSynchronize(ctx->side));
    if (ctx->geo) SL_HIP(ctx, hipStreamSynchronize(ctx->geo));
    if (ctx->comp) SL_HIP(ctx, hipStreamSynchronize(ctx->comp));
    SL_HIP(ctx, hipFree(gb->p));
    sc->bytes -= gb->bytes;
    gb->p = nullptr;
    gb->bytes = 0;
  }
  sl_status s = dev_alloc(ctx, bytes, &gb->p);
  if (s != SL_OK) return s;
  gb->bytes = bytes;
  sc->bytes += bytes;
  return SL_OK;
}

}  // namespace

extern "C" sl_status sl_scene_create(sl_context *ctx, const sl_scene_desc *d, sl_scene **out) {
  SL_LIVE(ctx);
  if (!d || !out) return SL_ERR_INVALID;
  *out = nullptr;
  // the layout checks of checkHeader / checkBuffer (src/asset.js:18-38)
  if (!d->gaussians) return fail(ctx, SL_ERR_LAYOUT, "Asset has no gaussians buffer");
  if (d->gaussian_stride != kGaussianStride)
    return fail(ctx, SL_ERR_LAYOUT, "gaussians stride " + std::to_string(d->gaussian_stride) + ", renderer expects 64");
  if (d->gaussian_count == 0 || d->gaussian_count > kMaxCount)
    return fail(ctx, SL_ERR_LAYOUT, "Bad gaussians count " + std::to_string(d->gaussian_count));
  if (!d->parts) return fail(ctx, SL_ERR_LAYOUT, "Asset has no parts buffer");
  if (d->part_stride != kPartStride)
    return fail(ctx, SL_ERR_LAYOUT, "parts stride " + std::to_string(d->part_stride) + ", renderer expects 128");
  if (d->part_count == 0 || d->part_count > SL_MAX_PARTS)
    return fail(ctx, SL_ERR_LAYOUT, std::to_string(d->part_count) + " parts; the renderer handles 32");
  if (d->materials && d->material_stride != kMaterialStride)
    return fail(ctx, SL_ERR_LAYOUT, "materials stride " + std::to_string(d->material_stride) + ", renderer expects 32");
  if (d->materials && d->material_count > SL_MAX_MATERIALS)
    return fail(ctx, SL_ERR_CAPACITY, std::to_string(d->material_count) + " materials; the renderer handles " +
                                          std::to_string(SL_MAX_MATERIALS));
  if (d->nodes && (d->node_stride != kNodeStride || d->node_count < d->gaussian_count))
    return fail(ctx, SL_ERR_LAYOUT, "nodes must be 96 B records, one leaf per Gaussian first");
  if (d->vgrid && (d->vgrid_stride != kVgridStride || d->vgrid_count < d->gaussian_count))
    return fail(ctx, SL_ERR_LAYOUT, "vgrid must be 128 B records, one per Gaussian");
  if (d->emitter_count > 16384u || (d->emitter_count && !d->emitters))
    return fail(ctx, SL_ERR_LAYOUT, "Bad emitter count " + std::to_string(d->emitter_count));
  if (d->storage > SL_STORAGE_BF16) return fail(ctx, SL_ERR_INVALID, "unknown storage format");
  for (uint32_t i = 0; i < d->gaussian_count; i++) {
    const uint32_t part = rd_u32((const uint8_t *)d->gaussians + 64 * (size_t)i, 11) & 0xffffu;
    if (part >= d->part_count)
      return fail(ctx, SL_ERR_LAYOUT, "Gaussian " + std::to_string(i) + " names part " + std::to_string(part) +
                                          " of " + std::to_string(d->part_count));
  }
  SL_HIP(ctx, hipSetDevice(ctx->device));
  sl_scene *sc = new sl_scene();
  sc->count = d->gaussian_count;
  sc->part_count = d->part_count;
  sc->n_emit = d->emitter_count;
  sc->material_count = d->materials ? d->material_count : 0;
  sc->storage = d->storage;
  sl_status s = upload_streams(ctx, sc, d->gaussians, sc->count, sc->storage, sc->g);
  if (s == SL_OK) s = upload_streams(ctx, sc, d->emitters, sc->n_emit, SL_STORAGE_F32, sc->e);
  void *p = nullptr;
  if (s == SL_OK && (s = scene_alloc(ctx, sc, SL_MAX_PARTS * 128, &p)) == SL_OK) sc->parts = (float4 *)p;
  if (s == SL_OK && (s = scene_alloc(ctx, sc, (size_t)sc->count * 16, &p)) == SL_OK) sc->colour = (float4 *)p;
  if (s == SL_OK) {
    sc->parts_host.assign(SL_MAX_PARTS * 128, 0);
    memcpy(sc->parts_host.data(), d->parts, (size_t)sc->part_count * 128);
    hipError_t e = hipMemcpyAsync(sc->parts, sc->parts_host.data(), SL_MAX_PARTS * 128, hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(sc->colour, 0, (size_t)sc->count * 16, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) s = hip_fail(ctx, e, "scene upload");
  }
  // relighting records: nodes (leaves), vgrid (all 255 when absent), materials
  if (s == SL_OK && d->nodes) {
    if ((s = scene_alloc(ctx, sc, (size_t)sc->count * 96, &p)) == SL_OK) {
      sc->nodes = (float4 *)p;
      hipError_t e = hipMemcpyAsync(sc->nodes, d->nodes, (size_t)sc->count * 96, hipMemcpyHostToDevice, ctx->stream);
      if (e != hipSuccess) s = hip_fail(ctx, e, "node upload");
    }
    if (s == SL_OK && (s = scene_alloc(ctx, sc, (size_t)sc->count * 128, &p)) == SL_OK) {
      sc->vgrid = (uint8_t *)p;
      hipError_t e = d->vgrid ? hipMemcpyAsync(sc->vgrid, d->vgrid, (size_t)sc->count * 128, hipMemcpyHostToDevice,
                                               ctx->stream)
                              : hipMemsetAsync(sc->vgrid, 0xFF, (size_t)sc->count * 128, ctx->stream);
      if (e != hipSuccess) s = hip_fail(ctx, e, "vgrid upload");
    }
    if (s == SL_OK && (s = scene_alloc(ctx, sc, (size_t)sc->count * 4, &p)) == SL_OK) {
      sc->order = (uint32_t *)p;
      const std::vector<uint32_t> ord = spatial_order(d->gaussians, sc->count);
      hipError_t e = hipMemcpy(sc->order, ord.data(), (size_t)sc->count * 4, hipMemcpyHostToDevice);
      if (e != hipSuccess) s = hip_fail(ctx, e, "order upload");
    }
    if (s == SL_OK && (s = scene_alloc(ctx, sc, (size_t)sc->count * 4, &p)) == SL_OK) sc->iso_norm = (float *)p;
    if (s == SL_OK && (s = scene_alloc(ctx, sc, (size_t)sc->count * 4, &p)) == SL_OK) sc->lamp_T = (float *)p;
    if (s == SL_OK && (s = scene_alloc(ctx, sc, (size_t)sc->count * 16, &p)) == SL_OK) sc->env_T = (float4 *)p;
  }
  if (s == SL_OK && d->materials && d->material_count) {
    sc->materials_host.assign((const uint8_t *)d->materials, (const uint8_t *)d->materials + 32 * (size_t)d->material_count);
    if ((s = scene_alloc(ctx, sc, 32 * (size_t)d->material_count, &p)) == SL_OK) {
      sc->materials = (uint32_t *)p;
      sc->materials_cap = 32 * (size_t)d->material_count;
      hipError_t e = hipMemcpyAsync(sc->materials, sc->materials_host.data(), sc->materials_host.size(),
                                    hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) s = hip_fail(ctx, e, "material upload");
    }
  }
  if (s == SL_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) s = hip_fail(ctx, e, "scene upload");
  }
  if (s != SL_OK) {
    scene_free(sc);
    return s;
  }
  *out = sc;
  return SL_OK;
}

extern "C" void sl_scene_destroy(sl_context *ctx, sl_scene *scene) {
  if (!scene) return;
  if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx && ctx->side) (void)hipStreamSynchronize(ctx->side);
  if (ctx && ctx->geo) (void)hipStreamSynchronize(ctx->geo);
  if (ctx && ctx->comp) (void)hipStreamSynchronize(ctx->comp);
  if (ctx && ctx->has_last && std::find(ctx->last.scenes.begin(), ctx->last.scenes.end(), scene) != ctx->last.scenes.end())
    ctx->has_last = false;
  scene_free(scene);
}

extern "C" size_t sl_scene_device_bytes(const sl_scene *scene) { return scene ? scene->bytes : 0; }

extern "C" sl_status sl_scene_write_parts(sl_context *ctx, sl_scene *sc, const void *bytes, size_t n) {
  SL_LIVE(ctx);
  if (!sc || !bytes) return SL_ERR_INVALID;
  if (n != (size_t)sc->part_count * 128)
    return fail(ctx, SL_ERR_LAYOUT, "parts are " + std::to_string(n) + " B, expected " +
                                        std::to_string(sc->part_count) + " x 128 B");
  memcpy(sc->parts_host.data(), bytes, n);
  SL_HIP(ctx, hipMemcpyAsync(sc->parts, sc->parts_host.data(), n, hipMemcpyHostToDevice, ctx->stream));
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  sc->parts_version++;
  return SL_OK;
}

extern "C" sl_status sl_scene_write_colour(sl_context *ctx, sl_scene *sc, const float *rgba, size_t count) {
  SL_LIVE(ctx);
  if (!sc || !rgba || count != sc->count) return fail(ctx, SL_ERR_INVALID, "colour needs one vec4 per Gaussian");
  SL_HIP(ctx, hipMemcpyAsync(sc->colour, rgba, count * 16, hipMemcpyHostToDevice, ctx->stream));
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SL_OK;
}

extern "C" sl_status sl_scene_write_bounce(sl_context *ctx, sl_scene *sc, const float *lo_b, size_t count) {
  SL_LIVE(ctx);
  if (!sc) return SL_ERR_INVALID;
  if (!lo_b) {
    sc->lo_b = nullptr;   // the allocation (if any) stays with the scene
    return SL_OK;
  }
  if (count != sc->count) return fail(ctx, SL_ERR_INVALID, "bounce needs 3 vec4 per Gaussian");
  void *p = nullptr;
  sl_status s = scene_alloc(ctx, sc, count * 48, &p);
  if (s != SL_OK) return s;
  SL_HIP(ctx, hipMemcpyAsync(p, lo_b, count * 48, hipMemcpyHostToDevice, ctx->stream));
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  sc->lo_b = (float4 *)p;
  return SL_OK;
}

namespace {
// Waits for every stream of the context, so a rewrite of a scene buffer never races a kernel still reading it.
sl_status quiesce(sl_context *ctx) {
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->side) SL_HIP(ctx, hipStreamSynchronize(ctx->side));
  if (ctx->geo) SL_HIP(ctx, hipStreamSynchronize(ctx->geo));
  if (ctx->comp) SL_HIP(ctx, hipStreamSynchronize(ctx->comp));
  return SL_OK;
}
}  // namespace

extern "C" sl_status sl_scene_write_gaussians(sl_context *ctx, sl_scene *sc, const void *bytes, size_t n) {
  SL_LIVE(ctx);
  if (!sc || !bytes) return SL_ERR_INVALID;
  if (n != (size_t)sc->count * 64)
    return fail(ctx, SL_ERR_LAYOUT, "gaussians are " + std::to_string(n) + " B, expected " +
                                        std::to_string(sc->count) + " x 64 B");
  for (uint32_t i = 0; i < sc->count; i++) {
    const uint32_t part = rd_u32((const uint8_t *)bytes + 64 * (size_t)i, 11) & 0xffffu;
    if (part >= sc->part_count)
      return fail(ctx, SL_ERR_LAYOUT, "Gaussian " + std::to_string(i) + " names part " + std::to_string(part) +
                                          " of " + std::to_string(sc->part_count));
  }
  SL_HIP(ctx, hipSetDevice(ctx->device));
  sl_status s = quiesce(ctx);
  if (s == SL_OK) s = fill_streams(ctx, bytes, sc->count, sc->storage, sc->g);
  if (s != SL_OK) return s;
  if (sc->order) {
    const std::vector<uint32_t> ord = spatial_order(bytes, sc->count);
    SL_HIP(ctx, hipMemcpy(sc->order, ord.data(), (size_t)sc->count * 4, hipMemcpyHostToDevice));
  }
  // what was formed from the records is formed again: iso_norm at the next lighting encode, the bounce's leaf
  // weights at the next bounce encode (as after a part-table write)
  sc->iso_done = false;
  sc->visit.valid = false;
  sc->parts_version++;
  return SL_OK;
}

extern "C" sl_status sl_scene_write_materials(sl_context *ctx, sl_scene *sc, const void *bytes, size_t n) {
  SL_LIVE(ctx);
  if (!sc || !bytes) return SL_ERR_INVALID;
  if (n == 0 || n % 32 != 0 || n / 32 > SL_MAX_MATERIALS)
    return fail(ctx, SL_ERR_LAYOUT, "materials are " + std::to_string(n) + " B, expected 1 to " +
                                        std::to_string(SL_MAX_MATERIALS) + " x 32 B");
  SL_HIP(ctx, hipSetDevice(ctx->device));
  sl_status s = quiesce(ctx);
  if (s != SL_OK) return s;
  if (!sc->materials || n > sc->materials_cap) {
    void *p = nullptr;
    if ((s = scene_alloc(ctx, sc, n, &p)) != SL_OK) return s;
    sc->materials = (uint32_t *)p;
    sc->materials_cap = n;
  }
  sc->materials_host.assign((const uint8_t *)bytes, (const uint8_t *)bytes + n);
  sc->material_count = (uint32_t)(n / 32);
  SL_HIP(ctx, hipMemcpy(sc->materials, sc->materials_host.data(), n, hipMemcpyHostToDevice));
  return SL_OK;
}

// ---- frame --------------------------------------------------------------------------------------------------

namespace {

// Decode, validate and lay out a batch; grow the workspace (fixed: lay it over the device's no-alloc scratch
// instead, allocating nothing); record it as the context's last batch and launch it.
sl_status encode_batch(sl_context *ctx, const sl_view *views, uint32_t n_views, uint32_t flags, bool fixed = false) {
  if (!views || n_views == 0) return fail(ctx, SL_ERR_INVALID, "no views");
  if (n_views > kMaxViews) return fail(ctx, SL_ERR_CAPACITY, "more than 256 views in one batch");
  if (fixed && !ctx->fixed_base) return fail(ctx, SL_ERR_DEVICE, "no-alloc scratch unavailable: " + ctx->fixed_err);
  Batch b;
  b.flags = flags;
  b.graph = !fixed && (flags & SL_BATCH_GRAPH) != 0;
  b.fixed = fixed;
  b.counting = (flags & SL_FRAME_COUNTING_SORT) != 0;
  // the tuned tile and tile-sort digit width (the no-alloc frame keeps 8-bit digits: its scratch holds no wide
  // look-back arrays)
  b.tile_shift = ctx->tuning.tile_px == 32u ? 5u : 4u;
  b.tile_bits = fixed ? 8u : ctx->tuning.radix_bits;
  if (b.counting && n_views > 1)
    return fail(ctx, SL_ERR_UNSUPPORTED, "the counting scatter orders one view's 16-bit depth keys, not a batch");
  b.views.resize(n_views);
  b.scenes.resize(n_views);
  b.slot_base.assign(n_views + 1, 0);
  b.tile_base.assign(n_views + 1, 0);
  b.tiles_x.assign(n_views, 0);
  uint64_t slots = 0, tiles = 0, want_pairs = 0;
  for (uint32_t v = 0; v < n_views; v++) {
    const sl_view &in = views[v];
    sl_scene *sc = in.scene;
    if (!sc || !in.frame_bytes || !in.out_device) return fail(ctx, SL_ERR_INVALID, "null argument");
    if (in.frame_len != kFrameBytes) return fail(ctx, SL_ERR_INVALID, "RenderFrame must be 1200 B");
    if (in.out_format > SL_OUT_RGBA16F) return fail(ctx, SL_ERR_INVALID, "unknown output format");
    const sl_frame_opts o = {in.width, in.height, in.out_format, 0u};
    ViewDev &vd = b.views[v];
    memset(&vd, 0, sizeof(vd));
    sl_status s = decode_frame(ctx, sc, in.frame_bytes, &o, &vd.f, b.tile_shift);
    if (s != SL_OK) return s;
    const uint32_t nt = vd.f.tiles_x * vd.f.tiles_y;
    if (nt > 65536u) return fail(ctx, SL_ERR_CAPACITY, "more than 65536 tiles in a view");
    vd.g0 = sc->g[0]; vd.g1 = sc->g[1]; vd.g2 = sc->g[2]; vd.g3 = sc->g[3];
    vd.e0 = sc->e[0]; vd.e1 = sc->e[1]; vd.e2 = sc->e[2]; vd.e3 = sc->e[3];
    vd.parts = sc->parts; vd.colour = sc->colour; vd.lo_b = sc->lo_b;
    vd.storage = sc->storage;
    if (vd.f.n_nodes) {   // the cut lives in the context's LOD buffers: one view at a time, with its own workspace
      if (n_views > 1 || fixed || (flags & SL_BATCH_GRAPH))
        return fail(ctx, SL_ERR_UNSUPPORTED, "a screen-space LOD frame renders alone (sl_frame_encode)");
      const uint32_t N = sc->lod.N, ni = N - sc->count;
      if ((s = grow_ctx(ctx, &ctx->lod_flag, 4 * (size_t)N)) != SL_OK) return s;
      if ((s = grow_ctx(ctx, &ctx->lod_open, ni)) != SL_OK) return s;
      if ((s = grow_ctx(ctx, &ctx->lod_cnt, 4 * (size_t)ni)) != SL_OK) return s;
      if ((s = grow_ctx(ctx, &ctx->lod_col, 16 * (size_t)ni)) != SL_OK) return s;
      vd.lod_flag = (const uint32_t *)ctx->lod_flag.p;
      vd.n0 = sc->lod.ng[0]; vd.n1 = sc->lod.ng[1]; vd.n2 = sc->lod.ng[2]; vd.n3 = sc->lod.ng[3];
      vd.node_colour = (const float4 *)ctx->lod_col.p;
      vd.node_slot0 = sc->count + 1u + sc->n_emit;
      vd.n_nodes = ni;
      b.lod = true;
      LodDev &L = b.lod_dev;
      L.na = sc->lod.na; L.nb = sc->lod.nb; L.nc = sc->lod.nc;
      L.up = sc->lod.up; L.part = sc->lod.part; L.child = sc->lod.child; L.kids = sc->lod.kids;
      L.mass = sc->lod.mass;
      L.parts = sc->parts;
      L.N = N;
      L.G = sc->count;
      memcpy(L.view, vd.f.view, sizeof(L.view));
      L.threshold = vd.f.lod_threshold;
      L.focal = vd.f.fx;
      L.near_ = vd.f.near_;
      L.emit_part = sc->lod.emit_part;
      L.open_parts = 0u;
      L.size_rule = sc->lod.size_rule;
    }
    vd.out = in.out_device;
    vd.out_format = in.out_format;
    vd.slot_base = (uint32_t)slots;
    vd.tile_base = (uint32_t)tiles;
    b.scenes[v] = sc;
    b.slot_base[v] = (uint32_t)slots;
    b.tile_base[v] = (uint32_t)tiles;
    b.tiles_x[v] = vd.f.tiles_x;
    b.blur = b.blur || vd.f.shutter > 0.0f;
    slots += ((uint64_t)vd.f.nslots + 255u) / 256u * 256u;
    tiles += nt;
    want_pairs += 8ull * vd.f.nslots;
  }
  if (slots >= (1ull << 31)) return fail(ctx, SL_ERR_CAPACITY, "more than 2^31 splat slots in one batch");
  if (tiles > (1ull << 24)) return fail(ctx, SL_ERR_CAPACITY, "more than 2^24 tiles in one batch");
  b.slot_base[n_views] = (uint32_t)slots;
  b.tile_base[n_views] = (uint32_t)tiles;
  b.total_slots = (uint32_t)slots;
  b.total_tiles = (uint32_t)tiles;
  b.depth_passes = n_views > 1 ? 3u : 2u;
  uint32_t tile_id_bits = 1;            // bits of the largest global tile id
  while (tile_id_bits < 24u && (tiles - 1u) >> tile_id_bits) tile_id_bits++;
  b.tile_passes = (tile_id_bits + b.tile_bits - 1u) / b.tile_bits;
  SL_HIP(ctx, hipSetDevice(ctx->device));
  if (fixed) {
    // this frame's slots and tiles, then as many pairs as the rest of the scratch holds (at least one per slot:
    // a batch's three-pass depth sort borrows the second pair buffers)
    const size_t cap = SL_NOALLOC_SCRATCH_BYTES;
    const size_t fixed_bytes = layout_workspace(nullptr, b.total_slots, b.total_slots, b.total_tiles, 8u, ctx->fws);
    if (fixed_bytes > cap)
      return fail(ctx, SL_ERR_CAPACITY, "no-alloc frame: " + std::to_string(b.total_slots) + " slots and " +
                                            std::to_string(b.total_tiles) + " tiles need " +
                                            std::to_string(fixed_bytes) + " B, more than the scratch's " +
                                            std::to_string(cap) + " B");
    uint32_t lo = b.total_slots, hi = (uint32_t)std::min<size_t>(cap / 16, 1u << 30);
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo + 1) / 2;
      if (layout_workspace(nullptr, b.total_slots, mid, b.total_tiles, 8u, ctx->fws) <= cap) lo = mid;
      else hi = mid - 1;
    }
    layout_workspace(ctx->fixed_base, b.total_slots, lo, b.total_tiles, 8u, ctx->fws);
    ctx->has_last = true;
    ctx->last = std::move(b);
    return launch_batch(ctx, ctx->last);
  }
  const uint32_t pairs = std::max<uint32_t>(ctx->ws.pairs_cap, (uint32_t)std::min<uint64_t>(want_pairs + 4096, 1u << 29));
  sl_status s;
  if ((s = ensure_workspace(ctx, b.total_slots, std::max(pairs, b.total_slots), b.total_tiles, b.tile_bits)) != SL_OK)
    return s;
  ctx->has_last = true;
  ctx->last = std::move(b);
  return launch_batch(ctx, ctx->last);
}

// Wait for the last batch; on pair overflow grow the pair buffers and run it again; per-view stats.
sl_status finish_batch(sl_context *ctx, sl_frame_stats *stats, uint32_t n_stats) {
  if (!ctx->has_last) return fail(ctx, SL_ERR_INVALID, "no frame encoded");
  // the batch ends on the composite's stream when it ran there: a lighting encoded since may still run on the
  // context's stream, and is not waited for
  SL_HIP(ctx, hipStreamSynchronize(ctx->comp_last ? ctx->comp : ctx->stream));
  FrameCounters c = *ctx->counters_host;
  const Batch &b = ctx->last;
  if (c.overflow > 0 && b.fixed) {
    if (stats) {
      memset(stats, 0, sizeof(*stats));
      stats->slots = b.views[0].f.nslots;
      stats->pairs = c.pairs;
      stats->overflow = c.overflow;
      stats->tiles = b.total_tiles;
    }
    return fail(ctx, SL_ERR_CAPACITY, "no-alloc frame: " + std::to_string((uint64_t)c.pairs + c.overflow) +
                                          " pairs do not fit the scratch's " + std::to_string(ctx->fws.pairs_cap) +
                                          " at this frame's slots and tiles (the image is incomplete)");
  }
  if (c.overflow > 0) {
    const uint64_t need = (uint64_t)c.pairs + c.overflow;
    const uint64_t grown = std::min<uint64_t>((need * 5) / 4 + 4096, 1u << 30);
    if (need > grown) return fail(ctx, SL_ERR_CAPACITY, "batch needs more than 2^30 pairs");
    sl_status s = join_composite(ctx);
    if (s != SL_OK) return s;
    s = ensure_workspace(ctx, b.total_slots, (uint32_t)grown, b.total_tiles, b.tile_bits);
    if (s != SL_OK) return s;
    if (ctx->light_seq != ctx->last_light_seq) {
      // a lighting encoded after this batch has rewritten the colours the batch was drawn with: the buffers are
      // grown for the next batch, but this one is not drawn again
      if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->slots = b.views[0].f.nslots;
        stats->pairs = c.pairs;
        stats->overflow = c.overflow;
        stats->tiles = b.total_tiles;
      }
      return fail(ctx, SL_ERR_CAPACITY, std::to_string(need) + " pairs overflowed the pair buffers and a lighting "
                                        "was encoded after the frame, so it is not re-run (the image is incomplete); "
                                        "the buffers are grown for the next frame");
    }
    if ((s = launch_batch(ctx, b)) != SL_OK) return s;
    SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    c = *ctx->counters_host;
    if (c.overflow > 0) return fail(ctx, SL_ERR_CAPACITY, "pair buffers still too small");
  }
  const uint32_t V = (uint32_t)b.views.size();
  if (stats)
    for (uint32_t v = 0; v < std::min(V, n_stats); v++) {
      const ViewCounters &vc = ctx->view_counters_host[v];
      memset(&stats[v], 0, sizeof(stats[v]));
      stats[v].slots = b.views[v].f.nslots;
      stats[v].visible = vc.visible;
      stats[v].pairs = V == 1 ? c.pairs : vc.pairs;
      stats[v].overflow = c.overflow;
      stats[v].tiles = b.views[v].f.tiles_x * b.views[v].f.tiles_y;
    }
  if (ctx->timers_valid) {
    for (int k = 0; k < kPasses; k++) {
      float ms = 0.0f;
      // overlapped: "splat" is the composite alone (its start waited for the lighting, then the colour pass)
      (void)hipEventElapsedTime(&ms, k == kPasses - 1 && ctx->timed_overlap ? ctx->ev_comp : ctx->ev[k], ctx->ev[k + 1]);
      ctx->ms[k] = ms;
    }
  }
  return SL_OK;
}

}  // namespace

extern "C" sl_status sl_frame_encode(sl_context *ctx, sl_scene *sc, const void *frame_bytes, size_t frame_len,
                                     const sl_frame_opts *opts, void *out_device) {
  SL_LIVE(ctx);
  if (!sc || !frame_bytes || !opts || !out_device) return fail(ctx, SL_ERR_INVALID, "null argument");
  const sl_view v = {sc, frame_bytes, frame_len, opts->width, opts->height, opts->out_format, out_device};
  return encode_batch(ctx, &v, 1, opts->flags & (SL_FRAME_TIMERS | SL_FRAME_COUNTING_SORT));
}

extern "C" sl_status sl_frame_encode_noalloc(sl_context *ctx, sl_scene *sc, const void *frame_bytes,
                                             size_t frame_len, const sl_frame_opts *opts, void *out_device) {
  SL_LIVE(ctx);
  if (!sc || !frame_bytes || !opts || !out_device) return fail(ctx, SL_ERR_INVALID, "null argument");
  const sl_view v = {sc, frame_bytes, frame_len, opts->width, opts->height, opts->out_format, out_device};
  return encode_batch(ctx, &v, 1, opts->flags & (SL_FRAME_TIMERS | SL_FRAME_COUNTING_SORT), true);
}

extern "C" sl_status sl_frame_finish(sl_context *ctx, sl_frame_stats *stats) {
  SL_LIVE(ctx);
  return finish_batch(ctx, stats, 1);
}

// ---- the real-asset path: screen-space LOD tree and display fill ----------------------------------------------

extern "C" sl_status sl_scene_set_lod(sl_context *ctx, sl_scene *sc, const sl_lod_tree *t) {
#pragma clang fp contract(off)
  SL_LIVE(ctx);
  if (!sc || !t || !t->nodes || !t->kids || !t->node_gaussians) return fail(ctx, SL_ERR_INVALID, "null argument");
  if (sc->lod.on) return fail(ctx, SL_ERR_INVALID, "the scene already holds its LOD tree");
  const uint32_t G = sc->count, N = t->node_count;
  if (t->node_stride != 96u) return fail(ctx, SL_ERR_LAYOUT, "nodes stride " + std::to_string(t->node_stride) +
                                                                 ", renderer expects 96");
  if (N <= G || N > (1u << 25)) return fail(ctx, SL_ERR_LAYOUT, "Bad nodes count " + std::to_string(N));
  if (t->node_gaussian_count != N - G)
    return fail(ctx, SL_ERR_LAYOUT, "node Gaussian records " + std::to_string(t->node_gaussian_count) +
                                        " != internal nodes " + std::to_string(N - G));
  if (t->size_rule > 1u) return fail(ctx, SL_ERR_INVALID, "size_rule is 0 (radius) or 1 (projected area)");
  const uint8_t *nd = (const uint8_t *)t->nodes;
  auto nw = [&](uint32_t n, int k) { return rd_u32(nd + 96 * (size_t)n, k); };
  auto nfl = [&](uint32_t n, int k) { return rd_f32(nd + 96 * (size_t)n, k); };
  const uint32_t ni = N - G, K = t->kid_count;
  std::vector<uint32_t> up(N, 0xFFFFFFFFu), part(N);
  std::vector<uint2> child(ni);
  std::vector<float4> na(ni), nb(ni), nc(ni);
  std::vector<float> mass(N);
  for (uint32_t n = 0; n < N; n++) {
    const uint32_t lo = nw(n, 19), hi = nw(n, 20), d = nw(n, 21);
    part[n] = nw(n, 15);
    if (d > 64u) return fail(ctx, SL_ERR_LAYOUT, "node " + std::to_string(n) + " deeper than 64");
    if (n < G) {
      if (hi != lo) return fail(ctx, SL_ERR_LAYOUT, "node " + std::to_string(n) + " < G has children: leaves first");
    } else {
      if (lo >= hi || hi > K) return fail(ctx, SL_ERR_LAYOUT, "node " + std::to_string(n) + " has bad child range");
      child[n - G] = make_uint2(lo, hi);
      for (uint32_t k = lo; k < hi; k++) {
        const uint32_t c = t->kids[k];
        if (c >= N || up[c] != 0xFFFFFFFFu || nw(c, 21) != d + 1u)
          return fail(ctx, SL_ERR_LAYOUT, "kids[" + std::to_string(k) + "] is not a child one depth below node " +
                                              std::to_string(n) + " with a single parent");
        up[c] = n;
      }
      na[n - G] = make_float4(nfl(n, 0), nfl(n, 1), nfl(n, 2), nfl(n, 3));
      nb[n - G] = make_float4(nfl(n, 4), nfl(n, 5), nfl(n, 6), nfl(n, 7));
      nc[n - G] = make_float4(nfl(n, 8), nfl(n, 9), nfl(n, 10), rd_f32(nd + 96 * (size_t)n, 15));   // part bits
    }
    // node_mass (lod.wgsl:61-65)
    const float ax = nfl(n, 4), ay = nfl(n, 5), az = nfl(n, 6), ox = nfl(n, 8), oy = nfl(n, 9), oz = nfl(n, 10);
    const float det_inv = ax * (ay * az - oz * oz) - ox * (ox * az - oz * oy) + oy * (ox * oz - ay * oy);
    mass[n] = nfl(n, 7) / sqrtf(std::max(det_inv, 1e-30f));
  }
  sl_status s;
  void *p = nullptr;
  auto put = [&](const void *src, size_t bytes, void **dst) -> sl_status {
    sl_status r = scene_alloc(ctx, sc, std::max<size_t>(bytes, 16), dst);
    if (r != SL_OK) return r;
    SL_HIP(ctx, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return SL_OK;
  };
  if ((s = put(na.data(), 16 * (size_t)ni, &p)) != SL_OK) return s;
  sc->lod.na = (float4 *)p;
  if ((s = put(nb.data(), 16 * (size_t)ni, &p)) != SL_OK) return s;
  sc->lod.nb = (float4 *)p;
  if ((s = put(nc.data(), 16 * (size_t)ni, &p)) != SL_OK) return s;
  sc->lod.nc = (float4 *)p;
  if ((s = put(child.data(), 8 * (size_t)ni, &p)) != SL_OK) return s;
  sc->lod.child = (uint2 *)p;
  if ((s = put(up.data(), 4 * (size_t)N, &p)) != SL_OK) return s;
  sc->lod.up = (uint32_t *)p;
  if ((s = put(part.data(), 4 * (size_t)N, &p)) != SL_OK) return s;
  sc->lod.part = (uint32_t *)p;
  if ((s = put(mass.data(), 4 * (size_t)N, &p)) != SL_OK) return s;
  sc->lod.mass = (float *)p;
  if ((s = put(t->kids, 4 * (size_t)K, &p)) != SL_OK) return s;
  sc->lod.kids = (uint32_t *)p;
  s = upload_streams(ctx, sc, t->node_gaussians, ni, sc->storage, sc->lod.ng);
  if (s != SL_OK) return s;
  sc->lod.N = N;
  sc->lod.emit_part = t->emit_part;
  sc->lod.size_rule = t->size_rule;
  sc->lod.on = true;
  return SL_OK;
}

extern "C" sl_status sl_scene_set_fill(sl_context *ctx, sl_scene *sc, const uint32_t *fill, size_t words) {
  SL_LIVE(ctx);
  if (!sc || !fill) return fail(ctx, SL_ERR_INVALID, "null argument");
  const uint32_t G = sc->count;
  if (words < (size_t)G + 1 || fill[0] != 0u) return fail(ctx, SL_ERR_LAYOUT, "fill offsets must start at 0");
  for (uint32_t g = 0; g < G; g++)
    if (fill[g + 1] < fill[g]) return fail(ctx, SL_ERR_LAYOUT, "fill offsets are not ascending");
  const size_t pairs = fill[G];
  if (words != (size_t)G + 1 + 2 * pairs)
    return fail(ctx, SL_ERR_LAYOUT, "fill table has " + std::to_string(words) + " words, its offsets span " +
                                        std::to_string(pairs) + " pairs");
  std::vector<uint8_t> target(G, 0);
  for (uint32_t g = 0; g < G; g++) target[g] = fill[g + 1] > fill[g];
  for (size_t k = 0; k < pairs; k++) {
    const uint32_t j = fill[G + 1 + 2 * k];
    if (j >= G || target[j]) return fail(ctx, SL_ERR_LAYOUT, "fill source " + std::to_string(j) +
                                                                 " is out of range or itself a target");
  }
  void *p = nullptr;
  sl_status s = scene_alloc(ctx, sc, 4 * words, &p);
  if (s != SL_OK) return s;
  SL_HIP(ctx, hipMemcpy(p, fill, 4 * words, hipMemcpyHostToDevice));
  sc->fill = (uint32_t *)p;
  return SL_OK;
}

extern "C" sl_status sl_fill_encode(sl_context *ctx, sl_scene *sc) {
  SL_LIVE(ctx);
  if (!sc) return fail(ctx, SL_ERR_INVALID, "null argument");
  if (!sc->fill) return SL_OK;   // assets without a table skip it (src/lighting.js:994-995)
  sl_display_fill<<<(sc->count + 255u) / 256u, 256, 0, ctx->stream>>>(sc->fill, sc->count, sc->colour, sc->lo_b);
  return check_launch(ctx, "sl_display_fill");
}

// ---- display: the HDR target to 8-bit display colour (shaders/display.wgsl) ------------------------------------

namespace {
// The background sky's coefficient table (DP_SH_ENTRIES float4, the layout sl_display reads): per band l and
// order m >= 0 at e = l (l + 1) / 2 + m, {sh[l*l + l + m].rgb, f} and at 45 + e {sh[l*l + l - m].rgb, 0}, where f
// is the recurrence's factor (display.wgsl:49-59): kf = (l - m)! / (l + m)! / (4 pi) built up in the shader's float
// steps and order, k = sqrt((2l + 1) kf), times 1.41421356 for m > 0. bg: the SkyBg words. Returns lmax.
uint32_t display_sh_table(uint32_t n, const float *bg, float4 *tab) {
#pragma clang fp contract(off)
  const uint32_t lmax = (uint32_t)roundf(sqrtf((float)n)) - 1u;
  for (uint32_t m = 0; m <= lmax; m++) {
    float kf = (float)(1.0 / (4.0 * 3.14159265));
    for (uint32_t j = 1; j <= 2u * m; j++) kf /= (float)j;
    for (uint32_t l = m; l <= lmax; l++) {
      if (l > m + 1u) kf *= (float)(l - m) / (float)(l + m);
      if (l == m + 1u) kf *= 1.0f / (float)(2u * m + 1u);
      const float k = sqrtf((float)(2u * l + 1u) * kf);
      const uint32_t e = l * (l + 1u) / 2u + m, ip = l * l + l + m, im = l * l + l - m;
      tab[e] = make_float4(bg[4 * ip], bg[4 * ip + 1], bg[4 * ip + 2], m == 0u ? k : 1.41421356f * k);
      tab[45u + e] = make_float4(bg[4 * im], bg[4 * im + 1], bg[4 * im + 2], 0.0f);
    }
  }
  return lmax;
}
}  // namespace

extern "C" sl_status sl_display_encode(sl_context *ctx, const void *frame_bytes, size_t frame_len, const void *env,
                                       size_t env_len, const void *sky_bg, size_t sky_bg_len,
                                       const sl_display_opts *opts, const void *hdr_device, void *out_device) {
  SL_LIVE(ctx);
  if (!frame_bytes || !opts || !hdr_device || !out_device) return fail(ctx, SL_ERR_INVALID, "null argument");
  if (frame_len != 1200u) return fail(ctx, SL_ERR_LAYOUT, "RenderFrame must be 1200 bytes");
  if (env && env_len != 416u) return fail(ctx, SL_ERR_LAYOUT, "Env must be 416 bytes");
  if (sky_bg && sky_bg_len != 1520u) return fail(ctx, SL_ERR_LAYOUT, "SkyBg must be 1520 bytes");
  if (opts->hdr_format != SL_OUT_RGBA32F && opts->hdr_format != SL_OUT_RGBA16F)
    return fail(ctx, SL_ERR_INVALID, "hdr_format: SL_OUT_RGBA32F or SL_OUT_RGBA16F");
  if (opts->out_format != SL_DISPLAY_RGBA8 && opts->out_format != SL_DISPLAY_BGRA8)
    return fail(ctx, SL_ERR_INVALID, "out_format: SL_DISPLAY_RGBA8 or SL_DISPLAY_BGRA8");
  if (opts->flags & ~SL_DISPLAY_EXACT) return fail(ctx, SL_ERR_INVALID, "unknown display flags");
  const float *fr = (const float *)frame_bytes;
  const uint32_t *fu = (const uint32_t *)frame_bytes;
  const uint32_t W = opts->width, H = opts->height;
  if (W == 0u || H == 0u || W > 16384u || H > 16384u)
    return fail(ctx, SL_ERR_INVALID, "display size must be 1..16384 on each side");
  if ((float)W != fr[72] || (float)H != fr[73])
    return fail(ctx, SL_ERR_INVALID, "display size differs from the RenderFrame viewport");
  DisplayArgs a;
  memset(&a, 0, sizeof(a));
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) a.rot[3 * c + r] = fr[4 * c + r];
  a.vw = fr[72];
  a.vh = fr[73];
  a.fx = fr[74];
  a.fy = fr[75];
  a.view_mode = fu[79];
  a.exposure = fr[82];
  a.tone = fu[260];
  a.width = W;
  a.height = H;
  float bgw[380] = {};
  if (env) memcpy(a.env, env, 416);
  if (sky_bg) memcpy(bgw, sky_bg, 1520);
  memcpy(a.bg_tail, bgw + 324, sizeof(a.bg_tail));
  float4 tab[DP_SH_ENTRIES] = {};
  const uint32_t nsh = std::min(((const uint32_t *)bgw)[4 * 81], 81u);
  if (nsh > 0u) a.lmax = display_sh_table(nsh, bgw, tab);
  const uint32_t nr = std::min(((const uint32_t *)bgw)[4 * 81 + 1], 8u);
  for (uint32_t k = 0; k < nr; k++) {   // the lobe directions in float as the shader forms them, rounded once
    const float ang = 6.28318531f * (float)k / (float)nr;
    a.ring_cs[2 * k] = (float)cos((double)ang);
    a.ring_cs[2 * k + 1] = (float)sin((double)ang);
  }
  const uint64_t npx = (uint64_t)W * H;
  const uint32_t blocks = (uint32_t)((npx + DP_PX - 1) / DP_PX);
  const bool half_in = opts->hdr_format == SL_OUT_RGBA16F, bgra = opts->out_format == SL_DISPLAY_BGRA8;
  const bool exact = (opts->flags & SL_DISPLAY_EXACT) != 0u;
  // the context's display block: the regions' count words, the coefficient table at 256 B, the lists at 2 KiB
  const uint32_t q_cap = DP_PX * ((blocks + DP_REGIONS - 1u) / DP_REGIONS);
  sl_status s = grow_ctx(ctx, &ctx->disp_q, 2048 + 4 * (size_t)q_cap * DP_REGIONS);
  if (s != SL_OK) return s;
  uint32_t *q_count = (uint32_t *)ctx->disp_q.p, *q_list = q_count + 512;
  float4 *sh_table = (float4 *)(q_count + 64);
  uchar4 *out = (uchar4 *)out_device;
  hipStream_t st = ctx->stream;
  SL_HIP(ctx, hipMemcpyAsync(sh_table, tab, sizeof(tab), hipMemcpyHostToDevice, st));
  if (!exact) sl_fill_u32<<<1, 256, 0, st>>>(q_count, 0u, DP_REGIONS);
  const dim3 settle(std::max(1u, std::min(q_cap / DP_THREADS, 8u)), DP_REGIONS);
#define SL_DP_LAUNCH(H_, B_)                                                                                  \
  do {                                                                                                       \
    if (exact) {                                                                                             \
      sl_display<H_, B_, true><<<blocks, DP_THREADS, 0, st>>>(hdr_device, out, sh_table, q_count, q_list, q_cap, \
                                                              a);                                           \
    } else {                                                                                                 \
      sl_display<H_, B_, false><<<blocks, DP_THREADS, 0, st>>>(hdr_device, out, sh_table, q_count, q_list,   \
                                                               q_cap, a);                                    \
      sl_display_settle<H_, B_><<<settle, DP_THREADS, 0, st>>>(hdr_device, out, sh_table, q_count, q_list,   \
                                                               q_cap, a);                                    \
    }                                                                                                        \
  } while (0)
  if (half_in) {
    if (bgra) SL_DP_LAUNCH(true, true);
    else SL_DP_LAUNCH(true, false);
  } else {
    if (bgra) SL_DP_LAUNCH(false, true);
    else SL_DP_LAUNCH(false, false);
  }
#undef SL_DP_LAUNCH
  return check_launch(ctx, "sl_display");
}

extern "C" sl_status sl_ctx_set_tuning(sl_context *ctx, const sl_tuning *t) {
  SL_LIVE(ctx);
  if (!t) return fail(ctx, SL_ERR_INVALID, "null argument");
  sl_tuning u = *t;
  u.radix_bits = u.radix_bits ? u.radix_bits : 8u;
  u.tile_px = u.tile_px ? u.tile_px : 16u;
  u.pad = 0u;
  if ((u.composite_batch != 64u && u.composite_batch != 128u && u.composite_batch != 256u) ||
      (u.radix_items != 8u && u.radix_items != 16u) || u.grid_mul < 1u || u.grid_mul > 16u ||
      (u.radix_bits != 8u && u.radix_bits != 11u) || (u.tile_px != 16u && u.tile_px != 32u))
    return fail(ctx, SL_ERR_INVALID, "tuning: composite_batch 64/128/256, radix_items 8/16, grid_mul 1-16, "
                                     "radix_bits 8/11, tile_px 16/32");
  ctx->tuning = u;   // a workspace laid out for other digits is laid out again at the next encode
  return SL_OK;
}

extern "C" sl_status sl_ctx_get_tuning(sl_context *ctx, sl_tuning *t) {
  SL_LIVE(ctx);
  if (!t) return fail(ctx, SL_ERR_INVALID, "null argument");
  *t = ctx->tuning;
  return SL_OK;
}

extern "C" sl_status sl_batch_encode(sl_context *ctx, const sl_view *views, uint32_t n_views, uint32_t flags) {
  SL_LIVE(ctx);
  return encode_batch(ctx, views, n_views, flags);
}

extern "C" sl_status sl_batch_finish(sl_context *ctx, sl_frame_stats *stats, uint32_t n_stats) {
  SL_LIVE(ctx);
  return finish_batch(ctx, stats, n_stats);
}

extern "C" sl_status sl_frame_read(sl_context *ctx, uint32_t which, void *host, size_t bytes, size_t *written) {
  SL_LIVE(ctx);
  if (!ctx->has_last) return fail(ctx, SL_ERR_INVALID, "no frame encoded");
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const FrameCounters &c = *ctx->counters_host;
  const Batch &b = ctx->last;
  const Workspace &w = b.fixed ? ctx->fws : ctx->ws;
  const void *src = nullptr;
  size_t n = 0;
  // a single frame: its own slots; a batch: the global slot space (views at their slot_base)
  const size_t slots = b.views.size() == 1 ? b.views[0].f.nslots : b.total_slots;
  switch (which) {
    case SL_ART_KEYS: src = w.keys; n = slots * 4; break;
    case SL_ART_ORDER: src = w.order; n = slots * 4; break;
    case SL_ART_SPLATS: src = w.splats; n = slots * sizeof(Splat); break;
    case SL_ART_PAIR_TILES: src = ctx->final_keys; n = (size_t)c.pairs * 4; break;
    case SL_ART_PAIR_SLOTS: src = ctx->final_vals; n = (size_t)c.pairs * 4; break;
    case SL_ART_RANGES: src = w.ranges; n = (size_t)b.total_tiles * 8; break;
    case SL_ART_LOD_FLAGS:
      if (!b.lod) return fail(ctx, SL_ERR_INVALID, "the last frame drew no LOD cut");
      src = ctx->lod_flag.p;
      n = (size_t)b.lod_dev.N * 4;
      break;
    case SL_ART_LOD_COLOUR:
      if (!b.lod) return fail(ctx, SL_ERR_INVALID, "the last frame drew no LOD cut");
      src = ctx->lod_col.p;
      n = (size_t)(b.lod_dev.N - b.lod_dev.G) * 16;
      break;
    default: return fail(ctx, SL_ERR_INVALID, "unknown artefact");
  }
  if (written) *written = n;
  if (!host) return SL_OK;
  if (bytes < n) return fail(ctx, SL_ERR_INVALID, "host buffer too small");
  if (n) {
    SL_HIP(ctx, hipMemcpyAsync(host, src, n, hipMemcpyDeviceToHost, ctx->stream));
    SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return SL_OK;
}

// ---- relighting ---------------------------------------------------------------------------------------------

extern "C" sl_status sl_lighting_iso_init(sl_context *ctx, sl_scene *sc) {
  SL_LIVE(ctx);
  if (!sc) return SL_ERR_INVALID;
  if (!sc->nodes) return fail(ctx, SL_ERR_INVALID, "relighting needs the scene's leaf nodes");
  SL_HIP(ctx, hipSetDevice(ctx->device));
  sl_iso_init<<<(sc->count + 63) / 64, 64, 0, ctx->stream>>>(sc->count, sc->g[2], sc->storage, sc->nodes, sc->vgrid,
                                                               sc->iso_norm);
  sl_status s = check_launch(ctx, "iso_init");
  if (s != SL_OK) return s;
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  sc->iso_done = true;
  sc->visit.valid = false;
  return SL_OK;
}

namespace {

template <class T>
void read_struct(const void *src, T *dst) { memcpy(dst, src, sizeof(T)); }

// every mip level of a layer lies inside it, level 0 first (the resolve writes it at the layer's start)
bool levels_fit(uint32_t res, uint32_t levels, const uint32_t *level_off, uint32_t layer_texels) {
  if (res < 1 || res > 16384 || (res & (res - 1)) != 0 || levels < 1 || levels > 16 || level_off[0] != 0) return false;
  for (uint32_t k = 0; k < levels; k++) {
    const uint32_t rk = res >> k;
    if (rk < 1) return false;
    if ((uint64_t)level_off[k] + (uint64_t)rk * rk > layer_texels) return false;
  }
  return true;
}

// The largest stack sl_shade_emit's cut walk can hold for this emitter tree: the walk that expands every internal
// node (any cut holds a subset of its pending entries when it reaches a node), with the kernel's 48-entry cap.
// A tree that does not terminate within 4n + 8 pops (a cycle) gets the full 48.
uint32_t emit_stack_need(const uint8_t *nodes, uint32_t n, uint32_t root) {
  std::vector<uint32_t> st(1, std::min(root, 0xFFFFu));
  size_t most = 1, pops = 0;
  while (!st.empty()) {
    if (++pops > 4 * (size_t)n + 8) return 48u;
    const uint32_t ni = st.back();
    st.pop_back();
    if (ni >= n) continue;
    const uint32_t clo = rd_u32(nodes + 96 * (size_t)ni, 11), chi = rd_u32(nodes + 96 * (size_t)ni, 15);
    if (clo == NO_CHILD) continue;                       // a leaf
    if (st.size() + 2 <= 48) {
      st.push_back(std::min(clo, 0xFFFFu));
      st.push_back(std::min(chi, 0xFFFFu));
      most = std::max(most, st.size());
    }
  }
  return (uint32_t)most;
}

sl_status grow_ctx(sl_context *ctx, Grow *gb, size_t bytes) {
  if (gb->p && gb->bytes >= bytes) return SL_OK;
  if (gb->p) {
    SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->side) SL_HIP(ctx, hipStreamSynchronize(ctx->side));
    if (ctx->geo) SL_HIP(ctx, hipStreamSynchronize(ctx->geo));   // the sun maps' clears run there
    if (ctx->comp) SL_HIP(ctx, hipStreamSynchronize(ctx->comp));
    SL_HIP(ctx, hipFree(gb->p));
    gb->p = nullptr;
    gb->bytes = 0;
  }
  sl_status s = dev_alloc(ctx, bytes, &gb->p);
  if (s == SL_OK) gb->bytes = bytes;
  return s;
}

// The map scratch block (u32 words), cleared before every map splat: the pair emission's counters, tickets and
// tile-digit histograms, the one-view layout it reads (tiles_x, zero slot and tile bases), its chunk look-back
// words, and the onesweep look-back words of up to 3 passes.
struct MapScratch {
  size_t counters = 0, tickets = 64, hist = 80, tiles_x = 848, zeros = 849, chunk = 856, tstat[3] = {0, 0, 0};
  size_t touched = 0;   // one bit per map tile: the sums gave it records (cleared with the rest each binning)
  size_t words = 0, pass_words = 0;
  MapScratch(size_t n_rec, size_t cap, size_t n_tiles) {
    const size_t chunks = (n_rec + PE_CHUNK - 1) / PE_CHUNK;
    pass_words = (cap + 4095) / 4096 * 256;
    tstat[0] = align_up(chunk + 2 * (std::max<size_t>(chunks, 1) + 1), 64);
    tstat[1] = tstat[0] + pass_words;
    tstat[2] = tstat[1] + pass_words;
    touched = tstat[2] + pass_words;
    words = touched + (n_tiles + 31) / 32;
  }
};

// Buffers of map kind `which` for n_rec records, each touching at most span x span tiles. Called before any launch
// of the frame's lighting: growing synchronises both streams.
sl_status ensure_map(sl_context *ctx, uint32_t which, size_t n_rec, uint32_t span, size_t n_tiles) {
  sl_status s;
  const size_t cap = std::max<size_t>(n_rec * span * span, 1);
  if (cap >= (1ull << 32)) return fail(ctx, SL_ERR_CAPACITY, "opacity-map splat needs more than 2^32 pairs");
  const MapScratch ms(n_rec, cap, n_tiles);
  sl_context::MapBufs &m = ctx->map[which];
  if ((s = grow_ctx(ctx, &m.rec, std::max<size_t>(n_rec, 1) * sizeof(MapRec))) != SL_OK) return s;
  if ((s = grow_ctx(ctx, &m.tbox, std::max<size_t>(n_rec, 1) * 8)) != SL_OK) return s;
  if ((s = grow_ctx(ctx, &m.tmask, std::max<size_t>(n_rec, 1) * 8)) != SL_OK) return s;
  for (Grow *gb : {&m.k0, &m.v0, &m.k1, &m.v1})
    if ((s = grow_ctx(ctx, gb, 4 * cap)) != SL_OK) return s;
  return grow_ctx(ctx, &m.scratch, 4 * ms.words);
}

// Map splats in two steps, each on stream st. map_bin: the records and tile boxes written by a prep pass ->
// (tile, record) pairs (sl_pairs_emit over the layers' stacked tile grid) -> sorted by tile (onesweep). map_sum:
// per-tile sums into acc, whose dirty layers are zero. which: 0 emitter map, 1 sun maps (buffers, stats slot,
// timer events and the sum kernel's fixed point).
struct MapShape {
  uint32_t tiles_x, tpl, n_tiles, tile_passes;
  size_t cap;
  MapShape(size_t n_rec, uint32_t n_layers, uint32_t res, uint32_t span) {
    tiles_x = (res + MT - 1) / MT;
    tpl = tiles_x * tiles_x;
    n_tiles = n_layers * tpl;
    tile_passes = n_tiles > 65536u ? 3u : 2u;
    cap = std::max<size_t>(n_rec * span * span, 1);
  }
};

// The binning scratch of map kind `which` cleared, with the records as the "ranks" (counters->visible, word 0),
// the tile columns and the tile count set, on stream st.
void map_clear(sl_context *ctx, hipStream_t st, uint32_t which, size_t n_rec, uint32_t n_layers, uint32_t res,
               uint32_t span) {
  const MapShape sh(n_rec, n_layers, res, span);
  const MapScratch ms(n_rec, sh.cap, sh.n_tiles);
  uint32_t *w = (uint32_t *)ctx->map[which].scratch.p;
  uint32_t *stats = (uint32_t *)ctx->map_stats.p + 2 * which;
  static_assert(offsetof(FrameCounters, visible) == 0, "visible is the counters' first word");
  sl_fill_u32_set<<<grid_for(ms.words, 256, 2048), 256, 0, st>>>(w, 0u, (uint32_t)ms.words, (uint32_t)ms.counters,
                                                                 (uint32_t)n_rec, (uint32_t)ms.tiles_x, sh.tiles_x,
                                                                 stats + 1, 0u);
}

// Binning of map kind `which` (its scratch cleared here unless `cleared`: map_clear already queued before it).
void map_bin(sl_context *ctx, hipStream_t st, uint32_t which, size_t n_rec, uint32_t n_layers, uint32_t res,
             uint32_t span, bool cleared = false) {
  sl_context::MapBufs &m = ctx->map[which];
  const MapShape sh(n_rec, n_layers, res, span);
  const MapScratch ms(n_rec, sh.cap, sh.n_tiles);
  uint32_t *w = (uint32_t *)m.scratch.p;
  FrameCounters *counters = (FrameCounters *)(w + ms.counters);
  if (!cleared) map_clear(ctx, st, which, n_rec, n_layers, res, span);
  uint32_t *k0 = (uint32_t *)m.k0.p, *v0 = (uint32_t *)m.v0.p;
  uint32_t *k1 = (uint32_t *)m.k1.p, *v1 = (uint32_t *)m.v1.p;
  const uint32_t chunks = (uint32_t)std::max<size_t>((n_rec + PE_CHUNK - 1) / PE_CHUNK, 1);
  uint64_t *chunk_pairs = (uint64_t *)(w + ms.chunk);
  sl_pairs_count<true><<<chunks, PE_THREADS, 0, st>>>(nullptr, (const uint2 *)m.tbox.p, (const uint64_t *)m.tmask.p,
                                                      counters, chunk_pairs);
  sl_pairs_scan<<<1, 256, 0, st>>>(counters, chunk_pairs, (uint32_t)sh.cap);
  sl_pairs_emit<true, 8><<<chunks, PE_THREADS, 0, st>>>(
      nullptr, (const uint2 *)m.tbox.p, (const uint64_t *)m.tmask.p, w + ms.zeros, w + ms.zeros, w + ms.tiles_x, 1u, sh.tile_passes,
      (uint32_t)sh.cap, counters, k0, v0, chunk_pairs, w + ms.hist);
  sl_radix_bases<8><<<sh.tile_passes, RS_THREADS, 0, st>>>(w + ms.hist);
  const uint32_t grid = std::max(std::min<uint32_t>((uint32_t)((sh.cap + 4095) / 4096), 4u * ctx->cus), 1u);
  uint32_t *kin = k0, *vin = v0, *kout = k1, *vout = v1;
  for (uint32_t p = 0; p < sh.tile_passes; p++) {
    sl_radix_pass<16, 8><<<grid, RS_THREADS, 0, st>>>(kin, vin, kout, vout, &counters->pairs, 0u, 8u * p,
                                                   w + ms.hist + 256 * p, w + ms.tstat[p], w + ms.tickets + 1 + p);
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
}

void map_sum(sl_context *ctx, hipStream_t st, uint32_t which, size_t n_rec, uint32_t n_layers, uint32_t res,
             uint32_t span, uint32_t *acc, bool timers) {
  sl_context::MapBufs &m = ctx->map[which];
  const MapShape sh(n_rec, n_layers, res, span);
  const MapScratch ms(n_rec, sh.cap, sh.n_tiles);
  uint32_t *w = (uint32_t *)m.scratch.p;
  const FrameCounters *counters = (const FrameCounters *)(w + ms.counters);
  uint32_t *stats = (uint32_t *)ctx->map_stats.p + 2 * which;
  // the sorted pairs are in (k0, v0) after an even number of passes, else in (k1, v1)
  const uint32_t *keys = (const uint32_t *)(sh.tile_passes % 2u == 0u ? m.k0.p : m.k1.p);
  const uint32_t *vals = (const uint32_t *)(sh.tile_passes % 2u == 0u ? m.v0.p : m.v1.p);
  if (timers) (void)hipEventRecord(ctx->lt[ctx->lt_cur].mev[2 * which], st);
  const uint32_t chunks = (uint32_t)((sh.cap + MAP_CHUNK - 1) / MAP_CHUNK);
  uint32_t *touched = w + ms.touched;
  if (which == 0)
    sl_map_accumulate<0><<<chunks, 256, 0, st>>>((const MapRec *)m.rec.p, keys, vals, counters, res, sh.tiles_x,
                                                 sh.tpl, acc, stats, touched);
  else
    sl_map_accumulate<1><<<chunks, 256, 0, st>>>((const MapRec *)m.rec.p, keys, vals, counters, res, sh.tiles_x,
                                                 sh.tpl, acc, stats, touched);
  ctx->map_touched[which] = touched;
  if (timers) (void)hipEventRecord(ctx->lt[ctx->lt_cur].mev[2 * which + 1], st);
  ctx->lt[ctx->lt_cur].map_timed[which] = timers;
  ctx->map_records[which] = (uint32_t)n_rec;
  ctx->map_tiles[which] = sh.n_tiles;
}

// tiles a box of at most 2 * max_half + 1 texels can span per axis
uint32_t map_span(uint32_t max_half, uint32_t res) {
  return std::min<uint32_t>((2u * max_half) / MT + 2u, (res + MT - 1) / MT);
}

void run_opacity_maps(hipStream_t st, float4 *mip, uint32_t n_layers, uint32_t res, uint32_t levels,
                      const uint32_t *level_off, uint32_t layer_texels, float fix, const uint32_t *acc,
                      const uint32_t *dirty_dev, const uint32_t *touched) {
  LevelOffs lo = {};
  for (uint32_t k = 0; k < levels && k < 16u; k++) lo.off[k] = level_off[k];
  if (res % 32u == 0u) {
    // levels 0-5 per 32x32 region, then full-size downs while a level is wider than 32, then the rest per layer
    sl_opacity_pyramid<<<dim3((res / 32u) * (res / 32u), n_layers), 256, 0, st>>>(acc, mip, res, levels, lo,
                                                                                  layer_texels, fix, dirty_dev,
                                                                                  touched);
    uint32_t k = 6;
    for (; k < levels && (res >> k) > 32u; k++) {
      const uint32_t rk = res >> k;
      sl_opacity_down<<<(uint32_t)(((size_t)rk * rk * n_layers + 255) / 256), 256, 0, st>>>(
          mip, n_layers, rk, level_off[k - 1], level_off[k], layer_texels, dirty_dev);
    }
    if (k < levels) sl_opacity_tail<<<n_layers, 256, 0, st>>>(mip, res, k, levels, lo, layer_texels, dirty_dev);
    return;
  }
  const size_t texels = (size_t)res * res;
  sl_opacity_resolve<<<(uint32_t)((texels * n_layers + 255) / 256), 256, 0, st>>>(acc, mip, n_layers, (uint32_t)texels,
                                                                                 layer_texels, fix, dirty_dev);
  for (uint32_t k = 1; k < levels; k++) {
    const uint32_t rk = res >> k;
    sl_opacity_down<<<(uint32_t)(((size_t)rk * rk * n_layers + 255) / 256), 256, 0, st>>>(
        mip, n_layers, rk, level_off[k - 1], level_off[k], layer_texels, dirty_dev);
  }
}

}  // namespace

namespace {
// The visiting-order copies (sl_scene::Visit), allocated on first use and formed on the context's stream.
sl_status visit_build(sl_context *ctx, sl_scene *sc) {
  sl_scene::Visit &v = sc->visit;
  const uint32_t n = sc->count;
  void *p = nullptr;
  sl_status s = SL_OK;
  for (int k = 0; k < 4 && s == SL_OK; k++)
    if (sc->g[k] && !v.g[k] && (s = scene_alloc(ctx, sc, (size_t)n * 16, &p)) == SL_OK) v.g[k] = (float4 *)p;
  if (s == SL_OK && !v.nodes && (s = scene_alloc(ctx, sc, (size_t)n * 96, &p)) == SL_OK) v.nodes = (float4 *)p;
  if (s == SL_OK && !v.vgrid && (s = scene_alloc(ctx, sc, (size_t)n * 128, &p)) == SL_OK) v.vgrid = (uint8_t *)p;
  if (s == SL_OK && !v.iso && (s = scene_alloc(ctx, sc, (size_t)n * 4, &p)) == SL_OK) v.iso = (float *)p;
  if (s != SL_OK) return s;
  auto rows = [&](const void *src, void *dst, uint32_t words) {
    sl_gather_rows<<<grid_for((size_t)n * words, 256, 1u << 30), 256, 0, ctx->stream>>>(
        n, sc->order, (const uint4 *)src, (uint4 *)dst, words);
  };
  for (int k = 0; k < 4; k++)
    if (sc->g[k]) rows(sc->g[k], v.g[k], 1);
  rows(sc->nodes, v.nodes, 6);
  rows(sc->vgrid, v.vgrid, 8);
  sl_gather_f32<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, sc->order, sc->iso_norm, v.iso);
  if ((s = check_launch(ctx, "visiting-order copies")) != SL_OK) return s;
  v.valid = true;
  return SL_OK;
}
}  // namespace

extern "C" sl_status sl_lighting_encode(sl_context *ctx, sl_scene *sc, const sl_lighting_desc *d) {
  SL_LIVE_BESIDE(ctx);
  if (!sc || !d || !d->frame_uniforms || !d->emit_params || !d->emit_nodes)
    return fail(ctx, SL_ERR_INVALID, "null argument");
  if (!sc->nodes || !sc->materials) return fail(ctx, SL_ERR_INVALID, "relighting needs the scene's nodes and materials");
  FrameUniformsDev U;
  EmitParamsDev EP;
  read_struct(d->frame_uniforms, &U);
  read_struct(d->emit_params, &EP);
  if (U.gaussian_count != sc->count) return fail(ctx, SL_ERR_INVALID, "FrameUniforms.gaussian_count != scene count");
  if (U.material_count == 0 || U.material_count > sc->materials_host.size() / 32)
    return fail(ctx, SL_ERR_INVALID, "FrameUniforms.material_count does not match the scene's materials");
  if ((U.iso & 1u) == 0u || (U.iso & (8u | 16u)) != 0u)
    return fail(ctx, SL_ERR_UNSUPPORTED, "only the M39 transfer without transmission or reflection split is on this path");
  // shiny receivers (material flags bit 3, not emissive): the shading and env_direct take their specular variants
  bool shiny = false;
  for (uint32_t m = 0; m < U.material_count; m++) {
    const uint32_t fl = rd_u32(sc->materials_host.data() + 32 * (size_t)m, 5);
    shiny = shiny || ((fl & 1u) == 0u && (fl & 8u) != 0u);
  }
  if (shiny && (EP.n5 & 16u) != 0u)
    return fail(ctx, SL_ERR_UNSUPPORTED, "the spec cut (EmitParams.n5 bit 4) is not on this path");
  if (EP.lights != 0u) return fail(ctx, SL_ERR_UNSUPPORTED, "unified lights are not on this path");
  if (d->emit_node_count == 0 || EP.root >= d->emit_node_count || (EP.count && d->emit_node_count < 2 * EP.count - 1))
    return fail(ctx, SL_ERR_INVALID, "emitter tree does not match EmitParams");
  if (d->emit_node_count > 65535u) return fail(ctx, SL_ERR_CAPACITY, "emitter tree over 65535 nodes");
  SL_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  sl_status s;
  if (!sc->iso_done && (s = sl_lighting_iso_init(ctx, sc)) != SL_OK) return s;
  if (!sc->visit.valid && (s = visit_build(ctx, sc)) != SL_OK) return s;
  const sl_scene::Visit &vis = sc->visit;
  // Validate and size both shadow chains before the first launch (a grow synchronises and frees). The chains then
  // run at once: emitter map -> lamp_T -> shade on the context's stream, sun maps -> env_T on the side stream; they
  // meet before env_direct, the first kernel that reads both.
  if ((s = grow(ctx, sc, &sc->emit_nodes, (size_t)d->emit_node_count * 96)) != SL_OK) return s;
  const bool fresh_stats = !ctx->map_stats.p;
  if (fresh_stats && (s = grow_ctx(ctx, &ctx->map_stats, 16)) != SL_OK) return s;
  // emitter map plan
  LSParamsDev LS{};
  uint32_t ls_flags[32] = {};
  uint32_t ls_span = 0;
  if (d->ls_params) {
    read_struct(d->ls_params, &LS);
    if (LS.part_count != sc->part_count || !levels_fit(LS.res, LS.levels, LS.level_off, LS.layer_texels))
      return fail(ctx, SL_ERR_INVALID, "LSParams do not match the scene");
    const size_t texels = (size_t)LS.res * LS.res;
    if ((s = grow(ctx, sc, &sc->lsm_acc, texels * 12 * LS.part_count)) != SL_OK) return s;
    if ((s = grow(ctx, sc, &sc->lsm_mip, (size_t)LS.layer_texels * 16 * LS.part_count)) != SL_OK) return s;
    if ((s = grow(ctx, sc, &sc->lsm_dirty, 4 * 32)) != SL_OK) return s;
    if (sc->lsm_res != LS.res || sc->lsm_parts != LS.part_count) LS.dirty = (1u << LS.part_count) - 1u;
    for (uint32_t p = 0; p < 32; p++) ls_flags[p] = (LS.dirty >> p) & 1u;
    if (LS.dirty) {
      const uint32_t tiles_x = (LS.res + MT - 1) / MT;
      ls_span = map_span(48, LS.res);
      if ((uint64_t)LS.part_count * tiles_x > 65535u)
        return fail(ctx, SL_ERR_CAPACITY, "emitter map: parts x tile rows over 65535");
      if ((s = ensure_map(ctx, 0, sc->count, ls_span, (size_t)LS.part_count * tiles_x * tiles_x)) != SL_OK) return s;
    }
  }
  // sun map plan
  EnvDev ENV{};
  SunParamsDev SM{};
  std::vector<SunLayerDesc> desc;
  std::vector<uint32_t> sun_flags;
  bool have_sun = false, sun_any = false;
  uint32_t max_layers = 0, sun_span = 0;
  size_t sun_rec = 0;
  if (d->env) {
    read_struct(d->env, &ENV);
    if (d->sun_params && ENV.info[0] > 0u) {
      read_struct(d->sun_params, &SM);
      if (SM.n_layers != d->sun_layer_count || !d->sun_layers || SM.n_layers == 0 || SM.n_lights > 4 ||
          !levels_fit(SM.res, SM.levels, SM.level_off, SM.layer_texels))
        return fail(ctx, SL_ERR_INVALID, "SunParams / layer plan do not match");
      for (uint32_t p = 0; p < 32; p++)
        if ((uint64_t)SM.part_layers[2 * p] + SM.part_layers[2 * p + 1] > SM.n_layers)
          return fail(ctx, SL_ERR_INVALID, "SunParams part layer range outside the plan");
      desc.resize(SM.n_layers);
      memcpy(desc.data(), d->sun_layers, sizeof(SunLayerDesc) * SM.n_layers);
      for (const SunLayerDesc &ld : desc)
        if (ld.part >= sc->part_count || ld.light >= 4u) return fail(ctx, SL_ERR_INVALID, "bad sun layer");
      const size_t texels = (size_t)SM.res * SM.res;
      if ((s = grow(ctx, sc, &sc->sun_acc, texels * 12 * SM.n_layers)) != SL_OK) return s;
      if ((s = grow(ctx, sc, &sc->sun_mip, (size_t)SM.layer_texels * 16 * SM.n_layers)) != SL_OK) return s;
      if ((s = grow(ctx, sc, &sc->sun_desc, sizeof(SunLayerDesc) * SM.n_layers)) != SL_OK) return s;
      if ((s = grow(ctx, sc, &sc->sun_layers, sizeof(SunLayer) * SM.n_layers)) != SL_OK) return s;
      if ((s = grow(ctx, sc, &sc->sun_dirty, 4 * (size_t)SM.n_layers)) != SL_OK) return s;
      const bool fresh = sc->sun_res != SM.res || sc->sun_n != SM.n_layers;
      sun_flags.resize(SM.n_layers);
      for (uint32_t l = 0; l < SM.n_layers; l++) {
        if (fresh) desc[l].dirty = 1u;
        sun_flags[l] = desc[l].dirty ? 1u : 0u;
        sun_any = sun_any || sun_flags[l];
      }
      if (sun_any) {
        for (uint32_t p = 0; p < sc->part_count; p++) max_layers = std::max(max_layers, SM.part_layers[2 * p + 1]);
        const uint32_t tiles_x = (SM.res + MT - 1) / MT;
        sun_span = map_span(32, SM.res);
        sun_rec = (size_t)sc->count * max_layers;
        if ((uint64_t)SM.n_layers * tiles_x > 65535u)
          return fail(ctx, SL_ERR_CAPACITY, "sun maps: layers x tile rows over 65535");
        if ((s = ensure_map(ctx, 1, sun_rec, sun_span, (size_t)SM.n_layers * tiles_x * tiles_x)) != SL_OK) return s;
      }
      have_sun = true;
    }
  }
  if (!ctx->geo) {
    SL_HIP(ctx, hipStreamCreateWithFlags(&ctx->geo, hipStreamNonBlocking));
    SL_HIP(ctx, hipEventCreateWithFlags(&ctx->pre_light, hipEventDisableTiming));
    SL_HIP(ctx, hipEventCreateWithFlags(&ctx->geo_done, hipEventDisableTiming));
    SL_HIP(ctx, hipStreamCreateWithFlags(&ctx->comp, hipStreamNonBlocking));
    SL_HIP(ctx, hipEventCreateWithFlags(&ctx->colours_done, hipEventDisableTiming));
    SL_HIP(ctx, hipEventCreateWithFlags(&ctx->comp_done, hipEventDisableTiming));
  }
  if (have_sun && !ctx->side) {
    SL_HIP(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    SL_HIP(ctx, hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
    SL_HIP(ctx, hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming));
    SL_HIP(ctx, hipEventCreateWithFlags(&ctx->cleared, hipEventDisableTiming));
  }

  const uint32_t gblocks = (sc->count + 63) / 64;
  const bool timers = (d->flags & SL_LIGHT_TIMERS) != 0;
  ctx->lt_cur ^= 1u;                                    // this encode's timer set (the other may still be read)
  sl_context::LightTimers &LT = ctx->lt[ctx->lt_cur];
  LT.map_timed[0] = LT.map_timed[1] = false;
  LT.sun_timed = false;
  // the next batch's geometry may run from here on, beside this lighting
  SL_HIP(ctx, hipEventRecord(ctx->pre_light, st));
  ctx->light_pending = true;
  ctx->light_seq++;
  SL_HIP(ctx, hipMemcpyAsync(sc->emit_nodes.p, d->emit_nodes, (size_t)d->emit_node_count * 96, hipMemcpyHostToDevice, st));
  if (fresh_stats) sl_fill_u32<<<1, 256, 0, st>>>((uint32_t *)ctx->map_stats.p, 0u, 4u);
  if (timers) SL_HIP(ctx, hipEventRecord(LT.lev[0], st));
  // side stream: sun maps -> env_T, after everything already on the context's stream. Enqueued before the emitter
  // map, whose short kernels then fill in around the sun maps' binning and sums (measured faster than ordering the
  // two chains' sums or the shading against each other with events).
  if (have_sun) {
    hipStream_t ss = ctx->side;
    SL_HIP(ctx, hipEventRecord(ctx->fork, st));
    SL_HIP(ctx, hipStreamWaitEvent(ss, ctx->fork, 0));
    if (timers) SL_HIP(ctx, hipEventRecord(LT.sev[0], ss));
    const size_t texels = (size_t)SM.res * SM.res;
    sc->sun_res = SM.res;
    sc->sun_n = SM.n_layers;
    sc->sun_layer_texels = SM.layer_texels;
    SL_HIP(ctx, hipMemcpyAsync(sc->sun_desc.p, desc.data(), sizeof(SunLayerDesc) * SM.n_layers, hipMemcpyHostToDevice, ss));
    SL_HIP(ctx, hipMemcpyAsync(sc->sun_dirty.p, sun_flags.data(), 4 * (size_t)SM.n_layers, hipMemcpyHostToDevice, ss));
    // the dirty layers' sums and the binning scratch are cleared on the geometry stream (idle until the frame's
    // encode), beside the layers' set-up and records: after the fork, so after every reader of the last frame's
    // maps; the binning waits for them
    hipStream_t gs = ctx->geo;
    SL_HIP(ctx, hipStreamWaitEvent(gs, ctx->fork, 0));
    for (uint32_t l = 0; l < SM.n_layers;) {            // each run of dirty layers cleared by one launch
      if (!sun_flags[l]) { l++; continue; }
      uint32_t e = l + 1u;
      while (e < SM.n_layers && sun_flags[e] && texels * 3 * (e + 1u - l) < (1ull << 31)) e++;
      sl_fill_u32<<<grid_for(texels * 3 * (e - l), 256, 4096), 256, 0, gs>>>(
          (uint32_t *)sc->sun_acc.p + texels * 3 * l, 0u, (uint32_t)(texels * 3 * (e - l)));
      l = e;
    }
    if (sun_any) map_clear(ctx, gs, 1, sun_rec, SM.n_layers, SM.res, sun_span);
    SL_HIP(ctx, hipEventRecord(ctx->cleared, gs));
    sl_sun_frame<<<(SM.n_layers + 63) / 64, 64, 0, ss>>>(SM, (const SunLayerDesc *)sc->sun_desc.p, sc->parts,
                                                         (SunLayer *)sc->sun_layers.p);
    if (sun_any) {
      if (sun_rec)
        sl_sun_prep<<<gblocks, 64, 0, ss>>>(sc->count, SM, (const SunLayerDesc *)sc->sun_desc.p,
                                            (const SunLayer *)sc->sun_layers.p, vis.g[0], vis.g[1], vis.g[2],
                                            sc->storage, vis.nodes, sc->parts, max_layers,
                                            (MapRec *)ctx->map[1].rec.p, (uint2 *)ctx->map[1].tbox.p,
                                            (uint64_t *)ctx->map[1].tmask.p);
    }
    SL_HIP(ctx, hipStreamWaitEvent(ss, ctx->cleared, 0));
    if (sun_any) {
      map_bin(ctx, ss, 1, sun_rec, SM.n_layers, SM.res, sun_span, true);
      map_sum(ctx, ss, 1, sun_rec, SM.n_layers, SM.res, sun_span, (uint32_t *)sc->sun_acc.p, timers);
      run_opacity_maps(ss, (float4 *)sc->sun_mip.p, SM.n_layers, SM.res, SM.levels, SM.level_off, SM.layer_texels,
                       4096.0f, (const uint32_t *)sc->sun_acc.p, (const uint32_t *)sc->sun_dirty.p,
                       ctx->map_touched[1]);
    }
    sl_sun_lookup<<<gblocks, 64, 0, ss>>>(sc->count, SM, (const SunLayerDesc *)sc->sun_desc.p,
                                          (const SunLayer *)sc->sun_layers.p, sc->g[0], sc->g[2], sc->storage,
                                          sc->parts, (const float4 *)sc->sun_mip.p, sc->env_T);
    if (timers) SL_HIP(ctx, hipEventRecord(LT.sev[1], ss));
    SL_HIP(ctx, hipEventRecord(ctx->join, ss));
    LT.sun_timed = timers;
  }
  // context's stream: the emitter light's opacity map and its lookup -> lamp_T, then the emitter shading
  const float *lampT = nullptr;
  if (d->ls_params) {
    const size_t texels = (size_t)LS.res * LS.res;
    sc->lsm_res = LS.res;
    sc->lsm_parts = LS.part_count;
    sc->lsm_layer_texels = LS.layer_texels;
    SL_HIP(ctx, hipMemcpyAsync(sc->lsm_dirty.p, ls_flags, sizeof(ls_flags), hipMemcpyHostToDevice, st));
    if (LS.dirty) {
      for (uint32_t p = 0; p < LS.part_count;) {        // each run of dirty parts cleared by one launch
        if (!ls_flags[p]) { p++; continue; }
        uint32_t e = p + 1u;
        while (e < LS.part_count && ls_flags[e] && texels * 3 * (e + 1u - p) < (1ull << 31)) e++;
        sl_fill_u32<<<grid_for(texels * 3 * (e - p), 256, 4096), 256, 0, st>>>(
            (uint32_t *)sc->lsm_acc.p + texels * 3 * p, 0u, (uint32_t)(texels * 3 * (e - p)));
        p = e;
      }
      if (sc->count)
        sl_lsm_prep<<<gblocks, 64, 0, st>>>(sc->count, LS, vis.g[0], vis.g[1], vis.g[2], sc->storage, vis.nodes,
                                            sc->parts, (MapRec *)ctx->map[0].rec.p, (uint2 *)ctx->map[0].tbox.p,
                                            (uint64_t *)ctx->map[0].tmask.p);
      map_bin(ctx, st, 0, sc->count, LS.part_count, LS.res, ls_span);
      map_sum(ctx, st, 0, sc->count, LS.part_count, LS.res, ls_span, (uint32_t *)sc->lsm_acc.p, timers);
      run_opacity_maps(st, (float4 *)sc->lsm_mip.p, LS.part_count, LS.res, LS.levels, LS.level_off, LS.layer_texels,
                       65536.0f, (const uint32_t *)sc->lsm_acc.p, (const uint32_t *)sc->lsm_dirty.p,
                       ctx->map_touched[0]);
    }
    sl_lsm_lookup<<<gblocks, 64, 0, st>>>(sc->count, LS, sc->g[0], sc->g[2], sc->storage, sc->parts,
                                          (const float4 *)sc->lsm_mip.p, sc->lamp_T);
    lampT = sc->lamp_T;
  }
  if (timers) SL_HIP(ctx, hipEventRecord(LT.lev[1], st));
  // with a bounce set up, the shading also keeps each Gaussian's emitter light before its own emission (the light
  // direct_dirs_emit sends toward the probes) for the next sl_bounce_encode
  float4 *c0_out = sc->bounce.on ? sc->bounce.c0 : nullptr;
  const bool wave_walk = emit_stack_need((const uint8_t *)d->emit_nodes, d->emit_node_count, EP.root) <= 16u;
  auto shade = [&](auto kernel) {
    kernel<<<gblocks, 64, 0, st>>>(U, EP, sc->order, (const float4 *)sc->emit_nodes.p, d->emit_node_count, vis.g[0],
                                   vis.g[1], vis.g[2], vis.g[3], sc->storage, vis.nodes, vis.vgrid, sc->parts, lampT,
                                   sc->materials, vis.iso, sc->colour, c0_out);
  };
  if (wave_walk)
    shiny ? shade(sl_shade_emit_wave<true>) : shade(sl_shade_emit_wave<false>);
  else
    shiny ? shade(sl_shade_emit<48, true>) : shade(sl_shade_emit<48, false>);
  if (timers) SL_HIP(ctx, hipEventRecord(LT.lev[2], st));
  // distant lights: env_T (the side stream's, or 1 without sun maps), then env_direct
  if (d->env) {
    if (have_sun)
      SL_HIP(ctx, hipStreamWaitEvent(st, ctx->join, 0));
    else
      sl_fill_u32<<<grid_for((size_t)sc->count * 4, 256, 4096), 256, 0, st>>>((uint32_t *)sc->env_T, 0x3f800000u,
                                                                               sc->count * 4);
    if (timers) SL_HIP(ctx, hipEventRecord(LT.lev[3], st));
    auto env = shiny ? sl_env_direct<true> : sl_env_direct<false>;
    env<<<gblocks, 64, 0, st>>>(U, ENV, sc->g[0], sc->g[1], sc->g[2], sc->g[3], sc->storage, sc->nodes, sc->vgrid,
                                sc->parts, sc->materials, sc->iso_norm, sc->env_T, sc->colour);
  } else if (timers) {
    SL_HIP(ctx, hipEventRecord(LT.lev[3], st));
  }
  if (timers) SL_HIP(ctx, hipEventRecord(LT.lev[4], st));
  LT.valid = timers;
  sc->bounce.U = U;
  sc->bounce.ENV = ENV;
  sc->bounce.lit = true;
  return check_launch(ctx, "lighting kernels");
}

// ---- sky and bounce through probes (Lighting.initBounce / encodeBounce, src/lighting.js:723-840, 1020-1103) -----

namespace {

// Checks a ProbeParams block against the scene and tree: every probe index the passes form stays below the probe
// count (lattice, then the movable parts' 8 each, then the static volumes), the lattice has 2+ probes per axis,
// the gather blend is an EMA (0 < alpha <= 1) and near-field reflection is off. Returns the probe count or 0.
uint32_t check_probe_params(sl_context *ctx, const sl_scene *sc, const ProbeParamsDev &PP, uint32_t node_count) {
  auto bad = [&](const std::string &m) { fail(ctx, SL_ERR_INVALID, "ProbeParams: " + m); return 0u; };
  if (PP.node_count != node_count) return bad("node_count does not match the tree");
  if (PP.dims[0] < 2 || PP.dims[1] < 2 || PP.dims[2] < 2 || PP.dims[2] > 16 || (uint64_t)PP.dims[0] * PP.dims[1] > 1024)
    return bad("lattice dims");
  if (!(PP.alpha > 0.0f && PP.alpha <= 1.0f)) {
    fail(ctx, SL_ERR_UNSUPPORTED, "ProbeParams: only the EMA gather blend (0 < alpha <= 1) is on this path");
    return 0u;
  }
  if (PP.refl[0] != 0u) {
    fail(ctx, SL_ERR_UNSUPPORTED, "ProbeParams: near-field reflection is not on this path");
    return 0u;
  }
  if (PP.subset == 0 || PP.phase >= PP.subset) return bad("subset / phase");
  const uint64_t lat = (uint64_t)PP.dims[0] * PP.dims[1] * PP.dims[2];
  if (PP.vol_info[0] > PP.part_probes || PP.vol_info[0] % 8u) return bad("part probes");
  for (uint32_t q = 0; q < 32; q++) {
    const uint32_t *v = PP.vols[q];
    if (v[1] == 0u) continue;
    if (v[1] < 2 || v[2] < 2 || v[3] < 2 || q >= sc->part_count) return bad("volume dims");
    if ((uint64_t)v[0] + (uint64_t)v[1] * v[2] * v[3] > PP.part_probes || v[0] < PP.vol_info[0]) return bad("volume range");
  }
  const uint64_t np = lat + PP.part_probes;
  if (np * 128u * 2u > (1ull << 31)) return bad("too many probes");
  return (uint32_t)np;
}

}  // namespace

extern "C" sl_status sl_bounce_init(sl_context *ctx, sl_scene *sc, const sl_bounce_tree *t, const void *probe_params,
                                    size_t probe_params_len) {
  SL_LIVE(ctx);
  if (!sc || !t || !t->nodes || !t->kids || !probe_params) return fail(ctx, SL_ERR_INVALID, "null argument");
  if (!sc->nodes) return fail(ctx, SL_ERR_INVALID, "the bounce needs the scene's relighting records");
  if (sc->bounce.on) return fail(ctx, SL_ERR_INVALID, "the scene's bounce is already set up");
  for (size_t m = 0; m + 32 <= sc->materials_host.size(); m += 32) {
    const uint32_t fl = rd_u32(sc->materials_host.data() + m, 5);
    if ((fl & 1u) == 0u && (fl & 8u) != 0u)
      return fail(ctx, SL_ERR_UNSUPPORTED, "the bounce toward shiny receivers (per-cell gloss, bounce.wgsl) is not "
                                           "on this path");
  }
  if (t->node_stride != 96u || t->node_count < sc->count || t->node_count >= (1u << 31))
    return fail(ctx, SL_ERR_LAYOUT, "tree nodes: 96 B records, leaves first, at least one per Gaussian");
  if (probe_params_len != sizeof(ProbeParamsDev)) return fail(ctx, SL_ERR_LAYOUT, "ProbeParams is 9632 B");
  const uint32_t N = t->node_count, G = sc->count;
  const uint8_t *nb = (const uint8_t *)t->nodes;
  for (uint32_t k = 0; k < t->kid_count; k++)
    if (t->kids[k] >= N) return fail(ctx, SL_ERR_INVALID, "kids: node index out of range");
  // node_up's lists: the internal nodes of each depth <= max_depth that belong to a part (src/lighting.js:762-772)
  std::vector<std::vector<uint32_t>> by(t->max_depth + 1);
  for (uint32_t i = G; i < N; i++) {
    const uint32_t clo = rd_u32(nb + 96 * (size_t)i, 19), chi = rd_u32(nb + 96 * (size_t)i, 20);
    const uint32_t part = rd_u32(nb + 96 * (size_t)i, 15), dep = rd_u32(nb + 96 * (size_t)i, 21);
    if (clo > chi || chi > t->kid_count) return fail(ctx, SL_ERR_INVALID, "tree node: child range outside kids");
    if (part != 0xFFFFFFFFu && dep <= t->max_depth) by[dep].push_back(i);
  }
  for (uint32_t q = 0; q < sc->part_count; q++) {
    const uint8_t *pq = sc->parts_host.data() + 128 * (size_t)q;
    if (rd_u32(pq, 30) > rd_u32(pq, 29) && rd_u32(pq, 28) >= N) return fail(ctx, SL_ERR_INVALID, "part root outside the tree");
  }
  ProbeParamsDev PP;
  read_struct(probe_params, &PP);
  const uint32_t NP = check_probe_params(ctx, sc, PP, N);
  if (!NP) return ctx->poisoned != SL_OK ? ctx->poisoned : SL_ERR_INVALID;
  // the gather's stack bound: a node popped with P entries on the stack (itself included) pushes its k kids,
  // holding P - 1 + k; its i-th kid (pushed i-th, popped after the later ones) is popped with P + i
  uint32_t walk_stack = 1;
  {
    std::vector<std::pair<uint32_t, uint32_t>> todo;   // (node, entries when popped)
    std::vector<uint8_t> seen(N, 0);
    for (uint32_t q = 0; q < sc->part_count; q++) {
      const uint8_t *pq = sc->parts_host.data() + 128 * (size_t)q;
      if (rd_u32(pq, 30) > rd_u32(pq, 29)) todo.push_back({rd_u32(pq, 28), 1u});
    }
    while (!todo.empty()) {
      const auto [i, P] = todo.back();
      todo.pop_back();
      if (i < G || seen[i]) continue;
      seen[i] = 1;
      const uint32_t clo = rd_u32(nb + 96 * (size_t)i, 19), chi = rd_u32(nb + 96 * (size_t)i, 20);
      walk_stack = std::max(walk_stack, P - 1u + (chi - clo));
      for (uint32_t j = clo; j < chi; j++) todo.push_back({t->kids[j], P + (j - clo)});
    }
  }
  // the gather's compact tree: each part's subtree breadth first from its root, so a node's children get
  // consecutive walk indices in kids order; 32 B a node: mean, radius, smax, first child, child count, node id
  std::vector<float> walk;
  std::vector<uint32_t> walk_root(SL_MAX_PARTS, 0u);
  {
    std::vector<uint32_t> order;   // walk index -> node id
    for (uint32_t q = 0; q < sc->part_count; q++) {
      const uint8_t *pq = sc->parts_host.data() + 128 * (size_t)q;
      if (rd_u32(pq, 30) <= rd_u32(pq, 29)) continue;
      walk_root[q] = (uint32_t)order.size();
      size_t head = order.size();
      order.push_back(rd_u32(pq, 28));
      while (head < order.size()) {
        const uint32_t i = order[head++];
        if (i < G) continue;
        for (uint32_t j = rd_u32(nb + 96 * (size_t)i, 19); j < rd_u32(nb + 96 * (size_t)i, 20); j++) order.push_back(t->kids[j]);
        if (order.size() > (size_t)8 * N) return fail(ctx, SL_ERR_INVALID, "tree: the parts' subtrees are not trees");
      }
    }
    walk.resize(order.size() * 8);
    for (size_t w = 0; w < order.size(); w++) {
      const uint32_t i = order[w];
      const uint8_t *r = nb + 96 * (size_t)i;
      float *o = walk.data() + 8 * w;
      o[0] = rd_f32(r, 0); o[1] = rd_f32(r, 1); o[2] = rd_f32(r, 2); o[3] = rd_f32(r, 3);
      o[4] = rd_f32(r, 11);
      const uint32_t nk = i >= G ? rd_u32(r, 20) - rd_u32(r, 19) : 0u;
      memcpy(o + 6, &nk, 4);
      memcpy(o + 7, &i, 4);
    }
    // first-child indices: replay the breadth-first assignment per part (a node's children follow every child
    // block handed out before it)
    for (uint32_t q = 0; q < sc->part_count; q++) {
      const uint8_t *pq = sc->parts_host.data() + 128 * (size_t)q;
      if (rd_u32(pq, 30) <= rd_u32(pq, 29)) continue;
      size_t end = walk_root[q] + 1;
      for (size_t w = walk_root[q]; w < end; w++) {
        uint32_t nk;
        memcpy(&nk, walk.data() + 8 * w + 6, 4);
        const uint32_t fc = (uint32_t)end;
        memcpy(walk.data() + 8 * w + 5, &fc, 4);
        end += nk;
      }
    }
  }
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  sl_scene::Bounce &B = sc->bounce;
  B.walk_stack = walk_stack;
  // node_up's entries per depth: (node, kid count, first kid, second kid), or (node, kid count, first index into
  // kids, 0) for more than two kids
  std::vector<uint32_t> up;
  B.lv_off.clear();
  B.lv_cnt.clear();
  for (uint32_t d = 0; d <= t->max_depth; d++) {
    B.lv_off.push_back((uint32_t)(up.size() / 4));
    B.lv_cnt.push_back((uint32_t)by[d].size());
    for (uint32_t i : by[d]) {
      const uint32_t clo = rd_u32(nb + 96 * (size_t)i, 19), nk = rd_u32(nb + 96 * (size_t)i, 20) - clo;
      const bool two = nk <= 2u;
      up.insert(up.end(), {i, nk, two && nk > 0u ? t->kids[clo] : clo, two && nk > 1u ? t->kids[clo + 1] : 0u});
    }
  }
  B.tail_hi = 0;
  while (B.tail_hi + 1u < t->max_depth && B.tail_hi + 1u <= UP_TAIL_DEPTHS && B.lv_cnt[B.tail_hi + 1u] <= UP_TAIL_MAX) {
    B.tail_hi++;
    B.tail.off[B.tail_hi] = B.lv_off[B.tail_hi];
    B.tail.cnt[B.tail_hi] = B.lv_cnt[B.tail_hi];
  }
  // the 128 cells' directions and solid angles, in double rounded once (bounce.wgsl cell_dir / cell_omega)
  float cells[128][4];
  for (int c = 0; c < 128; c++) {
    const int i = c / 16, j = c % 16;
    const double PI = 3.14159265358979323846;
    const double el = 0.5 * PI - (i + 0.5) * PI / 8.0, az = -PI + (j + 0.5) * 2.0 * PI / 16.0;
    const double top = 0.5 * PI - i * PI / 8.0;
    cells[c][0] = (float)(cos(el) * cos(az));
    cells[c][1] = (float)(cos(el) * sin(az));
    cells[c][2] = (float)sin(el);
    cells[c][3] = (float)((2.0 * PI / 16.0) * (sin(top) - sin(top - PI / 8.0)));
  }
  void *p = nullptr;
  sl_status s;
  struct Buf { void **dst; size_t bytes; };
  void *nodes_new = nullptr;
  const Buf bufs[] = {{&nodes_new, (size_t)N * 96}, {(void **)&B.kids, (size_t)std::max(t->kid_count, 1u) * 4},
                      {(void **)&B.up_ent, std::max<size_t>(up.size(), 4) * 4},
                      {(void **)&B.node_acc, (size_t)N * 14 * 16}, {(void **)&B.node_sh, (size_t)N * 14 * 4},
                      {(void **)&B.probe_data, (size_t)NP * 128 * 32}, {(void **)&B.probe_sh, (size_t)NP * 9 * 16},
                      {(void **)&B.lo_dir, (size_t)G * 16}, {(void **)&B.lo_b, (size_t)G * 48},
                      {(void **)&B.xfer, (size_t)G * 48}, {(void **)&B.c0, (size_t)G * 16},
                      {(void **)&B.cells, sizeof(cells)}, {&B.pp, sizeof(ProbeParamsDev)},
                      {(void **)&B.wcache, (size_t)G * 14 * 4},
                      {(void **)&B.walk, std::max<size_t>(walk.size(), 8) * 4},
                      {(void **)&B.walk_root, (size_t)SL_MAX_PARTS * 4}};
  for (const Buf &b : bufs) {
    if ((s = scene_alloc(ctx, sc, b.bytes, &p)) != SL_OK) return s;
    *b.dst = p;
    SL_HIP(ctx, hipMemsetAsync(p, 0, b.bytes, ctx->stream));
  }
  SL_HIP(ctx, hipMemcpyAsync(nodes_new, t->nodes, (size_t)N * 96, hipMemcpyHostToDevice, ctx->stream));
  SL_HIP(ctx, hipMemcpyAsync(B.kids, t->kids, (size_t)t->kid_count * 4, hipMemcpyHostToDevice, ctx->stream));
  if (!up.empty())
    SL_HIP(ctx, hipMemcpyAsync(B.up_ent, up.data(), up.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  SL_HIP(ctx, hipMemcpyAsync(B.cells, cells, sizeof(cells), hipMemcpyHostToDevice, ctx->stream));
  SL_HIP(ctx, hipMemcpyAsync(B.pp, &PP, sizeof(PP), hipMemcpyHostToDevice, ctx->stream));
  if (!walk.empty())
    SL_HIP(ctx, hipMemcpyAsync(B.walk, walk.data(), walk.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  SL_HIP(ctx, hipMemcpyAsync(B.walk_root, walk_root.data(), (size_t)SL_MAX_PARTS * 4, hipMemcpyHostToDevice, ctx->stream));
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  B.w_version = 0xFFFFFFFFu;
  // the whole tree replaces the scene's leaf array (its leaves are the same records)
  auto it = std::find(sc->allocs.begin(), sc->allocs.end(), (void *)sc->nodes);
  if (it != sc->allocs.end()) {
    (void)hipFree(*it);
    sc->allocs.erase(it);
    sc->bytes -= (size_t)G * 96;
  }
  sc->nodes = (float4 *)nodes_new;
  sc->visit.valid = false;
  B.node_count = N;
  B.kid_count = t->kid_count;
  B.max_depth = t->max_depth;
  B.n_probes = NP;
  B.on = true;
  B.lit = false;
  sc->lo_b = B.lo_b;
  return SL_OK;
}

extern "C" sl_status sl_bounce_encode(sl_context *ctx, sl_scene *sc, const void *probe_params, size_t len,
                                      uint32_t flags) {
  SL_LIVE(ctx);
  if (!sc || !probe_params) return fail(ctx, SL_ERR_INVALID, "null argument");
  sl_scene::Bounce &B = sc->bounce;
  if (!B.on) return fail(ctx, SL_ERR_INVALID, "sl_bounce_init has not run for this scene");
  if (!B.lit) return fail(ctx, SL_ERR_INVALID, "the bounce reads the scene's last lighting encode: none yet");
  if (len != sizeof(ProbeParamsDev)) return fail(ctx, SL_ERR_LAYOUT, "ProbeParams is 9632 B");
  ProbeParamsDev PP;
  read_struct(probe_params, &PP);
  const uint32_t NP = check_probe_params(ctx, sc, PP, B.node_count);
  if (!NP) return ctx->poisoned != SL_OK ? ctx->poisoned : SL_ERR_INVALID;
  if (NP != B.n_probes) return fail(ctx, SL_ERR_INVALID, "ProbeParams: the probe count changed since sl_bounce_init");
  if ((B.U.iso & 3u) != 3u) return fail(ctx, SL_ERR_UNSUPPORTED, "the bounce path needs FrameUniforms.iso bits 0 and 1");
  const bool timers = (flags & SL_BOUNCE_TIMERS) != 0u;
  hipStream_t st = ctx->stream;
  const FrameUniformsDev &U = B.U;
  const ProbeParamsDev *PPd = (const ProbeParamsDev *)B.pp;
  const uint32_t G = sc->count, gblocks = (G + 63) / 64;
  if (timers) SL_HIP(ctx, hipEventRecord(ctx->bev[0], st));
  SL_HIP(ctx, hipMemcpyAsync(B.pp, &PP, sizeof(PP), hipMemcpyHostToDevice, st));
  if ((flags & SL_BOUNCE_LIGHT_STILL) == 0u) {
    if (flags & SL_BOUNCE_XFER)
      sl_iso_xfer<<<gblocks, 64, 0, st>>>(U, PPd, sc->g[0], sc->g[2], sc->storage, sc->nodes, sc->vgrid, sc->parts,
                                           sc->iso_norm, B.cells, B.xfer);
    const uint32_t nd = (U.iso & 2u) != 0u ? (G + 3) / 4 : G;
    sl_bounce_lo_dir<<<(nd + 63) / 64, 64, 0, st>>>(U, B.ENV, sc->g[1], sc->g[2], sc->g[3], sc->storage, sc->nodes,
                                                    sc->vgrid, sc->parts, sc->materials, sc->iso_norm,
                                                    sc->env_T, B.c0, B.lo_dir);
    // node_leaf, node_up per depth (deepest first) and node_sh2: the leaves' weights follow the part table, the
    // leaves' sums are formed where they are read
    if (B.w_version != sc->parts_version) {
      sl_leaf_weights<<<gblocks, 64, 0, st>>>(G, PPd, sc->g[2], sc->storage, sc->nodes, sc->vgrid, sc->parts, B.wcache);
      B.w_version = sc->parts_version;
    }
    sl_leaf_sh<<<gblocks, 64, 0, st>>>(G, PPd, sc->g[2], sc->storage, B.wcache, B.lo_dir, B.lo_b, B.node_sh);
    // the deep depths one launch each; the shallow ones (depth 1 up to tail_hi) in one workgroup
    for (int d = (int)B.max_depth - 1; d > (int)B.tail_hi; d--) {
      const uint32_t cnt = B.lv_cnt[d];
      if (cnt)
        sl_node_up_sh<<<(cnt + 63) / 64, 64, 0, st>>>(G, B.node_count, PPd, sc->g[2], sc->storage, B.kids, B.up_ent,
                                                       B.lv_off[d], cnt, B.wcache, B.lo_dir, B.lo_b, B.node_acc,
                                                       B.node_sh);
    }
    if (B.tail_hi >= 1u)
      sl_node_up_tail<<<1, UP_TAIL_THREADS, 0, st>>>(G, B.node_count, PPd, sc->g[2], sc->storage, B.kids, B.up_ent,
                                                     B.tail, B.tail_hi, B.wcache, B.lo_dir, B.lo_b, B.node_acc,
                                                     B.node_sh);
    if (timers) SL_HIP(ctx, hipEventRecord(ctx->bev[1], st));
    const uint32_t rays = (NP * 128u + PP.subset - 1u) / PP.subset;
    if (B.walk_stack <= 32u)
      sl_probe_gather<32><<<(rays + 63) / 64, 64, 0, st>>>(U, PPd, B.ENV, sc->nodes, sc->vgrid, sc->parts, B.walk,
                                                           B.walk_root, B.node_sh, B.cells, B.probe_data, rays);
    else
      sl_probe_gather<64><<<(rays + 63) / 64, 64, 0, st>>>(U, PPd, B.ENV, sc->nodes, sc->vgrid, sc->parts, B.walk,
                                                           B.walk_root, B.node_sh, B.cells, B.probe_data, rays);
    if (timers) SL_HIP(ctx, hipEventRecord(ctx->bev[2], st));
    sl_probe_finish<<<NP, 64, 0, st>>>(U, PPd, sc->parts, B.cells, B.probe_data, (float *)B.probe_sh);
  } else if (timers) {
    SL_HIP(ctx, hipEventRecord(ctx->bev[1], st));
    SL_HIP(ctx, hipEventRecord(ctx->bev[2], st));
  }
  const uint32_t ns = PP.sh_full ? G : (G + PP.subset - 1u) / PP.subset;
  sl_bounce_sh<<<(ns + 63) / 64, 64, 0, st>>>(U, PPd, sc->g[0], sc->g[1], sc->g[2], sc->g[3], sc->storage, sc->nodes,
                                              sc->parts, B.probe_data, B.probe_sh, B.xfer, B.lo_b);
  if (timers) SL_HIP(ctx, hipEventRecord(ctx->bev[3], st));
  ctx->bounce_timed = timers;
  sc->lo_b = B.lo_b;
  return check_launch(ctx, "bounce kernels");
}

extern "C" sl_status sl_bounce_read(sl_context *ctx, sl_scene *sc, uint32_t which, void *host, size_t bytes,
                                    size_t *written) {
  SL_LIVE(ctx);
  if (!sc) return fail(ctx, SL_ERR_INVALID, "null argument");
  sl_scene::Bounce &B = sc->bounce;
  if (!B.on) return fail(ctx, SL_ERR_INVALID, "sl_bounce_init has not run for this scene");
  const void *src = nullptr;
  size_t n = 0;
  switch (which) {
    case SL_BOUNCE_LO_DIR: src = B.lo_dir; n = (size_t)sc->count * 16; break;
    case SL_BOUNCE_LO_B: src = B.lo_b; n = (size_t)sc->count * 48; break;
    case SL_BOUNCE_NODE_ACC: src = B.node_acc; n = (size_t)B.node_count * 14 * 16; break;
    case SL_BOUNCE_NODE_SH: src = B.node_sh; n = (size_t)B.node_count * 14 * 4; break;
    case SL_BOUNCE_PROBE_DATA: src = B.probe_data; n = (size_t)B.n_probes * 128 * 32; break;
    case SL_BOUNCE_PROBE_SH: src = B.probe_sh; n = (size_t)B.n_probes * 9 * 16; break;
    case SL_BOUNCE_XFER_SH: src = B.xfer; n = (size_t)sc->count * 48; break;
    case SL_BOUNCE_C0: src = B.c0; n = (size_t)sc->count * 16; break;
    case SL_BOUNCE_LEAF_W: src = B.wcache; n = (size_t)sc->count * 56; break;
    default: return fail(ctx, SL_ERR_INVALID, "unknown bounce buffer");
  }
  if (!host) {   // the size alone
    if (written) *written = n;
    return SL_OK;
  }
  if (bytes < n) return fail(ctx, SL_ERR_INVALID, "host buffer too small");
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (which == SL_BOUNCE_NODE_ACC) {   // stored direction-major: handed out node-major
    std::vector<float4> tmp((size_t)B.node_count * 14);
    SL_HIP(ctx, hipMemcpy(tmp.data(), src, n, hipMemcpyDeviceToHost));
    float4 *o = (float4 *)host;
    for (size_t k = 0; k < 14; k++)
      for (size_t i = 0; i < B.node_count; i++) o[i * 14 + k] = tmp[k * B.node_count + i];
  } else {
    SL_HIP(ctx, hipMemcpy(host, src, n, hipMemcpyDeviceToHost));
  }
  if (written) *written = n;
  return SL_OK;
}

extern "C" sl_status sl_scene_read(sl_context *ctx, sl_scene *sc, uint32_t which, void *host, size_t bytes,
                                   size_t *written) {
  SL_LIVE(ctx);
  if (!sc) return SL_ERR_INVALID;
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const void *src = nullptr;
  size_t n = 0;
  switch (which) {
    case SL_SCENE_COLOUR: src = sc->colour; n = (size_t)sc->count * 16; break;
    case SL_SCENE_LAMP_T: src = sc->lamp_T; n = sc->lamp_T ? (size_t)sc->count * 4 : 0; break;
    case SL_SCENE_ENV_T: src = sc->env_T; n = sc->env_T ? (size_t)sc->count * 16 : 0; break;
    case SL_SCENE_ISO_NORM: src = sc->iso_norm; n = sc->iso_norm ? (size_t)sc->count * 4 : 0; break;
    case SL_SCENE_LSM_ACC: src = sc->lsm_acc.p; n = (size_t)sc->lsm_res * sc->lsm_res * 12 * sc->lsm_parts; break;
    case SL_SCENE_LSM_MIP: src = sc->lsm_mip.p; n = (size_t)sc->lsm_layer_texels * 16 * sc->lsm_parts; break;
    case SL_SCENE_SUN_ACC: src = sc->sun_acc.p; n = (size_t)sc->sun_res * sc->sun_res * 12 * sc->sun_n; break;
    case SL_SCENE_SUN_MIP: src = sc->sun_mip.p; n = (size_t)sc->sun_layer_texels * 16 * sc->sun_n; break;
    default: return fail(ctx, SL_ERR_INVALID, "unknown scene state");
  }
  if (written) *written = n;
  if (!host || n == 0) return SL_OK;
  if (bytes < n) return fail(ctx, SL_ERR_INVALID, "host buffer too small");
  SL_HIP(ctx, hipMemcpyAsync(host, src, n, hipMemcpyDeviceToHost, ctx->stream));
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SL_OK;
}

extern "C" sl_status sl_lighting_stats_read(sl_context *ctx, sl_scene *sc, sl_lighting_stats *out) {
  SL_LIVE(ctx);
  if (!sc || !out) return fail(ctx, SL_ERR_INVALID, "null argument");
  memset(out, 0, sizeof(*out));
  uint32_t dev[4] = {0, 0, 0, 0};
  if (ctx->map_stats.p) {
    SL_HIP(ctx, hipMemcpyAsync(dev, ctx->map_stats.p, sizeof(dev), hipMemcpyDeviceToHost, ctx->stream));
    SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  out->lsm_records = ctx->map_records[0];
  out->lsm_pairs = dev[0];
  out->lsm_tiles = ctx->map_tiles[0];
  out->lsm_tiles_used = dev[1];
  out->sun_records = ctx->map_records[1];
  out->sun_pairs = dev[2];
  out->sun_tiles = ctx->map_tiles[1];
  out->sun_tiles_used = dev[3];
  return SL_OK;
}

extern "C" sl_status sl_timers_read(sl_context *ctx, const char **names, float *ms, uint32_t cap, uint32_t *n) {
  SL_LIVE(ctx);
  uint32_t k = 0;
  if (ctx->timers_valid)
    for (; k < (uint32_t)kPasses && k < cap; k++) {
      if (names) names[k] = kPassNames[k];
      if (ms) ms[k] = ctx->ms[k];
    }
  // the last sl_lighting_encode's set, or the one before it while the last still runs (a frame's lighting is
  // complete once its sl_frame_finish returned, while the next frame's lighting may already run)
  uint32_t use = ctx->lt_cur;
  if (ctx->lt[use].valid && ctx->lt[use ^ 1u].valid && hipEventQuery(ctx->lt[use].lev[4]) == hipErrorNotReady) {
    (void)hipGetLastError();                           // not ready is no error: keep it from the next launch check
    use ^= 1u;
  }
  const sl_context::LightTimers &LT = ctx->lt[use];
  if (LT.valid) {
    // the lighting passes of that sl_lighting_encode with SL_LIGHT_TIMERS (PASSES names: shadow, shade).
    // "sunmap" runs on the side stream alongside shadow and shade; "sun_wait" is the context stream's wait for it
    // before env_direct, and "lighting" the whole encode.
    static const char *const lnames[] = {"shadow", "shade", "sun_wait", "env_direct"};
    SL_HIP(ctx, hipEventSynchronize(LT.lev[4]));
    for (int j = 0; j < 4 && k < cap; j++, k++) {
      float v = 0.0f;
      SL_HIP(ctx, hipEventElapsedTime(&v, LT.lev[j], LT.lev[j + 1]));
      if (names) names[k] = lnames[j];
      if (ms) ms[k] = v;
    }
    if (LT.sun_timed && k < cap) {
      float v = 0.0f;
      SL_HIP(ctx, hipEventElapsedTime(&v, LT.sev[0], LT.sev[1]));
      if (names) names[k] = "sunmap";
      if (ms) ms[k] = v;
      k++;
    }
    if (k < cap) {
      float v = 0.0f;
      SL_HIP(ctx, hipEventElapsedTime(&v, LT.lev[0], LT.lev[4]));
      if (names) names[k] = "lighting";
      if (ms) ms[k] = v;
      k++;
    }
    // the map-sum kernel alone (sl_map_accumulate), inside "shadow" and "sunmap"
    static const char *const mnames[] = {"lsm_sum", "sun_sum"};
    for (int j = 0; j < 2 && k < cap; j++)
      if (LT.map_timed[j]) {
        float v = 0.0f;
        SL_HIP(ctx, hipEventElapsedTime(&v, LT.mev[2 * j], LT.mev[2 * j + 1]));
        if (names) names[k] = mnames[j];
        if (ms) ms[k] = v;
        k++;
      }
  }
  if (ctx->bounce_timed) {
    // the last sl_bounce_encode with SL_BOUNCE_TIMERS: the whole encode and its probe gather
    SL_HIP(ctx, hipEventSynchronize(ctx->bev[3]));
    const int ev[2][2] = {{0, 3}, {1, 2}};
    static const char *const bnames[] = {"bounce", "probe_gather"};
    for (int j = 0; j < 2 && k < cap; j++, k++) {
      float v = 0.0f;
      SL_HIP(ctx, hipEventElapsedTime(&v, ctx->bev[ev[j][0]], ctx->bev[ev[j][1]]));
      if (names) names[k] = bnames[j];
      if (ms) ms[k] = v;
    }
  }
  if (n) *n = k;
  return SL_OK;
}

#include "server.hip"
