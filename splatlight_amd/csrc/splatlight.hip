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

uint32_t rd_u32(const void *p, int word) { uint32_t v; memcpy(&v, (const char *)p + 4 * word, 4); return v; }
float rd_f32(const void *p, int word) { float v; memcpy(&v, (const char *)p + 4 * word, 4); return v; }

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

struct sl_scene {
  uint32_t count = 0, part_count = 0, n_emit = 0, material_count = 0, storage = SL_STORAGE_F32;
  float4 *g[4] = {nullptr, nullptr, nullptr, nullptr};
  float4 *e[4] = {nullptr, nullptr, nullptr, nullptr};
  float4 *parts = nullptr;
  float4 *colour = nullptr;
  float4 *lo_b = nullptr;
  std::vector<uint8_t> parts_host;
  std::vector<void *> allocs;
  size_t bytes = 0;
};

struct Workspace {
  uint32_t slots_cap = 0, pairs_cap = 0, tiles_cap = 0;
  void *base = nullptr;           // one allocation for everything below
  size_t bytes = 0;
  Splat *splats = nullptr;
  uint32_t *keys = nullptr, *keys_alt = nullptr, *vals_alt = nullptr, *order = nullptr;
  uint2 *tbox = nullptr;
  uint2 *partials = nullptr;      // project's per-block (visible, colour max)
  uint32_t *pk0 = nullptr, *pv0 = nullptr, *pk1 = nullptr, *pv1 = nullptr;
  // zeroed at the start of every frame:
  uint8_t *scratch = nullptr;
  size_t scratch_words = 0;
  FrameCounters *counters = nullptr;
  uint32_t *tickets = nullptr, *depth_hist = nullptr, *tile_hist = nullptr;
  uint32_t *dstatus[2] = {nullptr, nullptr}, *tstatus[2] = {nullptr, nullptr};
  uint64_t *chunk_status = nullptr;
  uint2 *ranges = nullptr;
  uint32_t *tile_order = nullptr;  // composite launch order (written every frame)
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
  bool timers_valid = false;
  float ms[kPasses] = {};
  // the last frame, for sl_frame_finish (re-run on overflow) and sl_frame_read
  bool has_last = false;
  FrameDev last_f{};
  sl_scene *last_scene = nullptr;
  void *last_out = nullptr;
  uint32_t last_fmt = 0, last_flags = 0;
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

#define SL_LIVE(ctx)                                                                              \
  do {                                                                                            \
    if (!(ctx)) return SL_ERR_INVALID;                                                            \
    if ((ctx)->poisoned != SL_OK) return SL_ERR_POISONED;                                         \
  } while (0)

sl_status check_launch(sl_context *ctx, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(ctx, e, what);
  return SL_OK;
}

sl_status dev_alloc(sl_context *ctx, size_t bytes, void **p) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  SL_HIP(ctx, hipMalloc(p, bytes));
  return SL_OK;
}

// Lay out the frame workspace for `slots` splat slots, `pairs` pairs and `tiles` tiles.
sl_status ensure_workspace(sl_context *ctx, uint32_t slots, uint32_t pairs, uint32_t tiles) {
  Workspace &w = ctx->ws;
  if (w.base && slots <= w.slots_cap && pairs <= w.pairs_cap && tiles <= w.tiles_cap) return SL_OK;
  const uint32_t s_cap = std::max(slots, w.slots_cap);
  const uint32_t p_cap = std::max(pairs, w.pairs_cap);
  const uint32_t t_cap = std::max(tiles, w.tiles_cap);
  const uint32_t d_tiles = (s_cap + RS_TILE - 1) / RS_TILE;
  const uint32_t p_tiles = (p_cap + RS_TILE - 1) / RS_TILE;
  const uint32_t chunks = (s_cap + PE_CHUNK - 1) / PE_CHUNK;
  size_t off = 0;
  auto take = [&](size_t bytes) { const size_t o = off; off = align_up(off + bytes, 256); return o; };
  const size_t o_splats = take((size_t)s_cap * sizeof(Splat));
  const size_t o_keys = take((size_t)s_cap * 4), o_keys_alt = take((size_t)s_cap * 4);
  const size_t o_vals_alt = take((size_t)s_cap * 4), o_order = take((size_t)s_cap * 4);
  const size_t o_tbox = take((size_t)s_cap * 8);
  const size_t o_partials = take((size_t)((s_cap + 255) / 256) * 8);
  const size_t o_pk0 = take((size_t)p_cap * 4), o_pv0 = take((size_t)p_cap * 4);
  const size_t o_pk1 = take((size_t)p_cap * 4), o_pv1 = take((size_t)p_cap * 4);
  const size_t o_scratch = off;
  const size_t o_counters = take(sizeof(FrameCounters));
  const size_t o_tickets = take(16 * 4);
  const size_t o_dhist = take(2 * 256 * 4), o_thist = take(2 * 256 * 4);
  const size_t o_ds0 = take((size_t)d_tiles * 256 * 4), o_ds1 = take((size_t)d_tiles * 256 * 4);
  const size_t o_ts0 = take((size_t)p_tiles * 256 * 4), o_ts1 = take((size_t)p_tiles * 256 * 4);
  const size_t o_chunk = take((size_t)chunks * 8);
  const size_t o_ranges = take((size_t)t_cap * 8);
  const size_t o_torder = take((size_t)t_cap * 4);
  const size_t total = off;
  void *base = nullptr;
  if (w.base) {
    SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SL_HIP(ctx, hipFree(w.base));
    w.base = nullptr;
  }
  sl_status s = dev_alloc(ctx, total, &base);
  if (s != SL_OK) return s;
  uint8_t *b = (uint8_t *)base;
  w.base = base;
  w.bytes = total;
  w.slots_cap = s_cap;
  w.pairs_cap = p_cap;
  w.tiles_cap = t_cap;
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
  w.scratch = b + o_scratch;
  w.scratch_words = (total - o_scratch) / 4;
  w.counters = (FrameCounters *)(b + o_counters);
  w.tickets = (uint32_t *)(b + o_tickets);
  w.depth_hist = (uint32_t *)(b + o_dhist);
  w.tile_hist = (uint32_t *)(b + o_thist);
  w.dstatus[0] = (uint32_t *)(b + o_ds0);
  w.dstatus[1] = (uint32_t *)(b + o_ds1);
  w.tstatus[0] = (uint32_t *)(b + o_ts0);
  w.tstatus[1] = (uint32_t *)(b + o_ts1);
  w.chunk_status = (uint64_t *)(b + o_chunk);
  w.ranges = (uint2 *)(b + o_ranges);
  w.tile_order = (uint32_t *)(b + o_torder);
  return SL_OK;
}

// RenderFrame (shaders/common.wgsl:44-77) -> FrameDev, refusing frames that turn on features outside this path.
sl_status decode_frame(sl_context *ctx, const sl_scene *sc, const void *fb, const sl_frame_opts *o, FrameDev *f) {
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
  if (rd_u32(fb, 180) != 0u) return fail(ctx, SL_ERR_UNSUPPORTED, "screen-space LOD (frame.lod) is not on this path");
  if (rd_f32(fb, 194) > 0.5f) return fail(ctx, SL_ERR_UNSUPPORTED, "explode is not on this path");
  if (rd_u32(fb, 296) != 0u) return fail(ctx, SL_ERR_UNSUPPORTED, "particles are not on this path");
  if (shutter > 0.0f) {
    for (uint32_t p = 0; p < sc->part_count; p++)
      if (rd_u32(sc->parts_host.data() + 128 * p, 31) & 2u)
        return fail(ctx, SL_ERR_UNSUPPORTED, "motion blur of a moving part is not on this path");
  }
  if (!(f->near_ > 0.0f) || !(f->far_ > f->near_)) return fail(ctx, SL_ERR_INVALID, "bad near / far");
  if ((uint32_t)f->vw != o->width || (uint32_t)f->vh != o->height || o->width == 0 || o->height == 0)
    return fail(ctx, SL_ERR_INVALID, "RenderFrame viewport does not match the output size");
  f->n_emit = sc->n_emit;
  f->nslots = sc->count + 1u + sc->n_emit;
  f->width = o->width;
  f->height = o->height;
  f->tiles_x = (o->width + 15u) / 16u;
  f->tiles_y = (o->height + 15u) / 16u;
  return SL_OK;
}

uint32_t grid_for(uint64_t n, uint32_t per_block, uint32_t cap) {
  uint64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return (uint32_t)std::min<uint64_t>(g, cap);
}

sl_status launch_frame(sl_context *ctx, const FrameDev &f, sl_scene *sc, void *out, uint32_t fmt, bool timers) {
  Workspace &w = ctx->ws;
  hipStream_t st = ctx->stream;
  const uint32_t ntiles = f.tiles_x * f.tiles_y;
  auto mark = [&](int k) -> sl_status {
    if (timers) SL_HIP(ctx, hipEventRecord(ctx->ev[k], st));
    return SL_OK;
  };
  sl_status s;
  if ((s = mark(0)) != SL_OK) return s;
  sl_fill_u32<<<grid_for(w.scratch_words, 256, 2048), 256, 0, st>>>((uint32_t *)w.scratch, 0u,
                                                                    (uint32_t)w.scratch_words);
  if ((s = mark(1)) != SL_OK) return s;
  sl_project_kernel<<<grid_for(f.nslots, 256, 1u << 30), 256, 0, st>>>(
      f, sc->g[0], sc->g[1], sc->g[2], sc->g[3], sc->e[0], sc->e[1], sc->e[2], sc->e[3], sc->parts, sc->colour,
      sc->lo_b, w.splats, w.keys, w.tbox, w.partials);
  sl_frame_totals<<<1, 1024, 0, st>>>(w.partials, (f.nslots + 255u) / 256u, w.counters);
  if ((s = mark(2)) != SL_OK) return s;
  // depth order: stable sort of the 16-bit keys, values = slot
  const uint32_t d_tiles = (f.nslots + RS_TILE - 1) / RS_TILE;
  const uint32_t d_grid = std::min<uint32_t>(d_tiles, 4u * ctx->cus);
  sl_radix_hist<<<grid_for(f.nslots, 256 * 16, 1024), RS_THREADS, 0, st>>>(w.keys, nullptr, f.nslots, 0u, 2u,
                                                                          w.depth_hist);
  sl_radix_bases<<<2, RS_THREADS, 0, st>>>(w.depth_hist);
  sl_radix_pass<<<d_grid, RS_THREADS, 0, st>>>(w.keys, nullptr, w.keys_alt, w.vals_alt, nullptr, f.nslots, 0u,
                                               w.depth_hist, w.dstatus[0], w.tickets + 0);
  sl_radix_pass<<<d_grid, RS_THREADS, 0, st>>>(w.keys_alt, w.vals_alt, nullptr, w.order, nullptr, f.nslots, 8u,
                                               w.depth_hist + 256, w.dstatus[1], w.tickets + 1);
  if ((s = mark(3)) != SL_OK) return s;
  sl_pairs_emit<<<grid_for(f.nslots, PE_CHUNK, 2u * ctx->cus), PE_THREADS, 0, st>>>(
      w.order, w.tbox, f.tiles_x, w.pairs_cap, w.counters, w.pk0, w.pv0, w.chunk_status, w.tickets + 2,
      w.tile_hist);
  if ((s = mark(4)) != SL_OK) return s;
  const uint32_t p_grid = std::min<uint32_t>((w.pairs_cap + RS_TILE - 1) / RS_TILE, 4u * ctx->cus);
  sl_radix_bases<<<2, RS_THREADS, 0, st>>>(w.tile_hist);
  sl_radix_pass<<<std::max(p_grid, 1u), RS_THREADS, 0, st>>>(w.pk0, w.pv0, w.pk1, w.pv1, &w.counters->pairs, 0u,
                                                             0u, w.tile_hist, w.tstatus[0], w.tickets + 3);
  sl_radix_pass<<<std::max(p_grid, 1u), RS_THREADS, 0, st>>>(w.pk1, w.pv1, w.pk0, w.pv0, &w.counters->pairs, 0u,
                                                             8u, w.tile_hist + 256, w.tstatus[1], w.tickets + 4);
  if ((s = mark(5)) != SL_OK) return s;
  sl_tile_ranges<<<grid_for(w.pairs_cap, 256, 4096), 256, 0, st>>>(w.pk0, w.counters, w.ranges);
  sl_tile_order<<<1, 1024, 0, st>>>(w.ranges, ntiles, w.tile_order);
  if ((s = mark(6)) != SL_OK) return s;
  sl_composite<<<ntiles, CP_THREADS, 0, st>>>(f, w.splats, w.pv0, w.ranges, w.tile_order, w.counters, out, fmt);
  if ((s = mark(7)) != SL_OK) return s;
  if ((s = check_launch(ctx, "frame kernels")) != SL_OK) return s;
  SL_HIP(ctx, hipMemcpyAsync(ctx->counters_host, w.counters, sizeof(FrameCounters), hipMemcpyDeviceToHost, st));
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
  if ((e = hipHostMalloc((void **)&ctx->counters_host, sizeof(FrameCounters), hipHostMallocDefault)) != hipSuccess)
    return bail(e);
  memset(ctx->counters_host, 0, sizeof(FrameCounters));
  for (int k = 0; k <= kPasses; k++)
    if ((e = hipEventCreate(&ctx->ev[k])) != hipSuccess) return bail(e);
  *out = ctx;
  return SL_OK;
}

extern "C" void sl_ctx_destroy(sl_context *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->ws.base) (void)hipFree(ctx->ws.base);
  for (int k = 0; k <= kPasses; k++)
    if (ctx->ev[k]) (void)hipEventDestroy(ctx->ev[k]);
  if (ctx->counters_host) (void)hipHostFree(ctx->counters_host);
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

sl_status upload_soa(sl_context *ctx, sl_scene *sc, const void *aos, uint32_t n, float4 *dst[4]) {
  for (int k = 0; k < 4; k++) {
    void *p = nullptr;
    sl_status s = dev_alloc(ctx, (size_t)std::max(n, 1u) * 16, &p);
    if (s != SL_OK) return s;
    sc->allocs.push_back(p);
    sc->bytes += (size_t)std::max(n, 1u) * 16;
    dst[k] = (float4 *)p;
  }
  if (n == 0) return SL_OK;
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

sl_status scene_alloc(sl_context *ctx, sl_scene *sc, size_t bytes, void **p) {
  sl_status s = dev_alloc(ctx, bytes, p);
  if (s != SL_OK) return s;
  sc->allocs.push_back(*p);
  sc->bytes += bytes;
  return SL_OK;
}

void scene_free(sl_scene *sc) {
  for (void *p : sc->allocs) (void)hipFree(p);
  delete sc;
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
  if (d->nodes && (d->node_stride != kNodeStride || d->node_count < d->gaussian_count))
    return fail(ctx, SL_ERR_LAYOUT, "nodes must be 96 B records, one leaf per Gaussian first");
  if (d->vgrid && (d->vgrid_stride != kVgridStride || d->vgrid_count < d->gaussian_count))
    return fail(ctx, SL_ERR_LAYOUT, "vgrid must be 128 B records, one per Gaussian");
  if (d->emitter_count > 16384u || (d->emitter_count && !d->emitters))
    return fail(ctx, SL_ERR_LAYOUT, "Bad emitter count " + std::to_string(d->emitter_count));
  if (d->storage > SL_STORAGE_BF16) return fail(ctx, SL_ERR_INVALID, "unknown storage format");
  SL_HIP(ctx, hipSetDevice(ctx->device));
  sl_scene *sc = new sl_scene();
  sc->count = d->gaussian_count;
  sc->part_count = d->part_count;
  sc->n_emit = d->emitter_count;
  sc->material_count = d->materials ? d->material_count : 0;
  sc->storage = d->storage;
  sl_status s = upload_soa(ctx, sc, d->gaussians, sc->count, sc->g);
  if (s == SL_OK) s = upload_soa(ctx, sc, d->emitters, sc->n_emit, sc->e);
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
  if (ctx && ctx->last_scene == scene) { ctx->last_scene = nullptr; ctx->has_last = false; }
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

// ---- frame --------------------------------------------------------------------------------------------------

extern "C" sl_status sl_frame_encode(sl_context *ctx, sl_scene *sc, const void *frame_bytes, size_t frame_len,
                                     const sl_frame_opts *opts, void *out_device) {
  SL_LIVE(ctx);
  if (!sc || !frame_bytes || !opts || !out_device) return fail(ctx, SL_ERR_INVALID, "null argument");
  if (frame_len != kFrameBytes) return fail(ctx, SL_ERR_INVALID, "RenderFrame must be 1200 B");
  if (opts->out_format > SL_OUT_RGBA16F) return fail(ctx, SL_ERR_INVALID, "unknown output format");
  FrameDev f;
  sl_status s = decode_frame(ctx, sc, frame_bytes, opts, &f);
  if (s != SL_OK) return s;
  const uint32_t ntiles = f.tiles_x * f.tiles_y;
  if (ntiles > 65536u) return fail(ctx, SL_ERR_CAPACITY, "more than 65536 16x16 tiles");
  SL_HIP(ctx, hipSetDevice(ctx->device));
  const uint32_t want_pairs = std::max<uint32_t>(ctx->ws.pairs_cap, std::min<uint64_t>(8ull * f.nslots + 4096, 1u << 29));
  if ((s = ensure_workspace(ctx, f.nslots, want_pairs, ntiles)) != SL_OK) return s;
  ctx->has_last = true;
  ctx->last_f = f;
  ctx->last_scene = sc;
  ctx->last_out = out_device;
  ctx->last_fmt = opts->out_format;
  ctx->last_flags = opts->flags;
  return launch_frame(ctx, f, sc, out_device, opts->out_format, (opts->flags & SL_FRAME_TIMERS) != 0);
}

extern "C" sl_status sl_frame_finish(sl_context *ctx, sl_frame_stats *stats) {
  SL_LIVE(ctx);
  if (!ctx->has_last) return fail(ctx, SL_ERR_INVALID, "no frame encoded");
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  FrameCounters c = *ctx->counters_host;
  if (c.overflow > 0) {
    // the pair buffers were too small: grow them and render the frame again
    const uint64_t need = (uint64_t)c.pairs + c.overflow;
    const uint64_t grown = std::min<uint64_t>((need * 5) / 4 + 4096, 1u << 30);
    if (need > grown) return fail(ctx, SL_ERR_CAPACITY, "frame needs more than 2^30 pairs");
    sl_status s = ensure_workspace(ctx, ctx->last_f.nslots, (uint32_t)grown, ctx->last_f.tiles_x * ctx->last_f.tiles_y);
    if (s != SL_OK) return s;
    s = launch_frame(ctx, ctx->last_f, ctx->last_scene, ctx->last_out, ctx->last_fmt,
                     (ctx->last_flags & SL_FRAME_TIMERS) != 0);
    if (s != SL_OK) return s;
    SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
    c = *ctx->counters_host;
    if (c.overflow > 0) return fail(ctx, SL_ERR_CAPACITY, "pair buffers still too small");
  }
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->slots = ctx->last_f.nslots;
    stats->visible = c.visible;
    stats->pairs = c.pairs;
    stats->overflow = c.overflow;
    stats->tiles = ctx->last_f.tiles_x * ctx->last_f.tiles_y;
  }
  if (ctx->timers_valid) {
    for (int k = 0; k < kPasses; k++) {
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, ctx->ev[k], ctx->ev[k + 1]);
      ctx->ms[k] = ms;
    }
  }
  return SL_OK;
}

extern "C" sl_status sl_frame_read(sl_context *ctx, uint32_t which, void *host, size_t bytes, size_t *written) {
  SL_LIVE(ctx);
  if (!ctx->has_last) return fail(ctx, SL_ERR_INVALID, "no frame encoded");
  SL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const FrameCounters &c = *ctx->counters_host;
  const Workspace &w = ctx->ws;
  const void *src = nullptr;
  size_t n = 0;
  switch (which) {
    case SL_ART_KEYS: src = w.keys; n = (size_t)ctx->last_f.nslots * 4; break;
    case SL_ART_ORDER: src = w.order; n = (size_t)ctx->last_f.nslots * 4; break;
    case SL_ART_SPLATS: src = w.splats; n = (size_t)ctx->last_f.nslots * sizeof(Splat); break;
    case SL_ART_PAIR_TILES: src = w.pk0; n = (size_t)c.pairs * 4; break;
    case SL_ART_PAIR_SLOTS: src = w.pv0; n = (size_t)c.pairs * 4; break;
    case SL_ART_RANGES: src = w.ranges; n = (size_t)ctx->last_f.tiles_x * ctx->last_f.tiles_y * 8; break;
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

extern "C" sl_status sl_timers_read(sl_context *ctx, const char **names, float *ms, uint32_t cap, uint32_t *n) {
  SL_LIVE(ctx);
  uint32_t k = 0;
  if (ctx->timers_valid)
    for (; k < (uint32_t)kPasses && k < cap; k++) {
      if (names) names[k] = kPassNames[k];
      if (ms) ms[k] = ctx->ms[k];
    }
  if (n) *n = k;
  return SL_OK;
}
