// N-API addon over the splat frame path's C ABI (include/splatlight.h): what the JavaScript host binds in place of
// its WebGPU calls. Handles are napi_external values with finalizers; byte arguments are ArrayBuffers or typed
// arrays passed straight through (the C ABI copies them); errors throw with sl_last_error's message. Frame
// completion (sl_frame_finish) runs on the libuv pool through napi_async_work and resolves a Promise on the JS
// thread; nothing is called from HIP threads. A context is not thread-safe, so every call into one first waits
// until no frameFinish of it is queued or running on the pool and then holds the context's lock: a host that encodes
// the next frame before awaiting the last (src/main.js:886-906) gets each frame's own counts, and the pool never
// touches a context while the JS thread does.
#include <node_api.h>

#include <stdint.h>
#include <string.h>

#include <condition_variable>
#include <mutex>
#include <string>

#include "splatlight.h"

namespace {

#define NAPI_OK(call)                                                            \
  do {                                                                           \
    if ((call) != napi_ok) {                                                     \
      napi_throw_error(env, nullptr, "splatlight addon: N-API call failed");     \
      return nullptr;                                                            \
    }                                                                            \
  } while (0)

// A context is destroyed when its JS handle and every scene and buffer made on it are gone (finalizers may run in
// any order when the environment shuts down, so the count is kept here rather than by JS references).
struct Ctx {
  sl_context *ctx = nullptr;
  int refs = 1;                 // touched on the JS thread only
  std::mutex mu;
  std::condition_variable idle;
  int finishing = 0;            // frameFinish calls queued or running on the pool
};
// The context's lock, taken once no frameFinish of it is pending (see the file comment).
std::unique_lock<std::mutex> quiet(Ctx *c) {
  std::unique_lock<std::mutex> lk(c->mu);
  c->idle.wait(lk, [c] { return c->finishing == 0; });
  return lk;
}
void ctx_release(Ctx *c) {
  if (--c->refs == 0) {
    sl_ctx_destroy(c->ctx);
    delete c;
  }
}
struct Scene {
  Ctx *owner = nullptr;
  sl_context *ctx = nullptr;
  sl_scene *scene = nullptr;
};
struct Buf {
  Ctx *owner = nullptr;
  sl_context *ctx = nullptr;
  void *dev = nullptr;
  size_t bytes = 0;
};

napi_value throw_sl(napi_env env, const sl_context *ctx, sl_status s, const char *what) {
  std::string msg = std::string(what) + ": ";
  msg += ctx ? sl_last_error(ctx) : sl_session_error();
  msg += " (status " + std::to_string(s) + ")";
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

// An ArrayBuffer, a typed array or a DataView: its bytes (null for undefined / null).
bool bytes_of(napi_env env, napi_value v, const void **p, size_t *len) {
  *p = nullptr;
  *len = 0;
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok || t == napi_undefined || t == napi_null) return true;
  bool is = false;
  void *data = nullptr;
  if (napi_is_arraybuffer(env, v, &is) == napi_ok && is)
    return napi_get_arraybuffer_info(env, v, &data, len) == napi_ok && ((*p = data), true);
  if (napi_is_typedarray(env, v, &is) == napi_ok && is) {
    napi_typedarray_type tt;
    size_t n, off;
    napi_value ab;
    if (napi_get_typedarray_info(env, v, &tt, &n, &data, &ab, &off) != napi_ok) return false;
    size_t el = 1;
    switch (tt) {
      case napi_int16_array: case napi_uint16_array: el = 2; break;
      case napi_int32_array: case napi_uint32_array: case napi_float32_array: el = 4; break;
      case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: el = 8; break;
      default: el = 1;
    }
    *p = data;
    *len = n * el;
    return true;
  }
  if (napi_is_dataview(env, v, &is) == napi_ok && is) {
    napi_value ab;
    size_t off;
    if (napi_get_dataview_info(env, v, len, &data, &ab, &off) != napi_ok) return false;
    *p = data;
    return true;
  }
  return false;
}

template <class T>
T *unwrap(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "splatlight addon: expected a handle");
    return nullptr;
  }
  return static_cast<T *>(p);
}

uint32_t u32_arg(napi_env env, napi_value v, uint32_t dflt = 0) {
  uint32_t x = dflt;
  napi_valuetype t;
  if (napi_typeof(env, v, &t) == napi_ok && t == napi_number) napi_get_value_uint32(env, v, &x);
  return x;
}

napi_value prop(napi_env env, napi_value obj, const char *name) {
  napi_value v = nullptr;
  napi_get_named_property(env, obj, name, &v);
  return v;
}

void set_u32(napi_env env, napi_value obj, const char *name, uint32_t x) {
  napi_value v;
  napi_create_uint32(env, x, &v);
  napi_set_named_property(env, obj, name, v);
}

// sessionAcquire(name?) -> token (number); throws with the reference's message when held
napi_value SessionAcquire(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  char name[256] = {0};
  size_t n = 0;
  napi_valuetype t = napi_undefined;
  if (argc) napi_typeof(env, a[0], &t);
  if (t == napi_string) napi_get_value_string_utf8(env, a[0], name, sizeof(name), &n);
  int64_t token = 0;
  if (sl_status s = sl_session_acquire(t == napi_string ? name : nullptr, &token)) return throw_sl(env, nullptr, s, "acquireGpuSession");
  napi_value out;
  NAPI_OK(napi_create_int64(env, token, &out));
  return out;
}

napi_value SessionRelease(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  int64_t token = 0;
  if (argc) napi_get_value_int64(env, a[0], &token);
  sl_session_release(token);
  return nullptr;
}

// ctxCreate(device = 0) -> handle
napi_value CtxCreate(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  const int dev = argc ? (int)u32_arg(env, a[0]) : 0;
  Ctx *c = new Ctx();
  if (sl_status s = sl_ctx_create(dev, &c->ctx)) {
    delete c;
    return throw_sl(env, nullptr, s, "ctxCreate");
  }
  napi_value out;
  NAPI_OK(napi_create_external(
      env, c,
      [](napi_env, void *p, void *) { ctx_release(static_cast<Ctx *>(p)); },
      nullptr, &out));
  return out;
}

// bufAlloc(ctx, bytes) -> handle; bufRead(buf) -> ArrayBuffer
napi_value BufAlloc(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Ctx *c = unwrap<Ctx>(env, a[0]);
  if (!c) return nullptr;
  double bytes = 0;
  napi_get_value_double(env, a[1], &bytes);
  Buf *b = new Buf();
  b->ctx = c->ctx;
  b->bytes = (size_t)bytes;
  auto lk = quiet(c);
  if (sl_status s = sl_buf_alloc(c->ctx, b->bytes, &b->dev)) {
    delete b;
    return throw_sl(env, c->ctx, s, "bufAlloc");
  }
  b->owner = c;
  c->refs++;
  napi_value out;
  NAPI_OK(napi_create_external(
      env, b,
      [](napi_env, void *p, void *) {
        Buf *b = static_cast<Buf *>(p);
        {
          auto lk = quiet(b->owner);
          sl_buf_free(b->ctx, b->dev);
        }
        ctx_release(b->owner);
        delete b;
      },
      nullptr, &out));
  return out;
}

napi_value BufRead(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Buf *b = unwrap<Buf>(env, a[0]);
  if (!b) return nullptr;
  void *data = nullptr;
  napi_value ab;
  NAPI_OK(napi_create_arraybuffer(env, b->bytes, &data, &ab));
  auto lk = quiet(b->owner);
  if (sl_status s = sl_buf_read(b->ctx, b->dev, 0, data, b->bytes)) return throw_sl(env, b->ctx, s, "bufRead");
  return ab;
}

// sceneCreate(ctx, {gaussians, parts, materials?, nodes?, vgrid?, emitters?, storage?}) -> handle. Record counts
// follow from the byte sizes (64 B Gaussians, 128 B parts, 32 B materials, 96 B nodes, 128 B grids).
napi_value SceneCreate(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Ctx *c = unwrap<Ctx>(env, a[0]);
  if (!c) return nullptr;
  sl_scene_desc d;
  memset(&d, 0, sizeof(d));
  size_t n;
  const void *p;
  if (!bytes_of(env, prop(env, a[1], "gaussians"), &p, &n)) return nullptr;
  d.gaussians = p; d.gaussian_count = (uint32_t)(n / 64); d.gaussian_stride = 64;
  if (!bytes_of(env, prop(env, a[1], "parts"), &p, &n)) return nullptr;
  d.parts = p; d.part_count = (uint32_t)(n / 128); d.part_stride = 128;
  if (!bytes_of(env, prop(env, a[1], "materials"), &p, &n)) return nullptr;
  d.materials = p; d.material_count = (uint32_t)(n / 32); d.material_stride = 32;
  if (!bytes_of(env, prop(env, a[1], "nodes"), &p, &n)) return nullptr;
  d.nodes = p; d.node_count = (uint32_t)(n / 96); d.node_stride = 96;
  if (!bytes_of(env, prop(env, a[1], "vgrid"), &p, &n)) return nullptr;
  d.vgrid = p; d.vgrid_count = (uint32_t)(n / 128); d.vgrid_stride = 128;
  if (!bytes_of(env, prop(env, a[1], "emitters"), &p, &n)) return nullptr;
  d.emitters = p; d.emitter_count = (uint32_t)(n / 64);
  d.storage = u32_arg(env, prop(env, a[1], "storage"));
  Scene *s = new Scene();
  s->ctx = c->ctx;
  auto lk = quiet(c);
  if (sl_status rc = sl_scene_create(c->ctx, &d, &s->scene)) {
    delete s;
    return throw_sl(env, c->ctx, rc, "sceneCreate");
  }
  s->owner = c;
  c->refs++;
  napi_value out;
  NAPI_OK(napi_create_external(
      env, s,
      [](napi_env, void *p, void *) {
        Scene *s = static_cast<Scene *>(p);
        {
          auto lk = quiet(s->owner);
          sl_scene_destroy(s->ctx, s->scene);
        }
        ctx_release(s->owner);
        delete s;
      },
      nullptr, &out));
  return out;
}

// sceneWriteParts(scene, partBytes) / sceneWriteColour(scene, Float32Array rgba)
napi_value SceneWriteParts(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Scene *s = unwrap<Scene>(env, a[0]);
  if (!s) return nullptr;
  const void *p;
  size_t n;
  if (!bytes_of(env, a[1], &p, &n)) return nullptr;
  auto lk = quiet(s->owner);
  if (sl_status rc = sl_scene_write_parts(s->ctx, s->scene, p, n)) return throw_sl(env, s->ctx, rc, "writeParts");
  return nullptr;
}

napi_value SceneWriteColour(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Scene *s = unwrap<Scene>(env, a[0]);
  if (!s) return nullptr;
  const void *p;
  size_t n;
  if (!bytes_of(env, a[1], &p, &n)) return nullptr;
  auto lk = quiet(s->owner);
  if (sl_status rc = sl_scene_write_colour(s->ctx, s->scene, (const float *)p, n / 16))
    return throw_sl(env, s->ctx, rc, "bindColour");
  return nullptr;
}

// sceneWriteGaussians(scene, bytes) / sceneWriteMaterials(scene, bytes): the Lighting constructor's re-upload of the
// records and its material table (src/lighting.js:467, 1152-1227)
napi_value SceneWriteRecords(napi_env env, napi_callback_info info, bool materials) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Scene *s = unwrap<Scene>(env, a[0]);
  if (!s) return nullptr;
  const void *p;
  size_t n;
  if (!bytes_of(env, a[1], &p, &n) || !p) {
    napi_throw_type_error(env, nullptr, "expected an ArrayBuffer or typed array of records");
    return nullptr;
  }
  auto lk = quiet(s->owner);
  sl_status rc = materials ? sl_scene_write_materials(s->ctx, s->scene, p, n) : sl_scene_write_gaussians(s->ctx, s->scene, p, n);
  if (rc) return throw_sl(env, s->ctx, rc, materials ? "writeMaterials" : "writeGaussians");
  return nullptr;
}
napi_value SceneWriteGaussians(napi_env env, napi_callback_info info) { return SceneWriteRecords(env, info, false); }
napi_value SceneWriteMaterials(napi_env env, napi_callback_info info) { return SceneWriteRecords(env, info, true); }

// sceneWriteBounce(scene, Float32Array of 3 vec4 per Gaussian, or null to clear)
napi_value SceneWriteBounce(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Scene *s = unwrap<Scene>(env, a[0]);
  if (!s) return nullptr;
  const void *p = nullptr;
  size_t n = 0;
  if (argc > 1 && !bytes_of(env, a[1], &p, &n)) return nullptr;
  auto lk = quiet(s->owner);
  if (sl_status rc = sl_scene_write_bounce(s->ctx, s->scene, (const float *)p, n / 48))
    return throw_sl(env, s->ctx, rc, "bindColour");
  return nullptr;
}

// sceneSetLod(scene, {nodes, kids, nodeGaussians, emitPart, sizeRule}): Lod's constructor (src/lod.js:114-140)
napi_value SceneSetLod(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Scene *s = unwrap<Scene>(env, a[0]);
  if (!s) return nullptr;
  const void *nodes = nullptr, *kids = nullptr, *ng = nullptr;
  size_t nn = 0, nk = 0, nng = 0;
  if (!bytes_of(env, prop(env, a[1], "nodes"), &nodes, &nn) || !bytes_of(env, prop(env, a[1], "kids"), &kids, &nk) ||
      !bytes_of(env, prop(env, a[1], "nodeGaussians"), &ng, &nng)) {
    napi_throw_type_error(env, nullptr, "sceneSetLod: nodes, kids and nodeGaussians must be ArrayBuffers or typed arrays");
    return nullptr;
  }
  sl_lod_tree t;
  t.nodes = nodes;
  t.node_count = (uint32_t)(nn / 96);
  t.node_stride = 96;
  t.kids = (const uint32_t *)kids;
  t.kid_count = (uint32_t)(nk / 4);
  t.node_gaussians = ng;
  t.node_gaussian_count = (uint32_t)(nng / 64);
  t.emit_part = u32_arg(env, prop(env, a[1], "emitPart"), 0xFFFFFFFFu);
  t.size_rule = u32_arg(env, prop(env, a[1], "sizeRule"), 1u);
  auto lk = quiet(s->owner);
  if (sl_status rc = sl_scene_set_lod(s->ctx, s->scene, &t)) return throw_sl(env, s->ctx, rc, "Lod");
  return nullptr;
}

// sceneSetFill(scene, Uint32Array table) / fillEncode(scene): Lighting.loadFill / encodeFill
// (src/lighting.js:961-1000)
napi_value SceneSetFill(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Scene *s = unwrap<Scene>(env, a[0]);
  if (!s) return nullptr;
  const void *p = nullptr;
  size_t n = 0;
  if (!bytes_of(env, a[1], &p, &n) || !p) {
    napi_throw_type_error(env, nullptr, "sceneSetFill: the table must be an ArrayBuffer or typed array");
    return nullptr;
  }
  auto lk = quiet(s->owner);
  if (sl_status rc = sl_scene_set_fill(s->ctx, s->scene, (const uint32_t *)p, n / 4))
    return throw_sl(env, s->ctx, rc, "loadFill");
  return nullptr;
}

napi_value FillEncode(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Scene *s = unwrap<Scene>(env, a[0]);
  if (!s) return nullptr;
  auto lk = quiet(s->owner);
  if (sl_status rc = sl_fill_encode(s->ctx, s->scene)) return throw_sl(env, s->ctx, rc, "encodeFill");
  return nullptr;
}

// displayEncode(ctx, frameBytes, env, skyBg, {width, height, hdrFormat, outFormat, flags}, hdrBuf, outBuf): the display
// pass of SplatRenderer.encode (shaders/display.wgsl, src/renderer.js:475-485)
napi_value DisplayEncode(napi_env env, napi_callback_info info) {
  size_t argc = 7;
  napi_value a[7];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Ctx *c = unwrap<Ctx>(env, a[0]);
  Buf *hdr = c ? unwrap<Buf>(env, a[5]) : nullptr;
  Buf *out = hdr ? unwrap<Buf>(env, a[6]) : nullptr;
  if (!c || !hdr || !out) return nullptr;
  const void *fb = nullptr, *ev = nullptr, *bg = nullptr;
  size_t nfb = 0, nev = 0, nbg = 0;
  if (!bytes_of(env, a[1], &fb, &nfb) || !bytes_of(env, a[2], &ev, &nev) || !bytes_of(env, a[3], &bg, &nbg)) {
    napi_throw_type_error(env, nullptr, "displayEncode: byte arguments must be ArrayBuffers or typed arrays");
    return nullptr;
  }
  sl_display_opts o;
  o.width = u32_arg(env, prop(env, a[4], "width"));
  o.height = u32_arg(env, prop(env, a[4], "height"));
  o.hdr_format = u32_arg(env, prop(env, a[4], "hdrFormat"), SL_OUT_RGBA16F);
  o.out_format = u32_arg(env, prop(env, a[4], "outFormat"), SL_DISPLAY_RGBA8);
  o.flags = u32_arg(env, prop(env, a[4], "flags"));
  auto lk = quiet(c);
  if (sl_status rc = sl_display_encode(c->ctx, fb, nfb, ev, nev, bg, nbg, &o, hdr->dev, out->dev))
    return throw_sl(env, c->ctx, rc, "display");
  return nullptr;
}

// sync(ctx): onSubmittedWorkDone, synchronously
napi_value Sync(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Ctx *c = unwrap<Ctx>(env, a[0]);
  if (!c) return nullptr;
  auto lk = quiet(c);
  if (sl_status rc = sl_sync(c->ctx)) return throw_sl(env, c->ctx, rc, "sync");
  return nullptr;
}

napi_value LightingIsoInit(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Scene *s = unwrap<Scene>(env, a[0]);
  if (!s) return nullptr;
  auto lk = quiet(s->owner);
  if (sl_status rc = sl_lighting_iso_init(s->ctx, s->scene)) return throw_sl(env, s->ctx, rc, "runIsoInit");
  return nullptr;
}

// lightingEncode(scene, {frameUniforms, emitParams, emitNodes, lsParams?, env?, sunParams?, sunLayers?}, timers)
// bounceInit(scene, {nodes, kids, maxDepth, probeParams}): Lighting.initBounce (src/lighting.js:723-840) with the
// asset's tree (nodes.bin: 96 B records, leaves first; kids.bin: u32) and the ProbeParams bytes it writes
napi_value BounceInit(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Scene *s = unwrap<Scene>(env, a[0]);
  if (!s) return nullptr;
  const void *nodes = nullptr, *kids = nullptr, *pp = nullptr;
  size_t nn = 0, nk = 0, npp = 0;
  if (!bytes_of(env, prop(env, a[1], "nodes"), &nodes, &nn) || !bytes_of(env, prop(env, a[1], "kids"), &kids, &nk) ||
      !bytes_of(env, prop(env, a[1], "probeParams"), &pp, &npp)) {
    napi_throw_type_error(env, nullptr, "bounceInit: nodes, kids and probeParams must be ArrayBuffers or typed arrays");
    return nullptr;
  }
  uint32_t max_depth = 0;
  napi_get_value_uint32(env, prop(env, a[1], "maxDepth"), &max_depth);
  sl_bounce_tree t;
  t.nodes = nodes;
  t.node_count = (uint32_t)(nn / 96);
  t.node_stride = 96;
  t.kids = (const uint32_t *)kids;
  t.kid_count = (uint32_t)(nk / 4);
  t.max_depth = max_depth;
  auto lk = quiet(s->owner);
  if (sl_status rc = sl_bounce_init(s->ctx, s->scene, &t, pp, npp)) return throw_sl(env, s->ctx, rc, "Lighting.initBounce");
  return nullptr;
}

// bounceEncode(scene, probeParams, {timers, xfer, lightStill}): Lighting.encodeBounce (src/lighting.js:1020-1103)
napi_value BounceEncode(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Scene *s = unwrap<Scene>(env, a[0]);
  if (!s) return nullptr;
  const void *pp = nullptr;
  size_t npp = 0;
  if (!bytes_of(env, a[1], &pp, &npp)) {
    napi_throw_type_error(env, nullptr, "bounceEncode: probeParams must be an ArrayBuffer or typed array");
    return nullptr;
  }
  uint32_t flags = 0;
  if (argc > 2) {
    bool v = false;
    if (napi_get_value_bool(env, prop(env, a[2], "timers"), &v) == napi_ok && v) flags |= SL_BOUNCE_TIMERS;
    v = false;
    if (napi_get_value_bool(env, prop(env, a[2], "xfer"), &v) == napi_ok && v) flags |= SL_BOUNCE_XFER;
    v = false;
    if (napi_get_value_bool(env, prop(env, a[2], "lightStill"), &v) == napi_ok && v) flags |= SL_BOUNCE_LIGHT_STILL;
  }
  auto lk = quiet(s->owner);
  if (sl_status rc = sl_bounce_encode(s->ctx, s->scene, pp, npp, flags)) return throw_sl(env, s->ctx, rc, "Lighting.encodeBounce");
  return nullptr;
}

napi_value LightingEncode(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Scene *s = unwrap<Scene>(env, a[0]);
  if (!s) return nullptr;
  sl_lighting_desc d;
  memset(&d, 0, sizeof(d));
  size_t n = 0;
  const void *p = nullptr;
  bool ok = bytes_of(env, prop(env, a[1], "frameUniforms"), &d.frame_uniforms, &n) &&
            bytes_of(env, prop(env, a[1], "emitParams"), &d.emit_params, &n) &&
            bytes_of(env, prop(env, a[1], "emitNodes"), &p, &n);
  d.emit_nodes = p;
  d.emit_node_count = (uint32_t)(n / 96);
  ok = ok && bytes_of(env, prop(env, a[1], "lsParams"), &d.ls_params, &n) &&
       bytes_of(env, prop(env, a[1], "env"), &d.env, &n) && bytes_of(env, prop(env, a[1], "sunParams"), &d.sun_params, &n) &&
       bytes_of(env, prop(env, a[1], "sunLayers"), &d.sun_layers, &n);
  if (!ok) {
    napi_throw_type_error(env, nullptr, "lightingEncode: byte arguments must be ArrayBuffers or typed arrays");
    return nullptr;
  }
  d.sun_layer_count = (uint32_t)(n / 32);
  bool timers = false;
  if (argc > 2) napi_get_value_bool(env, a[2], &timers);
  d.flags = timers ? SL_LIGHT_TIMERS : 0u;
  auto lk = quiet(s->owner);
  if (sl_status rc = sl_lighting_encode(s->ctx, s->scene, &d)) return throw_sl(env, s->ctx, rc, "Lighting.encode");
  return nullptr;
}

// frameEncode(scene, frameBytes, width, height, target, format, timers)
napi_value FrameEncode(napi_env env, napi_callback_info info) {
  size_t argc = 7;
  napi_value a[7];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Scene *s = unwrap<Scene>(env, a[0]);
  Buf *t = s ? unwrap<Buf>(env, a[4]) : nullptr;
  if (!s || !t) return nullptr;
  const void *fb;
  size_t flen;
  if (!bytes_of(env, a[1], &fb, &flen)) return nullptr;
  sl_frame_opts o;
  o.width = u32_arg(env, a[2]);
  o.height = u32_arg(env, a[3]);
  o.out_format = argc > 5 ? u32_arg(env, a[5], SL_OUT_RGBA16F) : SL_OUT_RGBA16F;
  bool timers = false;
  if (argc > 6) napi_get_value_bool(env, a[6], &timers);
  o.flags = timers ? SL_FRAME_TIMERS : 0u;
  auto lk = quiet(s->owner);
  if (sl_status rc = sl_frame_encode(s->ctx, s->scene, fb, flen, &o, t->dev)) return throw_sl(env, s->ctx, rc, "encode");
  return nullptr;
}

// frameFinish(ctx, display?) -> Promise<{slots, visible, pairs, overflow, tiles}> of the frame encoded last, waited for
// on the libuv pool (which holds the context's lock meanwhile; the next call into the context waits for it). display:
// {frameBytes, env, skyBg, width, height, hdrFormat, outFormat, flags, hdr, out} -- the display pass into out, enqueued
// on the pool right after the finish (so after an overflow re-run, and before any later frame's encode can rewrite
// the HDR target); its inputs are copied at the call.
struct FinishWork {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  Ctx *owner = nullptr;
  sl_context *ctx = nullptr;
  sl_status status = SL_OK;
  sl_frame_stats stats{};
  std::string err;
  bool display = false;
  std::string frame, env, sky_bg;        // copies of the display's byte inputs
  bool has_env = false, has_bg = false;
  sl_display_opts opts{};
  const void *hdr = nullptr;
  void *out = nullptr;
  napi_ref hdr_ref = nullptr, out_ref = nullptr;   // keep both buffers alive until the work completes
};

bool copy_bytes(napi_env env, napi_value v, std::string *dst, bool *present) {
  const void *p = nullptr;
  size_t n = 0;
  if (!bytes_of(env, v, &p, &n)) return false;
  if (present) *present = p != nullptr;
  if (p) dst->assign((const char *)p, n);
  return true;
}

napi_value FrameFinish(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Ctx *c = unwrap<Ctx>(env, a[0]);
  if (!c) return nullptr;
  FinishWork *w = new FinishWork();
  napi_valuetype t = napi_undefined;
  if (argc > 1) napi_typeof(env, a[1], &t);
  if (t == napi_object) {
    napi_value hv = prop(env, a[1], "hdr"), ov = prop(env, a[1], "out");
    Buf *hdr = unwrap<Buf>(env, hv);
    Buf *out = hdr ? unwrap<Buf>(env, ov) : nullptr;
    bool ok = out && copy_bytes(env, prop(env, a[1], "frameBytes"), &w->frame, nullptr) &&
              copy_bytes(env, prop(env, a[1], "env"), &w->env, &w->has_env) &&
              copy_bytes(env, prop(env, a[1], "skyBg"), &w->sky_bg, &w->has_bg);
    if (!ok) {
      delete w;
      if (hdr && out) napi_throw_type_error(env, nullptr, "frameFinish: display byte arguments must be ArrayBuffers or typed arrays");
      return nullptr;
    }
    w->display = true;
    w->opts.width = u32_arg(env, prop(env, a[1], "width"));
    w->opts.height = u32_arg(env, prop(env, a[1], "height"));
    w->opts.hdr_format = u32_arg(env, prop(env, a[1], "hdrFormat"), SL_OUT_RGBA16F);
    w->opts.out_format = u32_arg(env, prop(env, a[1], "outFormat"), SL_DISPLAY_RGBA8);
    w->opts.flags = u32_arg(env, prop(env, a[1], "flags"));
    w->hdr = hdr->dev;
    w->out = out->dev;
    NAPI_OK(napi_create_reference(env, hv, 1, &w->hdr_ref));
    NAPI_OK(napi_create_reference(env, ov, 1, &w->out_ref));
  }
  {
    auto lk = quiet(c);    // one finish at a time: each reports the frame encoded before it
    c->finishing++;
  }
  w->ctx = c->ctx;
  w->owner = c;
  c->refs++;
  napi_value promise, name;
  NAPI_OK(napi_create_promise(env, &w->deferred, &promise));
  NAPI_OK(napi_create_string_utf8(env, "sl_frame_finish", NAPI_AUTO_LENGTH, &name));
  NAPI_OK(napi_create_async_work(
      env, nullptr, name,
      [](napi_env, void *p) {
        FinishWork *w = static_cast<FinishWork *>(p);
        {
          std::lock_guard<std::mutex> lk(w->owner->mu);
          w->status = sl_frame_finish(w->ctx, &w->stats);
          if (w->status == SL_OK && w->display)
            w->status = sl_display_encode(w->ctx, w->frame.data(), w->frame.size(), w->has_env ? w->env.data() : nullptr,
                                          w->env.size(), w->has_bg ? w->sky_bg.data() : nullptr, w->sky_bg.size(),
                                          &w->opts, w->hdr, w->out);
          if (w->status) w->err = sl_last_error(w->ctx);
          w->owner->finishing--;
        }
        w->owner->idle.notify_all();
      },
      [](napi_env env, napi_status, void *p) {
        FinishWork *w = static_cast<FinishWork *>(p);
        if (w->status == SL_OK) {
          napi_value o;
          napi_create_object(env, &o);
          set_u32(env, o, "slots", w->stats.slots);
          set_u32(env, o, "visible", w->stats.visible);
          set_u32(env, o, "pairs", w->stats.pairs);
          set_u32(env, o, "overflow", w->stats.overflow);
          set_u32(env, o, "tiles", w->stats.tiles);
          napi_resolve_deferred(env, w->deferred, o);
        } else {
          napi_value msg, e;
          napi_create_string_utf8(env, w->err.c_str(), NAPI_AUTO_LENGTH, &msg);
          napi_create_error(env, nullptr, msg, &e);
          napi_reject_deferred(env, w->deferred, e);
        }
        if (w->hdr_ref) napi_delete_reference(env, w->hdr_ref);
        if (w->out_ref) napi_delete_reference(env, w->out_ref);
        ctx_release(w->owner);
        napi_delete_async_work(env, w->work);
        delete w;
      },
      w, &w->work));
  NAPI_OK(napi_queue_async_work(env, w->work));
  return promise;
}

// timersRead(ctx) -> {pass: ms}
napi_value TimersRead(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  Ctx *c = unwrap<Ctx>(env, a[0]);
  if (!c) return nullptr;
  const char *names[32];
  float ms[32];
  uint32_t n = 0;
  auto lk = quiet(c);
  if (sl_status rc = sl_timers_read(c->ctx, names, ms, 32, &n)) return throw_sl(env, c->ctx, rc, "timers");
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  for (uint32_t k = 0; k < n; k++) {
    napi_value v;
    napi_create_double(env, ms[k], &v);
    napi_set_named_property(env, o, names[k], v);
  }
  return o;
}

}  // namespace

NAPI_MODULE_INIT() {
  const napi_property_attributes kMethod =
      static_cast<napi_property_attributes>(napi_writable | napi_enumerable | napi_configurable);
  const napi_property_descriptor d[] = {
      {"sessionAcquire", nullptr, SessionAcquire, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"sessionRelease", nullptr, SessionRelease, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"ctxCreate", nullptr, CtxCreate, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"bufAlloc", nullptr, BufAlloc, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"bufRead", nullptr, BufRead, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"sceneCreate", nullptr, SceneCreate, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"sceneWriteParts", nullptr, SceneWriteParts, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"sceneWriteColour", nullptr, SceneWriteColour, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"sceneWriteBounce", nullptr, SceneWriteBounce, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"sceneWriteGaussians", nullptr, SceneWriteGaussians, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"sceneWriteMaterials", nullptr, SceneWriteMaterials, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"sceneSetLod", nullptr, SceneSetLod, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"sceneSetFill", nullptr, SceneSetFill, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"fillEncode", nullptr, FillEncode, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"displayEncode", nullptr, DisplayEncode, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"sync", nullptr, Sync, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"lightingIsoInit", nullptr, LightingIsoInit, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"lightingEncode", nullptr, LightingEncode, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"bounceInit", nullptr, BounceInit, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"bounceEncode", nullptr, BounceEncode, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"frameEncode", nullptr, FrameEncode, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"frameFinish", nullptr, FrameFinish, nullptr, nullptr, nullptr, kMethod, nullptr},
      {"timersRead", nullptr, TimersRead, nullptr, nullptr, nullptr, kMethod, nullptr},
  };
  napi_define_properties(env, exports, sizeof(d) / sizeof(d[0]), d);
  return exports;
}
