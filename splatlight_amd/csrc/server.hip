// Serving: continuous batching of render requests over resident scenes (host code, part of the library's single
// translation unit).
//
// A worker thread owns `lanes` execution lanes, each a context of its own (stream, workspace, counters). Requests
// wait in arrival order; whenever a lane is free, the worker admits requests from the head of the queue into one
// batch while the batch stays within max_batch_views and the workspace the lane needs for it (its slots, tiles and
// pair lists, laid out exactly as the frame path lays them) fits beside the other lanes' workspaces within the HBM
// budget, and launches it on that lane (sl_batch_encode's path). The budget bounds what the lanes allocate: a lane's
// growth (its first batches, a pair-list overflow re-run) is capped at what the others leave, idle lanes give their
// workspaces back when a batch cannot fit otherwise, and a request that could not fit alone is refused at submit. A lane's completion is polled with an event,
// so requests that arrive while other batches run are admitted as soon as a lane and memory free up, and up to
// `lanes` batches overlap on the device. Results carry submit / start / done times (seconds on a steady clock).
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>

struct ServerRequest {
  uint64_t id = 0;
  sl_view view{};
  std::vector<uint8_t> frame;
  uint32_t slots = 0, tiles = 0;  // its splat slots (rounded to the 256 the batch layout gives a view) and tiles
  double t_submit = 0.0;
  hipEvent_t ready = nullptr;    // the owner context's queued work at submit (a lighting of the scene, uploads)
};

struct ServerLane {
  sl_context *ctx = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<ServerRequest> reqs;
  uint64_t reserved = 0;          // workspace bytes the lane holds or may grow to for its running batch
  double t_start = 0.0;
};

struct sl_server {
  sl_context *owner = nullptr;
  sl_server_opts opts{};
  std::vector<ServerLane> lanes;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<ServerRequest> pending;          // submitted, not yet admitted
  std::deque<sl_request_result> results;      // completed, not yet polled
  uint64_t next_id = 1;
  uint64_t in_flight_bytes = 0;
  bool stop = false;
  std::atomic<bool> fail_next{false};         // sl_server_fail_next: the next completion check reports an error
  std::string err;
  std::thread worker;
  std::chrono::steady_clock::time_point t0;
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

// The workspace bytes a lane whose workspace already holds `w` needs for a batch of `slots` slots and `tiles` tiles
// (the pair lists at 8 per slot + 4096, as the batch path sizes them; never below what the workspace has).
uint64_t lane_need(const Workspace &w, uint32_t tile_bits, uint64_t slots, uint64_t tiles, uint64_t nslots) {
  const uint64_t pairs = std::min<uint64_t>(8ull * nslots + 4096, 1u << 29);
  Workspace tmp;
  const uint32_t s_cap = (uint32_t)std::max<uint64_t>(slots, w.slots_cap);
  const uint32_t p_cap = (uint32_t)std::max<uint64_t>(std::max<uint64_t>(pairs, slots), w.pairs_cap);
  const uint32_t t_cap = (uint32_t)std::max<uint64_t>(tiles, w.tiles_cap);
  return layout_workspace(nullptr, s_cap, p_cap, t_cap, tile_bits, tmp);
}

// A lane whose completion check returned q (hipSuccess, or the error that ended it): its requests' results, with
// the batch's status (on an error, every request carries it and the server's error names it).
void finish_lane(sl_server *s, ServerLane &ln, hipError_t q) {
  std::vector<sl_frame_stats> st(ln.reqs.size());
  sl_status rc;
  if (q == hipSuccess) {
    // a pair-list overflow re-runs the batch with grown lists: at most what the other lanes leave of the budget
    uint64_t others = 0;
    for (const ServerLane &l2 : s->lanes) others += &l2 == &ln ? 0 : l2.reserved;
    ln.ctx->ws_limit = s->opts.hbm_budget > others ? s->opts.hbm_budget - others : 1;
    rc = finish_batch(ln.ctx, st.data(), (uint32_t)st.size());
    ln.ctx->ws_limit = 0;
  } else {
    (void)hipStreamSynchronize(ln.ctx->stream);   // nothing of the batch still runs on the lane
    rc = hip_fail(ln.ctx, q, "server lane: batch completion");
  }
  const double t = s->now();
  std::lock_guard<std::mutex> g(s->mu);
  if (rc != SL_OK) s->err = ln.ctx->err;
  for (size_t k = 0; k < ln.reqs.size(); k++) {
    sl_request_result r{};
    r.id = ln.reqs[k].id;
    r.status = rc;
    r.batch_views = (uint32_t)ln.reqs.size();
    if (rc == SL_OK) r.stats = st[k];
    r.submit_s = ln.reqs[k].t_submit;
    r.start_s = ln.t_start;
    r.done_s = t;
    s->results.push_back(r);
  }
  ln.reqs.clear();
  ln.reserved = ln.ctx->ws.base ? ln.ctx->ws.bytes : 0;   // what it holds now
  s->in_flight_bytes = 0;
  for (const ServerLane &l2 : s->lanes) s->in_flight_bytes += l2.reserved;
  ln.busy = false;
  s->cv.notify_all();
}

// An idle lane's workspace given back (it is laid out again by its next batch).
void release_workspace(ServerLane &ln) {
  Workspace &w = ln.ctx->ws;
  if (!w.base) return;
  (void)hipStreamSynchronize(ln.ctx->stream);
  (void)hipFree(w.base);
  w = Workspace();
  drop_graph(ln.ctx);
  ln.reserved = 0;
}

// Admit from the head of the queue into one batch for a free lane; false when nothing was admitted (or, when the
// head request cannot fit beside the other lanes, an idle lane's workspace was given back: true, try again).
bool launch_lane(sl_server *s, ServerLane &ln) {
  std::vector<ServerRequest> batch;
  uint64_t need = 0;
  {
    std::lock_guard<std::mutex> g(s->mu);
    uint64_t others = 0;
    for (const ServerLane &l2 : s->lanes) others += &l2 == &ln ? 0 : l2.reserved;
    const uint32_t bits = ln.ctx->tuning.radix_bits;
    uint64_t slots = 0, tiles = 0, nslots = 0;
    while (!s->pending.empty() && batch.size() < s->opts.max_batch_views) {
      const ServerRequest &r = s->pending.front();
      const uint64_t n = lane_need(ln.ctx->ws, bits, slots + r.slots, tiles + r.tiles, nslots + r.slots);
      if (others + n > s->opts.hbm_budget) break;   // wait for memory
      need = n;
      slots += r.slots;
      tiles += r.tiles;
      nslots += r.slots;
      batch.push_back(std::move(s->pending.front()));
      s->pending.pop_front();
    }
    if (batch.empty()) {
      if (s->pending.empty()) return false;
      // the head does not fit beside the others: give back the largest idle workspace, if any holds memory
      ServerLane *idle = nullptr;
      for (ServerLane &l2 : s->lanes)
        if (&l2 != &ln && !l2.busy && l2.reserved && (!idle || l2.reserved > idle->reserved)) idle = &l2;
      if (!idle) return false;                      // wait for a busy lane to finish
      release_workspace(*idle);
      return true;
    }
    ln.reserved = need;
    s->in_flight_bytes = 0;
    for (const ServerLane &l2 : s->lanes) s->in_flight_bytes += l2.reserved;
  }
  std::vector<sl_view> views(batch.size());
  for (size_t k = 0; k < batch.size(); k++) {
    views[k] = batch[k].view;
    views[k].frame_bytes = batch[k].frame.data();
  }
  ln.t_start = s->now();
  ln.reqs = std::move(batch);
  ln.busy = true;
  sl_status rc = ln.ctx->poisoned != SL_OK ? SL_ERR_POISONED : SL_OK;
  for (ServerRequest &r : ln.reqs) {             // ordered after what the owner had queued when it was submitted
    if (rc == SL_OK && hipStreamWaitEvent(ln.ctx->stream, r.ready, 0) != hipSuccess) rc = SL_ERR_DEVICE;
    (void)hipEventDestroy(r.ready);
    r.ready = nullptr;
  }
  if (rc == SL_OK) {
    ln.ctx->ws_limit = ln.reserved;                  // the workspace grows to at most what was reserved
    rc = encode_batch(ln.ctx, views.data(), (uint32_t)views.size(), 0u);
    ln.ctx->ws_limit = 0;
  }
  if (rc == SL_OK && hipEventRecord(ln.done, ln.ctx->stream) != hipSuccess) rc = SL_ERR_DEVICE;
  if (rc != SL_OK) {
    const double t = s->now();
    std::lock_guard<std::mutex> g(s->mu);
    for (const ServerRequest &r : ln.reqs) {
      sl_request_result res{};
      res.id = r.id;
      res.status = rc;
      res.batch_views = (uint32_t)ln.reqs.size();
      res.submit_s = r.t_submit;
      res.start_s = ln.t_start;
      res.done_s = t;
      s->results.push_back(res);
    }
    s->err = ln.ctx->err.empty() ? std::string("server lane: device error") : ln.ctx->err;
    ln.reqs.clear();
    ln.reserved = ln.ctx->ws.base ? ln.ctx->ws.bytes : 0;
    s->in_flight_bytes = 0;
    for (const ServerLane &l2 : s->lanes) s->in_flight_bytes += l2.reserved;
    ln.busy = false;
    s->cv.notify_all();
  }
  return true;
}

void server_loop(sl_server *s) {
  (void)hipSetDevice(s->owner->device);
  for (;;) {
    bool progress = false;
    for (ServerLane &ln : s->lanes) {
      if (!ln.busy) continue;
      // any status but not-ready ends the batch: a device error must not leave the lane busy forever
      const hipError_t q = s->fail_next.exchange(false) ? hipErrorLaunchFailure : hipEventQuery(ln.done);
      if (q == hipErrorNotReady) continue;
      finish_lane(s, ln, q);
      progress = true;
    }
    // a lane poisoned by a device error takes no more work while a healthy lane remains (with none left, requests
    // still run through it and complete with SL_ERR_POISONED rather than wait forever)
    bool any_live = false;
    for (const ServerLane &ln : s->lanes) any_live = any_live || ln.ctx->poisoned == SL_OK;
    for (ServerLane &ln : s->lanes)
      if (!ln.busy && (ln.ctx->poisoned == SL_OK || !any_live) && launch_lane(s, ln)) progress = true;
    std::unique_lock<std::mutex> g(s->mu);
    bool busy = false;
    for (const ServerLane &ln : s->lanes) busy = busy || ln.busy;
    if (s->stop && !busy && s->pending.empty()) return;
    if (!progress) {
      // idle: wait for a submission; running: poll the lanes' events again shortly
      if (busy)
        s->cv.wait_for(g, std::chrono::microseconds(50));
      else
        s->cv.wait_for(g, std::chrono::milliseconds(2), [&] { return s->stop || !s->pending.empty(); });
    }
  }
}

}  // namespace

extern "C" sl_status sl_server_create(sl_context *ctx, const sl_server_opts *o, sl_server **out) {
  SL_LIVE(ctx);
  if (!o || !out) return fail(ctx, SL_ERR_INVALID, "null argument");
  if (o->lanes < 1 || o->lanes > 8 || o->max_batch_views < 1 || o->max_batch_views > kMaxViews || o->hbm_budget == 0)
    return fail(ctx, SL_ERR_INVALID, "server: 1-8 lanes, 1-256 views per batch, a non-zero HBM budget");
  sl_server *s = new sl_server();
  s->owner = ctx;
  s->opts = *o;
  s->t0 = std::chrono::steady_clock::now();
  s->lanes.resize(o->lanes);
  for (ServerLane &ln : s->lanes) {
    sl_status rc = sl_ctx_create(ctx->device, &ln.ctx);
    hipError_t e = rc == SL_OK ? hipEventCreateWithFlags(&ln.done, hipEventDisableTiming) : hipSuccess;
    if (rc == SL_OK && e == hipSuccess) {
      ln.ctx->tuning = ctx->tuning;
      continue;
    }
    for (ServerLane &l2 : s->lanes) {
      if (l2.done) (void)hipEventDestroy(l2.done);
      if (l2.ctx) sl_ctx_destroy(l2.ctx);
    }
    delete s;
    return rc != SL_OK ? fail(ctx, rc, "server lane context") : hip_fail(ctx, e, "hipEventCreate");
  }
  s->worker = std::thread(server_loop, s);
  *out = s;
  return SL_OK;
}

// Validated here, on the caller's thread with the owner context (sl_last_error(owner) has the reason), so a
// request the batch path would refuse never joins a batch.
extern "C" sl_status sl_server_submit(sl_server *s, const sl_view *v, uint64_t *id) {
  if (!s) return SL_ERR_INVALID;
  sl_context *ctx = s->owner;
  if (!v || !v->scene || !v->frame_bytes || !v->out_device) return fail(ctx, SL_ERR_INVALID, "null argument");
  if (v->frame_len != kFrameBytes) return fail(ctx, SL_ERR_INVALID, "RenderFrame must be 1200 B");
  if (v->out_format > SL_OUT_RGBA16F) return fail(ctx, SL_ERR_INVALID, "unknown output format");
  uint32_t r_slots = 0, r_tiles = 0;
  {
    FrameDev f;
    const sl_frame_opts o = {v->width, v->height, v->out_format, 0u};
    const uint32_t shift = s->lanes[0].ctx->tuning.tile_px == 32u ? 5u : 4u;   // the lanes' tiling
    const sl_status rc = decode_frame(ctx, v->scene, v->frame_bytes, &o, &f, shift);
    if (rc != SL_OK) return rc;
    if (f.tiles_x * f.tiles_y > 65536u) return fail(ctx, SL_ERR_CAPACITY, "more than 65536 tiles in a view");
    if (f.n_nodes) return fail(ctx, SL_ERR_UNSUPPORTED, "a screen-space LOD frame renders alone (sl_frame_encode)");
    r_slots = (f.nslots + 255u) / 256u * 256u;
    r_tiles = f.tiles_x * f.tiles_y;
  }
  const uint64_t alone = lane_need(Workspace(), s->lanes[0].ctx->tuning.radix_bits, r_slots, r_tiles, r_slots);
  if (alone > s->opts.hbm_budget)
    return fail(ctx, SL_ERR_CAPACITY, "server: the request alone needs a " + std::to_string(alone) +
                                          " B workspace, more than the HBM budget of " +
                                          std::to_string(s->opts.hbm_budget) + " B");
  ServerRequest r;
  r.view = *v;
  r.frame.assign((const uint8_t *)v->frame_bytes, (const uint8_t *)v->frame_bytes + kFrameBytes);
  r.slots = r_slots;
  r.tiles = r_tiles;
  r.t_submit = s->now();
  // the lane that renders it waits for the work the owner context had queued (e.g. sl_lighting_encode of the scene)
  SL_HIP(ctx, hipEventCreateWithFlags(&r.ready, hipEventDisableTiming));
  const hipError_t er = hipEventRecord(r.ready, ctx->stream);
  if (er != hipSuccess) {
    (void)hipEventDestroy(r.ready);
    return hip_fail(ctx, er, "hipEventRecord");
  }
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) {
      (void)hipEventDestroy(r.ready);
      return SL_ERR_INVALID;
    }
    r.id = s->next_id++;
    if (id) *id = r.id;
    s->pending.push_back(std::move(r));
  }
  s->cv.notify_all();
  return SL_OK;
}

extern "C" sl_status sl_server_poll(sl_server *s, sl_request_result *out, uint32_t cap, uint32_t *n, uint32_t wait_us) {
  if (!s || !n) return SL_ERR_INVALID;
  std::unique_lock<std::mutex> g(s->mu);
  if (s->results.empty() && wait_us)
    s->cv.wait_for(g, std::chrono::microseconds(wait_us), [&] { return !s->results.empty(); });
  uint32_t k = 0;
  while (k < cap && !s->results.empty()) {
    if (out) out[k] = s->results.front();
    s->results.pop_front();
    k++;
  }
  *n = k;
  return SL_OK;
}

extern "C" sl_status sl_server_info(sl_server *s, uint32_t *pending, uint32_t *running, uint64_t *in_flight_bytes) {
  if (!s) return SL_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  uint32_t r = 0;
  for (const ServerLane &ln : s->lanes) r += ln.busy ? (uint32_t)ln.reqs.size() : 0u;
  if (pending) *pending = (uint32_t)s->pending.size();
  if (running) *running = r;
  if (in_flight_bytes) *in_flight_bytes = s->in_flight_bytes;
  return SL_OK;
}

// A copy taken under the lock, per calling thread (the worker may write the error meanwhile).
extern "C" const char *sl_server_error(sl_server *s) {
  thread_local std::string copy;
  if (!s) return "";
  std::lock_guard<std::mutex> g(s->mu);
  copy = s->err;
  return copy.c_str();
}

extern "C" sl_status sl_server_fail_next(sl_server *s) {
  if (!s) return SL_ERR_INVALID;
  s->fail_next.store(true);
  return SL_OK;
}

extern "C" void sl_server_destroy(sl_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (ServerLane &ln : s->lanes) {
    if (ln.done) (void)hipEventDestroy(ln.done);
    if (ln.ctx) sl_ctx_destroy(ln.ctx);
  }
  delete s;
}
