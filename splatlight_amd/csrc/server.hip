// Serving: continuous batching of render requests over resident scenes (host code, part of the library's single
// translation unit).
//
// A worker thread owns `lanes` execution lanes, each a context of its own (stream, workspace, counters). Requests
// wait in arrival order; whenever a lane is free, the worker admits requests from the head of the queue into one
// batch while the batch stays within max_batch_views and the HBM estimate of everything in flight stays within
// the budget, and launches it on that lane (sl_batch_encode's path). A lane's completion is polled with an event,
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
  uint64_t bytes = 0;            // HBM estimate while in flight
  double t_submit = 0.0;
  hipEvent_t ready = nullptr;    // the owner context's queued work at submit (a lighting of the scene, uploads)
};

struct ServerLane {
  sl_context *ctx = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<ServerRequest> reqs;
  uint64_t bytes = 0;
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

// HBM a request holds while in flight: its target, its splat slots (splat, keys, order, sort temporaries, tile
// box) and its pair lists at the workspace's initial 8 pairs per slot (key and value, two copies).
uint64_t request_bytes(const sl_view &v) {
  const uint64_t slots = (uint64_t)v.scene->count + 1 + v.scene->n_emit;
  const uint64_t tiles = (uint64_t)((v.width + 15) / 16) * ((v.height + 15) / 16);
  const uint64_t px = (uint64_t)v.width * v.height * (v.out_format == SL_OUT_RGBA16F ? 8 : 16);
  return px + slots * (48 + 4 * 4 + 8) + 8 * slots * 16 + tiles * 12;
}

// A lane whose completion check returned q (hipSuccess, or the error that ended it): its requests' results, with
// the batch's status (on an error, every request carries it and the server's error names it).
void finish_lane(sl_server *s, ServerLane &ln, hipError_t q) {
  std::vector<sl_frame_stats> st(ln.reqs.size());
  sl_status rc;
  if (q == hipSuccess) {
    rc = finish_batch(ln.ctx, st.data(), (uint32_t)st.size());
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
  s->in_flight_bytes -= ln.bytes;
  ln.reqs.clear();
  ln.bytes = 0;
  ln.busy = false;
  s->cv.notify_all();
}

// Admit from the head of the queue into one batch for a free lane; false when nothing was admitted.
bool launch_lane(sl_server *s, ServerLane &ln) {
  std::vector<ServerRequest> batch;
  uint64_t bytes = 0;
  {
    std::lock_guard<std::mutex> g(s->mu);
    while (!s->pending.empty() && batch.size() < s->opts.max_batch_views) {
      const ServerRequest &r = s->pending.front();
      if (s->in_flight_bytes + bytes + r.bytes > s->opts.hbm_budget && !(batch.empty() && s->in_flight_bytes == 0))
        break;                                   // wait for memory (an oversized request runs alone)
      bytes += r.bytes;
      batch.push_back(std::move(s->pending.front()));
      s->pending.pop_front();
    }
    if (batch.empty()) return false;
    s->in_flight_bytes += bytes;
  }
  std::vector<sl_view> views(batch.size());
  for (size_t k = 0; k < batch.size(); k++) {
    views[k] = batch[k].view;
    views[k].frame_bytes = batch[k].frame.data();
  }
  ln.t_start = s->now();
  ln.reqs = std::move(batch);
  ln.bytes = bytes;
  ln.busy = true;
  sl_status rc = ln.ctx->poisoned != SL_OK ? SL_ERR_POISONED : SL_OK;
  for (ServerRequest &r : ln.reqs) {             // ordered after what the owner had queued when it was submitted
    if (rc == SL_OK && hipStreamWaitEvent(ln.ctx->stream, r.ready, 0) != hipSuccess) rc = SL_ERR_DEVICE;
    (void)hipEventDestroy(r.ready);
    r.ready = nullptr;
  }
  if (rc == SL_OK) rc = encode_batch(ln.ctx, views.data(), (uint32_t)views.size(), 0u);
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
    s->in_flight_bytes -= ln.bytes;
    ln.reqs.clear();
    ln.bytes = 0;
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
  {
    FrameDev f;
    const sl_frame_opts o = {v->width, v->height, v->out_format, 0u};
    const sl_status rc = decode_frame(ctx, v->scene, v->frame_bytes, &o, &f, 4u);   // lanes composite 16x16 tiles
    if (rc != SL_OK) return rc;
    if (f.tiles_x * f.tiles_y > 65536u) return fail(ctx, SL_ERR_CAPACITY, "more than 65536 16x16 tiles in a view");
    if (f.n_nodes) return fail(ctx, SL_ERR_UNSUPPORTED, "a screen-space LOD frame renders alone (sl_frame_encode)");
  }
  ServerRequest r;
  r.view = *v;
  r.frame.assign((const uint8_t *)v->frame_bytes, (const uint8_t *)v->frame_bytes + kFrameBytes);
  r.bytes = request_bytes(*v);
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
