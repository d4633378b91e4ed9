// serve.hip -- the served form of the C ABI (pgvd_server_*): octree builds
// requested online, admitted against an HBM budget and overlapped on the
// device.
//
// The server owns `slots` contexts; each has its own stream (default
// priority) and a pool of device buffers kept across requests.  An engine
// thread inside the server runs the loop:
//   * a submitted request joins a FIFO queue;
//   * the head of the queue is admitted when a slot is free and the device
//     memory the server would then hold -- every slot's pool and points
//     buffer as allocated (a slot grows to a request's footprint), plus the
//     results not yet released -- fits the budget; an idle slot's pool is
//     released when that makes a request fit;
//   * admission enqueues, on the slot's stream, the copy of the points, the
//     whole build (no host waits: one thread feeds every slot) and a digest
//     of the octree (sized on the device), and records an event;
//   * a finished build gets a result buffer from the server's cache of
//     released ones (sizes rounded up to a quarter octave, so steady serving
//     allocates and frees nothing), the octree is copied there on the slot's
//     stream, and once that copy is done the request is complete and its
//     slot free.  Cached buffers count against the budget and are freed
//     first when a request would not fit.
// Completion is found by polling the slots' events, so a finished build
// never waits behind another.  A result's octree stays on the device until
// pgvd_server_release.  pgvd_server_run replays a trace of requests in real
// time through the same engine.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "capi_internal.hpp"

using namespace pgvd;
using namespace pgvd_capi;

namespace {

using Clock = std::chrono::steady_clock;

struct Slot {
  pgvd_ctx *ctx = nullptr;
  void *d_pts = nullptr;
  size_t pts_bytes = 0;
  uint64_t *d_digest = nullptr;
  uint64_t *h_digest = nullptr;   // pinned
  hipEvent_t ev = nullptr;
  int64_t ticket = -1;
  int phase = 0;                  // 1: building, 2: copying the octree out
};

enum TicketState { T_QUEUED, T_RUNNING, T_DONE };

struct Ticket {
  const float *host_pts = nullptr;
  int64_t n = 0;
  int dim = 0, bits = 0;
  size_t foot = 0;
  TicketState state = T_QUEUED;
  double arrival = 0, admit = 0, done = 0;
  int status = PGVD_OK;
  int slot = -1;
  int64_t m = 0, octree_size = 0;
  uint64_t digest = 0;
  void *result = nullptr;         // device octree, octree_size records
  size_t result_bytes = 0;        // capacity of that buffer
};

// Capacity of a result buffer for `bytes`: rounded up to a quarter octave
// (at most 1/4 more than asked), at least 64 KiB.
size_t result_class(size_t bytes) {
  size_t c = bytes < (size_t(64) << 10) ? (size_t(64) << 10) : bytes;
  int top = 63 - __builtin_clzll(c);
  const size_t q = top >= 2 ? size_t(1) << (top - 2) : 1;
  return (c + q - 1) / q * q;
}

size_t pool_bytes(const pgvd_ctx *ctx) { return ctx->pool_bytes; }

}  // namespace

struct pgvd_server {
  int device = 0;
  size_t budget = 0;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv;       // new work for the engine, or a completion
  std::thread engine;
  bool stop = false;
  int64_t next_ticket = 0;
  std::unordered_map<int64_t, Ticket> tickets;
  std::deque<int64_t> queue;
  size_t held_results = 0;          // bytes of result buffers, held or cached
  std::multimap<size_t, void *> spare_results;   // released, by capacity
  size_t spare_bytes = 0;
  Clock::time_point t0 = Clock::now();
  size_t free0 = 0;                 // device free memory at creation
  // statistics since the last reset
  int in_flight = 0, max_in_flight = 0;
  size_t max_reserved = 0, min_free = 0;
  int64_t completed = 0;
  int first_err = PGVD_OK;
};

namespace {

double since(const pgvd_server *srv) {
  return std::chrono::duration<double>(Clock::now() - srv->t0).count();
}

// Device bytes the server holds: every slot's pool and points buffer as
// allocated, and the results not yet released.
size_t reserved(const pgvd_server *srv) {
  size_t b = srv->held_results;
  for (const Slot &sl : srv->slots) b += pool_bytes(sl.ctx) + sl.pts_bytes;
  return b;
}

void note_memory(pgvd_server *srv) {
  const size_t r = reserved(srv);
  if (r > srv->max_reserved) srv->max_reserved = r;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr < srv->min_free) srv->min_free = fr;
}

// Frees an idle slot's pool and points (its stream has no work left).
int release_slot_memory(Slot &sl) {
  pgvd_ctx *ctx = sl.ctx;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (auto &kv : ctx->pool)
    if (kv.second.p) HIPCHK(hipFree(kv.second.p));
  ctx->pool.clear();
  ctx->pool_bytes = 0;
  ctx->octree_cap = 0;
  ctx->octree_stride = 0;
  ctx->last.valid = false;
  if (sl.d_pts) HIPCHK(hipFree(sl.d_pts));
  sl.d_pts = nullptr;
  sl.pts_bytes = 0;
  return PGVD_OK;
}

// Frees the cached result buffers (their last readers have released them).
int free_spare_results(pgvd_server *srv) {
  for (auto &kv : srv->spare_results) HIPCHK(hipFree(kv.second));
  srv->held_results -= srv->spare_bytes;
  srv->spare_results.clear();
  srv->spare_bytes = 0;
  return PGVD_OK;
}

int admit(pgvd_server *srv, Slot &sl, Ticket &r) {
  (void)srv;
  pgvd_ctx *ctx = sl.ctx;
  const size_t pb = (size_t)r.n * r.dim * 4;
  if (pb > sl.pts_bytes) {
    // the slot is idle: its stream has finished with the old buffer
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (sl.d_pts) HIPCHK(hipFree(sl.d_pts));
    sl.d_pts = nullptr;
    sl.pts_bytes = 0;
    HIPCHK(hipMalloc(&sl.d_pts, pb + pb / 8));
    sl.pts_bytes = pb + pb / 8;
  }
  HIPCHK(hipMemcpyAsync(sl.d_pts, r.host_pts, pb, hipMemcpyHostToDevice, ctx->stream));
  // one thread enqueues every slot's builds: none may wait for the device
  pgvd_build_opts o;
  std::memset(&o, 0, sizeof(o));
  o.no_host_wait = 1;
  CHK(pgvd_build_octree(ctx, sl.d_pts, r.n, r.dim, r.bits, nullptr, &o, nullptr));
  // digest of the octree as built; the node count is read on the device
  const int stride = (1 << r.dim) + 1;
  launch_digest_count(ctx->pool["octree"].p, &ctx->st->octree_size, (uint32_t)stride,
                      (size_t)ctx->octree_cap * stride, sl.d_digest, ctx->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(sl.h_digest, sl.d_digest, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipEventRecord(sl.ev, ctx->stream));
  return PGVD_OK;
}

// The build on this slot has completed: read its sizes and digest, and copy
// its octree into a result buffer of its own (asynchronously).
int collect(pgvd_server *srv, Slot &sl, Ticket &r) {
  pgvd_ctx *ctx = sl.ctx;
  pgvd_octree_result res;
  // a refill, or the order-exact path of a malformed tree, rewrites the
  // octree in finish, after the digest enqueued at admission
  const bool refilled = (ctx->st_host->status & ST_CAPACITY) != 0 || ctx->st_host->malformed;
  CHK(pgvd_build_octree_finish(ctx, &res));
  r.m = res.m;
  r.octree_size = res.octree_size;
  const int stride = (1 << r.dim) + 1;
  if (refilled) {
    launch_digest(res.octree, (size_t)res.octree_size * stride, sl.d_digest, true, ctx->stream);
    HIPCHK(hipMemcpyAsync(sl.h_digest, sl.d_digest, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  r.digest = *sl.h_digest;
  const size_t bytes = (size_t)res.octree_size * stride * 4;
  if (bytes) {
    const size_t cls = result_class(bytes);
    // the smallest cached buffer that holds it, if under twice the class
    auto it = srv->spare_results.lower_bound(cls);
    if (it != srv->spare_results.end() && it->first <= 2 * cls) {
      r.result = it->second;
      r.result_bytes = it->first;
      srv->spare_results.erase(it);
      srv->spare_bytes -= r.result_bytes;
    } else {
      if (hipMalloc(&r.result, cls) != hipSuccess) {
        (void)hipGetLastError();
        r.result = nullptr;
        return fail(PGVD_ERR_ALLOC, "no device memory for a served result");
      }
      srv->held_results += cls;
      r.result_bytes = cls;
    }
    launch_copy(r.result, res.octree, bytes, ctx->stream);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(sl.ev, ctx->stream));
  return PGVD_OK;
}

void finish_ticket(pgvd_server *srv, Slot &sl, Ticket &r, int status) {
  r.status = status;
  r.done = since(srv);
  r.state = T_DONE;
  if (status != PGVD_OK && srv->first_err == PGVD_OK) srv->first_err = status;
  sl.ticket = -1;
  sl.phase = 0;
  --srv->in_flight;
  ++srv->completed;
}

// One round of the engine under the lock; true when something happened.
bool engine_round(pgvd_server *srv) {
  bool progress = false;
  for (Slot &sl : srv->slots) {
    if (sl.ticket < 0) continue;
    const hipError_t q = hipEventQuery(sl.ev);
    if (q == hipErrorNotReady) continue;
    Ticket &r = srv->tickets[sl.ticket];
    progress = true;
    if (q != hipSuccess) {
      finish_ticket(srv, sl, r, fail(PGVD_ERR_HIP, hipGetErrorString(q)));
      continue;
    }
    if (sl.phase == 1) {
      const int rc = collect(srv, sl, r);
      if (rc != PGVD_OK) finish_ticket(srv, sl, r, rc);
      else sl.phase = 2;
    } else {
      finish_ticket(srv, sl, r, PGVD_OK);
    }
  }
  while (!srv->queue.empty()) {
    const int64_t id = srv->queue.front();
    Ticket &r = srv->tickets[id];
    // the free slot that needs to grow least
    Slot *pick = nullptr;
    size_t pick_grow = 0;
    for (Slot &sl : srv->slots) {
      if (sl.ticket >= 0) continue;
      const size_t have = pool_bytes(sl.ctx) + sl.pts_bytes;
      const size_t grow = r.foot > have ? r.foot - have : 0;
      if (!pick || grow < pick_grow) { pick = &sl; pick_grow = grow; }
    }
    if (!pick) break;
    size_t need = reserved(srv) + pick_grow;
    // free cached results, then idle pools (largest first), until the
    // request fits
    if (need > srv->budget && srv->spare_bytes) {
      if (free_spare_results(srv) != PGVD_OK) break;
      need = reserved(srv) + pick_grow;
    }
    while (need > srv->budget) {
      Slot *victim = nullptr;
      size_t vb = 0;
      for (Slot &sl : srv->slots) {
        if (sl.ticket >= 0 || &sl == pick) continue;
        const size_t b = pool_bytes(sl.ctx) + sl.pts_bytes;
        if (b > vb) { victim = &sl; vb = b; }
      }
      if (!victim) {
        // the chosen slot's own pool goes last: it regrows to the footprint
        const size_t have = pool_bytes(pick->ctx) + pick->pts_bytes;
        if (have && reserved(srv) - have + r.foot <= srv->budget) victim = pick;
        else break;
      }
      if (release_slot_memory(*victim) != PGVD_OK) break;
      const size_t have = pool_bytes(pick->ctx) + pick->pts_bytes;
      need = reserved(srv) + (r.foot > have ? r.foot - have : 0);
    }
    if (need > srv->budget) break;   // wait for results to be released or slots to finish
    srv->queue.pop_front();
    r.admit = since(srv);
    r.slot = (int)(pick - srv->slots.data());
    r.state = T_RUNNING;
    ++srv->in_flight;
    if (srv->in_flight > srv->max_in_flight) srv->max_in_flight = srv->in_flight;
    pick->ticket = id;
    pick->phase = 1;
    const int rc = admit(srv, *pick, r);
    if (rc != PGVD_OK) finish_ticket(srv, *pick, r, rc);
    note_memory(srv);
    progress = true;
  }
  if (progress) srv->cv.notify_all();
  return progress;
}

void engine_main(pgvd_server *srv) {
  (void)hipSetDevice(srv->device);
  std::unique_lock<std::mutex> lk(srv->mu);
  for (;;) {
    const bool progress = engine_round(srv);
    const bool busy = srv->in_flight > 0;
    if (srv->stop && !busy && srv->queue.empty()) break;
    if (progress) continue;
    if (busy) {
      // builds on the device: poll again soon, letting callers in
      lk.unlock();
      std::this_thread::yield();
      lk.lock();
    } else {
      // nothing in flight: wait for a submission (or a release freeing
      // budget for the queue's head)
      srv->cv.wait_for(lk, std::chrono::milliseconds(5));
    }
  }
}

int check_request(const float *pts, int64_t n, int dim, int bits) {
  if (!pts || n <= 0 || (dim != 2 && dim != 3) || bits < 1 || bits * dim > 128)
    return fail(PGVD_ERR_INVALID, "malformed request");
  if (n >= (int64_t(1) << 30) - 1) return fail(PGVD_ERR_TOO_LARGE, "n must be below 2^30 - 1");
  return PGVD_OK;
}

}  // namespace

extern "C" {

size_t pgvd_build_footprint(int64_t n, int dim, int bits) {
  return plain_footprint_bytes(n, dim, bits, false);
}

int pgvd_host_alloc(size_t bytes, void **out) {
  if (!out) return fail(PGVD_ERR_INVALID, "null argument");
  *out = nullptr;
  if (hipHostMalloc(out, bytes ? bytes : 1) != hipSuccess)
    return fail(PGVD_ERR_ALLOC, "hipHostMalloc failed");
  return PGVD_OK;
}

int pgvd_host_free(void *p) {
  if (p) HIPCHK(hipHostFree(p));
  return PGVD_OK;
}

int pgvd_server_create(int device, int slots, size_t hbm_budget_bytes, pgvd_server **out) {
  if (!out || slots < 1) return fail(PGVD_ERR_INVALID, "need an output pointer and at least one slot");
  *out = nullptr;
  pgvd_server *srv = new pgvd_server();
  srv->device = device;
  srv->budget = hbm_budget_bytes;
  srv->slots.resize(slots);
  for (Slot &sl : srv->slots) {
    int rc = pgvd_ctx_create(device, &sl.ctx);
    if (rc == PGVD_OK) sl.ctx->own_stream_only = true;
    if (rc == PGVD_OK && (hipMalloc((void **)&sl.d_digest, 8) != hipSuccess ||
                          hipHostMalloc((void **)&sl.h_digest, 8) != hipSuccess ||
                          hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess))
      rc = fail(PGVD_ERR_HIP, "server slot set-up failed");
    if (rc != PGVD_OK) {
      pgvd_server_destroy(srv);
      return rc;
    }
  }
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) == hipSuccess) srv->free0 = fr;
  srv->min_free = srv->free0;
  srv->t0 = Clock::now();
  srv->engine = std::thread(engine_main, srv);
  *out = srv;
  return PGVD_OK;
}

int pgvd_server_destroy(pgvd_server *srv) {
  if (!srv) return PGVD_OK;
  if (srv->engine.joinable()) {
    {
      std::lock_guard<std::mutex> g(srv->mu);
      srv->stop = true;
      srv->queue.clear();          // requests not admitted are dropped
    }
    srv->cv.notify_all();
    srv->engine.join();
  }
  for (auto &kv : srv->tickets)
    if (kv.second.result) (void)hipFree(kv.second.result);
  (void)free_spare_results(srv);
  for (Slot &sl : srv->slots) {
    if (!sl.ctx) continue;
    (void)hipSetDevice(sl.ctx->device);
    (void)hipStreamSynchronize(sl.ctx->stream);
    if (sl.d_pts) (void)hipFree(sl.d_pts);
    if (sl.d_digest) (void)hipFree(sl.d_digest);
    if (sl.h_digest) (void)hipHostFree(sl.h_digest);
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    pgvd_ctx_destroy(sl.ctx);
  }
  delete srv;
  return PGVD_OK;
}

int pgvd_server_submit(pgvd_server *srv, const float *host_pts, int64_t n, int dim, int bits,
                       int64_t *ticket) {
  if (!srv || !ticket) return fail(PGVD_ERR_INVALID, "null argument");
  CHK(check_request(host_pts, n, dim, bits));
  const size_t foot = pgvd_build_footprint(n, dim, bits);
  if (foot > srv->budget) return fail(PGVD_ERR_CAPACITY, "request exceeds the HBM budget alone");
  {
    std::lock_guard<std::mutex> g(srv->mu);
    const int64_t id = srv->next_ticket++;
    Ticket &t = srv->tickets[id];
    t.host_pts = host_pts;
    t.n = n;
    t.dim = dim;
    t.bits = bits;
    t.foot = foot;
    t.arrival = since(srv);
    srv->queue.push_back(id);
    *ticket = id;
  }
  srv->cv.notify_all();
  return PGVD_OK;
}

int pgvd_server_poll(pgvd_server *srv, int64_t ticket, int wait, pgvd_served_result *out) {
  if (!srv || !out) return fail(PGVD_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->tickets.find(ticket);
  if (it == srv->tickets.end()) return fail(PGVD_ERR_INVALID, "unknown or released ticket");
  if (it->second.state != T_DONE) {
    if (!wait) return PGVD_NOT_READY;
    srv->cv.wait(lk, [&] { return it->second.state == T_DONE; });
  }
  const Ticket &t = it->second;
  std::memset(out, 0, sizeof(*out));
  out->status = t.status;
  out->n = t.n;
  out->dim = t.dim;
  out->bits = t.bits;
  out->m = t.m;
  out->octree_size = t.octree_size;
  out->octree_digest = t.digest;
  out->arrival_s = t.arrival;
  out->admit_s = t.admit;
  out->done_s = t.done;
  out->slot = t.slot;
  out->octree = static_cast<const int32_t *>(t.result);
  return PGVD_OK;
}

int pgvd_server_release(pgvd_server *srv, int64_t ticket) {
  if (!srv) return fail(PGVD_ERR_INVALID, "null argument");
  {
    std::lock_guard<std::mutex> g(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return fail(PGVD_ERR_INVALID, "unknown or released ticket");
    if (it->second.state != T_DONE) return fail(PGVD_ERR_INVALID, "the request is not complete");
    if (it->second.result) {
      // kept for a later result of its size class
      srv->spare_results.emplace(it->second.result_bytes, it->second.result);
      srv->spare_bytes += it->second.result_bytes;
    }
    srv->tickets.erase(it);
  }
  srv->cv.notify_all();   // the freed bytes may admit the queue's head
  return PGVD_OK;
}

int pgvd_server_run(pgvd_server *srv, pgvd_request *reqs, int64_t count, pgvd_server_stats *stats) {
  if (!srv || (!reqs && count)) return fail(PGVD_ERR_INVALID, "null argument");
  for (int64_t i = 0; i < count; ++i) {
    const pgvd_request &r = reqs[i];
    if (check_request(r.host_pts, r.n, r.dim, r.bits) != PGVD_OK)
      return fail(PGVD_ERR_INVALID, "request " + std::to_string(i) + " is malformed");
    if (i && r.arrival_s < reqs[i - 1].arrival_s)
      return fail(PGVD_ERR_INVALID, "requests must be sorted by arrival time");
    if (pgvd_build_footprint(r.n, r.dim, r.bits) > srv->budget)
      return fail(PGVD_ERR_CAPACITY, "request " + std::to_string(i) + " exceeds the HBM budget alone");
  }
  {
    std::lock_guard<std::mutex> g(srv->mu);
    srv->max_in_flight = srv->in_flight;
    srv->max_reserved = reserved(srv);
    srv->min_free = srv->free0;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) srv->min_free = fr;
    srv->completed = 0;
    srv->first_err = PGVD_OK;
  }
  const double base = since(srv);
  std::vector<int64_t> ids((size_t)count);
  int first = PGVD_OK;
  double last_done = base;
  int64_t collected = 0;
  // results are collected (and their octrees released) as they complete,
  // in submission order, between submissions
  auto collect_one = [&](int64_t i, int wait) -> int {
    pgvd_served_result res;
    const int rc = pgvd_server_poll(srv, ids[i], wait, &res);
    if (rc != PGVD_OK) return rc;
    pgvd_request &r = reqs[i];
    r.status = res.status;
    r.m = res.m;
    r.octree_size = res.octree_size;
    r.octree_digest = res.octree_digest;
    r.admit_s = res.admit_s - base;
    r.done_s = res.done_s - base;
    r.slot = res.slot;
    if (res.done_s > last_done) last_done = res.done_s;
    if (res.status != PGVD_OK && first == PGVD_OK) first = res.status;
    return pgvd_server_release(srv, ids[i]);
  };
  for (int64_t i = 0; i < count; ++i) {
    // submit each request at its arrival time
    for (;;) {
      const double wait = base + reqs[i].arrival_s - since(srv);
      if (wait <= 0) break;
      if (collected < i) {
        const int rc = collect_one(collected, 0);
        if (rc == PGVD_OK) { ++collected; continue; }
        if (rc != PGVD_NOT_READY) return rc;
      }
      if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(
                           std::min(wait - 1e-4, 1e-4)));
      else std::this_thread::yield();
    }
    CHK(pgvd_server_submit(srv, reqs[i].host_pts, reqs[i].n, reqs[i].dim, reqs[i].bits, &ids[i]));
  }
  for (; collected < count; ++collected) CHK(collect_one(collected, 1));
  if (stats) {
    std::lock_guard<std::mutex> g(srv->mu);
    stats->elapsed_s = last_done - base - (count ? reqs[0].arrival_s : 0.0);
    stats->max_in_flight = srv->max_in_flight;
    stats->max_in_flight_bytes = srv->max_reserved;
    stats->completed = count;
    stats->max_device_bytes = srv->free0 > srv->min_free ? srv->free0 - srv->min_free : 0;
  }
  return first;
}

}  // extern "C"
