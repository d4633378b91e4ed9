// serve.hip -- the served form of the C ABI (pgvd_server_*): a request loop
// over whole-path octree builds.
//
// The server owns `slots` contexts; each has its own stream (default
// priority) and pooled device buffers that stay allocated across requests.
// The loop replays a trace in real time:
//   * a request joins the FIFO queue at its arrival time;
//   * the head of the queue is admitted when a slot is free and its device
//     footprint (what one build allocates in a pool, plus its points) fits
//     the HBM budget beside the requests in flight;
//   * admission enqueues, on the slot's stream, the copy of the points, the
//     whole build and a digest of the octree (sized on the device), and
//     records an event;
//   * completion is found by polling the slots' events, so a finished build
//     never waits behind another, and the slot and its bytes are free again.
// Builds on different slots overlap on the device.
#include <chrono>
#include <deque>
#include <thread>
#include <vector>

#include "capi_internal.hpp"

using namespace pgvd;
using namespace pgvd_capi;

namespace {

struct Slot {
  pgvd_ctx *ctx = nullptr;
  void *d_pts = nullptr;
  size_t pts_bytes = 0;
  uint64_t *d_digest = nullptr;
  uint64_t *h_digest = nullptr;   // pinned
  hipEvent_t ev = nullptr;
  int64_t req = -1;
  size_t bytes = 0;
};

}  // namespace

struct pgvd_server {
  int device = 0;
  size_t budget = 0;
  std::vector<Slot> slots;
};

namespace {

using Clock = std::chrono::steady_clock;

double since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

int admit(pgvd_server *srv, Slot &sl, pgvd_request &r) {
  pgvd_ctx *ctx = sl.ctx;
  const size_t pb = (size_t)r.n * r.dim * 4;
  if (pb > sl.pts_bytes) {
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
  (void)srv;
  return PGVD_OK;
}

// The build on this slot has completed: collect its results.
int complete(Slot &sl, pgvd_request &r) {
  pgvd_ctx *ctx = sl.ctx;
  pgvd_octree_result res;
  const bool refilled = (ctx->st_host->status & ST_CAPACITY) != 0;
  const int rc = pgvd_build_octree_finish(ctx, &res);
  r.status = rc;
  if (rc != PGVD_OK) return rc;
  r.m = res.m;
  r.octree_size = res.octree_size;
  if (refilled) {
    // the digest enqueued at admission saw the unfilled buffer
    const int stride = (1 << r.dim) + 1;
    launch_digest(res.octree, (size_t)res.octree_size * stride, sl.d_digest, true, ctx->stream);
    HIPCHK(hipMemcpyAsync(sl.h_digest, sl.d_digest, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  r.octree_digest = *sl.h_digest;
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
    if (rc == PGVD_OK && (hipMalloc((void **)&sl.d_digest, 8) != hipSuccess ||
                          hipHostMalloc((void **)&sl.h_digest, 8) != hipSuccess ||
                          hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess))
      rc = fail(PGVD_ERR_HIP, "server slot set-up failed");
    if (rc != PGVD_OK) {
      pgvd_server_destroy(srv);
      return rc;
    }
  }
  *out = srv;
  return PGVD_OK;
}

int pgvd_server_destroy(pgvd_server *srv) {
  if (!srv) return PGVD_OK;
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

int pgvd_server_run(pgvd_server *srv, pgvd_request *reqs, int64_t count, pgvd_server_stats *stats) {
  if (!srv || (!reqs && count)) return fail(PGVD_ERR_INVALID, "null argument");
  for (int64_t i = 0; i < count; ++i) {
    const pgvd_request &r = reqs[i];
    if (!r.host_pts || r.n <= 0 || (r.dim != 2 && r.dim != 3) || r.bits < 1 || r.bits * r.dim > 128)
      return fail(PGVD_ERR_INVALID, "request " + std::to_string(i) + " is malformed");
    if (i && r.arrival_s < reqs[i - 1].arrival_s)
      return fail(PGVD_ERR_INVALID, "requests must be sorted by arrival time");
    if (pgvd_build_footprint(r.n, r.dim, r.bits) > srv->budget)
      return fail(PGVD_ERR_CAPACITY, "request " + std::to_string(i) + " exceeds the HBM budget alone");
  }
  HIPCHK(hipSetDevice(srv->device));
  std::deque<int64_t> queue;
  std::vector<size_t> foot(count);
  for (int64_t i = 0; i < count; ++i) foot[i] = pgvd_build_footprint(reqs[i].n, reqs[i].dim, reqs[i].bits);
  int64_t next = 0, done = 0;
  size_t in_bytes = 0;
  int in_flight = 0, max_in = 0;
  size_t max_bytes = 0;
  int first_err = PGVD_OK;
  const Clock::time_point t0 = Clock::now();
  while (done < count) {
    bool progress = false;
    const double now = since(t0);
    while (next < count && reqs[next].arrival_s <= now) queue.push_back(next++);
    for (Slot &sl : srv->slots) {
      if (sl.req < 0) continue;
      const hipError_t q = hipEventQuery(sl.ev);
      if (q == hipErrorNotReady) continue;
      pgvd_request &r = reqs[sl.req];
      int rc = q == hipSuccess ? complete(sl, r) : fail(PGVD_ERR_HIP, hipGetErrorString(q));
      if (rc != PGVD_OK && first_err == PGVD_OK) first_err = rc;
      r.status = rc;
      r.done_s = since(t0);
      in_bytes -= sl.bytes;
      --in_flight;
      ++done;
      sl.req = -1;
      sl.bytes = 0;
      progress = true;
    }
    while (!queue.empty()) {
      const int64_t i = queue.front();
      if (in_bytes + foot[i] > srv->budget) break;
      Slot *free_slot = nullptr;
      for (Slot &sl : srv->slots)
        if (sl.req < 0) { free_slot = &sl; break; }
      if (!free_slot) break;
      queue.pop_front();
      pgvd_request &r = reqs[i];
      r.admit_s = since(t0);
      r.slot = (int)(free_slot - srv->slots.data());
      const int rc = admit(srv, *free_slot, r);
      if (rc != PGVD_OK) {
        r.status = rc;
        r.done_s = since(t0);
        ++done;
        if (first_err == PGVD_OK) first_err = rc;
        continue;
      }
      free_slot->req = i;
      free_slot->bytes = foot[i];
      in_bytes += foot[i];
      ++in_flight;
      if (in_flight > max_in) max_in = in_flight;
      if (in_bytes > max_bytes) max_bytes = in_bytes;
      progress = true;
    }
    if (!progress) {
      // nothing to do until a build completes or a request arrives
      if (next < count && queue.empty() && in_flight == 0) {
        const double wait = reqs[next].arrival_s - since(t0);
        if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
      } else {
        std::this_thread::yield();
      }
    }
  }
  if (stats) {
    stats->elapsed_s = since(t0) - (count ? reqs[0].arrival_s : 0.0);
    stats->max_in_flight = max_in;
    stats->max_in_flight_bytes = max_bytes;
    stats->completed = done;
  }
  return first_err;
}

}  // extern "C"
