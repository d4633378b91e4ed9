// capi_internal.hpp -- context, error reporting and the keyed device-buffer
// pool shared by the C-ABI translation units of libpgvd.so.  Internal.
#pragma once

#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/pgvd.h"
#include "kernels.hpp"

namespace pgvd_capi {

using pgvd::RunState;

inline thread_local std::string g_err;
// Stage whose next launch check on this thread reports a failure
// (pgvd_debug_fail_launch); empty: none.
inline thread_local std::string g_fail_launch;

inline int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      return fail(PGVD_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));          \
  } while (0)

#define CHK(x)                    \
  do {                            \
    int r_ = (x);                 \
    if (r_ != PGVD_OK) return r_; \
  } while (0)

inline float fdec_host(uint32_t e) {
  const uint32_t b = (e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e;
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}

inline int key_bytes_for(int mbits) { return mbits <= 32 ? 4 : (mbits <= 64 ? 8 : 16); }

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// opencl/Kernels.cpp:20
inline int ref_next_pow2(int num) {
  const int p = (int)pow(2, ceil(log(num) / log(2)));
  return p > 8 ? p : 8;
}

struct PoolBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct TimingRec {
  std::string stage;
  hipEvent_t a, b;
};

struct LastBuild {
  bool valid = false;
  int64_t n = 0;
  int dim = 0, bits = 0, mbits = 0, kb = 0;
  bool labels = false;
  void *keys_u = nullptr;
  void *brt = nullptr;
  int32_t *parent = nullptr;
  uint32_t *ps = nullptr;
  uint32_t *lsx = nullptr;
  uint32_t *chunks = nullptr;
  uint32_t *defer = nullptr;
  uint32_t *conflict_bits = nullptr;
  uint64_t *lb_splits = nullptr;   // look-back words of the order-exact scan (zero)
  bool exact_deferred = false;     // the order-exact path waits for finish
  hipStream_t stream = nullptr;
};

}  // namespace pgvd_capi

struct pgvd_ctx {
  using PoolBuf = pgvd_capi::PoolBuf;
  using TimingRec = pgvd_capi::TimingRec;
  using LastBuild = pgvd_capi::LastBuild;
  int device = 0;
  hipStream_t stream = nullptr;
  std::unordered_map<std::string, PoolBuf> pool;
  pgvd::RunState *st = nullptr;       // device
  pgvd::RunState *st_host = nullptr;  // pinned mirror
  hipEvent_t done = nullptr;
  hipEvent_t sync_ev = nullptr;       // mid-build waits of the bucket sort
  int timing = 0;               // 0 off, 1 every stage, 2 sort span only
  std::vector<TimingRec> recs;
  std::vector<hipEvent_t> spare;
  LastBuild last;
  int64_t octree_cap = 0;       // nodes the pooled octree buffer holds
  int octree_stride = 0;        // ints per node of that buffer
  bool own_stream_only = false; // every call on the context uses its stream
  size_t pool_bytes = 0;        // device bytes the pool holds
  size_t pool_limit = 0;        // most it may hold (0: no limit)
};

namespace pgvd_capi {

inline int use_device(pgvd_ctx *ctx) {
  HIPCHK(hipSetDevice(ctx->device));
  return PGVD_OK;
}

inline hipStream_t pick(pgvd_ctx *ctx, pgvd_stream s) {
  return s ? static_cast<hipStream_t>(s) : ctx->stream;
}

// Keyed pool (CLFW::get, CLFW/src/clfw.cpp:212-226): reuse when large
// enough, else reallocate with headroom.
inline int pool_get(pgvd_ctx *ctx, const char *key, size_t bytes, void **out) {
  if (bytes == 0) bytes = 256;
  PoolBuf &b = ctx->pool[key];
  if (b.bytes < bytes) {
    if (b.p) {
      // the old buffer may be in use on any stream the caller passed (only
      // on the context's own where it is the only one: the served form's
      // slots, whose neighbours keep running)
      if (ctx->own_stream_only) HIPCHK(hipStreamSynchronize(ctx->stream));
      else HIPCHK(hipDeviceSynchronize());
      HIPCHK(hipFree(b.p));
      ctx->pool_bytes -= b.bytes;
      b.p = nullptr;
      b.bytes = 0;
    }
    const size_t want = align256(bytes + bytes / 8);
    if (ctx->pool_limit && ctx->pool_bytes + want > ctx->pool_limit)
      return fail(PGVD_ERR_ALLOC, std::string("pool buffer ") + key + " of " +
                                      std::to_string(want) + " bytes would pass the pool limit of " +
                                      std::to_string(ctx->pool_limit) + " bytes");
    if (hipMalloc(&b.p, want) != hipSuccess) {
      (void)hipGetLastError();   // not left for the next launch check to report
      b.p = nullptr;
      return fail(PGVD_ERR_ALLOC, std::string("hipMalloc failed for pool buffer ") + key);
    }
    b.bytes = want;
    ctx->pool_bytes += want;
  }
  *out = b.p;
  return PGVD_OK;
}

inline hipEvent_t take_event(pgvd_ctx *ctx) {
  if (!ctx->spare.empty()) {
    hipEvent_t e = ctx->spare.back();
    ctx->spare.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Records HIP events around a stage's launches when timing is on.
struct Stage {
  pgvd_ctx *ctx;
  hipStream_t s;
  hipEvent_t a = nullptr;
  const char *name;
  Stage(pgvd_ctx *c, const char *nm, hipStream_t st, int mode = 1) : ctx(c), s(st), name(nm) {
    if (ctx && ctx->timing == mode) {
      a = take_event(ctx);
      if (a) (void)hipEventRecord(a, s);
    }
  }
  ~Stage() {
    if (a) {
      hipEvent_t b = take_event(ctx);
      if (b) {
        (void)hipEventRecord(b, s);
        ctx->recs.push_back({name, a, b});
      }
    }
  }
};

// Device bytes a whole-path build of n points takes in a context's pool
// (every pooled buffer with its headroom, plus the input points).
size_t plain_footprint_bytes(int64_t n, int dim, int bits, bool labels);

}  // namespace pgvd_capi
