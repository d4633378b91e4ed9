// capi.hip -- the C ABI of libpgvd.so (include/pgvd.h): context, keyed
// device-buffer pool, stage entry points and the whole-path pipeline.
//
// The pipeline mirrors Kernels::BuildOctree_p (opencl/Kernels.cpp:413-429)
// with the bounding box and quantization of Karras::Quantize
// (Karras.cpp:69-99) in front, but keeps every intermediate on the device:
// the reference's two blocking reads (unique count, :214; octree size,
// :347) are replaced by device-side counts that later kernels read, so the
// whole path is one stream of launches and one small readback at the end.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/pgvd.h"
#include "kernels.hpp"

using namespace pgvd;

#include <type_traits>

#include "capi_internal.hpp"

using namespace pgvd_capi;

namespace {

// Events around each onesweep pass, recorded as separate "sort_pass" stage
// entries so that a per-launch average can be formed.
struct PassTimer {
  pgvd_ctx *ctx;
  std::vector<hipEvent_t> ev;
  PassTimer(pgvd_ctx *c, int passes) : ctx(c) {
    if (!ctx || ctx->timing != 1) return;
    for (int i = 0; i < 2 * passes; ++i) {
      hipEvent_t e = take_event(ctx);
      if (!e) { ev.clear(); return; }
      ev.push_back(e);
    }
  }
  hipEvent_t *events() { return ev.empty() ? nullptr : ev.data(); }
  ~PassTimer() {
    for (size_t i = 0; i + 1 < ev.size(); i += 2) ctx->recs.push_back({"sort_pass", ev[i], ev[i + 1]});
  }
};

int launch_status(const char *what) {
  if (!g_fail_launch.empty() && g_fail_launch == what) {
    g_fail_launch.clear();
    return fail(PGVD_ERR_HIP, std::string(what) + ": injected launch failure");
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(PGVD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return PGVD_OK;
}

// Workspace that must be zero at the start of a run, as one block.
struct ZeroLayout {
  size_t hist = 0, sort_lb = 0, lb_unique = 0, lb_splits = 0, lb_scan = 0, lb_brt = 0,
         cbits = 0, msd_hist = 0, total = 0;
};

ZeroLayout zero_layout(int64_t n, int kb, int passes, bool labels) {
  ZeroLayout z;
  size_t off = 0;
  z.hist = off;
  off += align256((size_t)kMaxPasses * 256 * 4);
  z.lb_unique = off;
  off += align256((size_t)scan_tiles(n + 1) * 8);
  z.lb_splits = off;
  off += align256((size_t)scan_tiles(n + 1) * 8);
  z.lb_scan = off;
  off += align256((size_t)scan_tiles(n + 1) * 8);
  z.lb_brt = off;
  off += align256((size_t)scan_tiles(n + 1) * 8);
  z.cbits = off;
  if (labels) off += align256((size_t)((n + 1 + 31) / 32) * 4);
  z.msd_hist = off;
  off += align256((size_t)kBkMaxBuckets * 4);
  // the sort passes' look-back words last: beside the bucket sort they are
  // zeroed only if the passes run (launch_histogram), not at the start
  z.sort_lb = off;
  const int64_t tiles = onesweep_tiles(kb, labels, n > 0 ? n : 1);
  off += align256((size_t)passes * (size_t)tiles * 256 * 4);
  z.total = off;
  return z;
}

}  // namespace

size_t pgvd_capi::plain_footprint_bytes(int64_t n, int dim, int bits, bool labels) {
  const int mbits = dim * bits;
  const int kb = key_bytes_for(mbits);
  const int64_t nk = n + 1;
  const size_t node = kb == 4 ? sizeof(BrtNodeT<uint32_t>)
                              : (kb == 8 ? sizeof(BrtNodeT<uint64_t>) : sizeof(BrtNodeT<Key128>));
  const int64_t oct_nodes = std::max<int64_t>(2 * nk, 1024);
  size_t b = 0;
  b += 2 * (size_t)nk * kb + 32;                         // keys_a, keys_b
  if (labels) b += 2 * (size_t)n * 4;                    // vals_a, vals_b
  b += (size_t)nk * node;                                // brt
  b += 3 * (size_t)nk * 4;                               // ps, parent, defer
  b += kBboxPartBytes;                                   // bbox_part
  b += kb == 4 ? bucket_ws_bytes<uint32_t>(n)            // the bucket sort's workspace
                : (kb == 8 ? bucket_ws_bytes<uint64_t>(n) : bucket_ws_bytes<Key128>(n));
  b += (size_t)((nk + kBlock - 1) / kBlock) * 4;         // chunks
  b += zero_layout(n, kb, (mbits + 7) / 8, labels).total;
  b += (size_t)oct_nodes * ((1u << dim) + 1) * 4;        // octree
  if (labels) b += (size_t)oct_nodes * 4;                // node_conflict
  b += b / 8;                                            // pool headroom
  return b + (size_t)n * dim * 4;                        // input points
}

namespace {

int copy_state_and_wait(pgvd_ctx *ctx, hipStream_t s) {
  HIPCHK(hipMemcpyAsync(ctx->st_host, ctx->st, sizeof(RunState), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return PGVD_OK;
}

int status_error(uint32_t st) {
  if (st & ST_NEG_COORD) return fail(PGVD_ERR_BAD_COORD, "bad coordinate: a point quantized below zero");
  if (st & ST_DUP_KEY) return fail(PGVD_ERR_DUP_KEY, "BRT input has equal neighbouring keys");
  if (st & (ST_STACK | ST_WALK)) return fail(PGVD_ERR_INTERNAL, "octree walk exceeded its bound");
  if (st & ST_TOO_LARGE) return fail(PGVD_ERR_TOO_LARGE, "count exceeds 2^30");
  return PGVD_OK;
}

int check_dims(int dim, int bits) {
  if (dim != 2 && dim != 3) return fail(PGVD_ERR_INVALID, "dim must be 2 or 3");
  if (bits < 1 || bits * dim > 128) return fail(PGVD_ERR_INVALID, "bits must satisfy 1 <= bits*dim <= 128");
  return PGVD_OK;
}

int check_n(int64_t n) {
  if (n <= 0) return fail(PGVD_ERR_EMPTY, "Zero points not supported");
  if (n >= (int64_t(1) << 30) - 1) return fail(PGVD_ERR_TOO_LARGE, "n must be below 2^30 - 1");
  return PGVD_OK;
}

template <typename K>
size_t brt_node_bytes() { return sizeof(BrtNodeT<K>); }

// ---------------------------------------------------------------------------
// whole path
// ---------------------------------------------------------------------------
template <typename K>
int ensure_octree(pgvd_ctx *ctx, int dim, int64_t nodes, bool labels, int32_t **oct,
                  uint32_t **nconf) {
  const int stride = (1 << dim) + 1;
  void *p = nullptr;
  CHK(pool_get(ctx, "octree", (size_t)nodes * stride * 4, &p));
  *oct = static_cast<int32_t *>(p);
  ctx->octree_cap = (int64_t)(ctx->pool["octree"].bytes / ((size_t)stride * 4));
  ctx->octree_stride = stride;
  *nconf = nullptr;
  if (labels) {
    CHK(pool_get(ctx, "node_conflict", (size_t)ctx->octree_cap * 4, &p));
    *nconf = static_cast<uint32_t *>(p);
  }
  return PGVD_OK;
}

// The device buffers of one whole-path build.
template <typename K>
struct BuildBufs {
  K *kA = nullptr, *kB = nullptr;          // sort ping-pong, 16 bytes after a front pad
  uint32_t *vA = nullptr, *vB = nullptr;   // labels ping-pong
  BrtNodeT<K> *brt = nullptr;
  uint32_t *ps = nullptr;
  int32_t *parent = nullptr;
  uint32_t *defer = nullptr, *chunks = nullptr;
  char *zb = nullptr;                      // the block zeroed at the build's start
  ZeroLayout z;
  float *bbox_part = nullptr;
  int bk_b = -1;                           // bucket sort's level-0 bits (-1: LSD passes)
  BucketMem bk;                            // the bucket sort's workspace
};

// What a build leaves for the caller (the unique keys, the packed local
// splits and the conflict bits live inside the carved buffers).
template <typename K>
struct BuildOut {
  K *uniq = nullptr;
  uint32_t *lsx = nullptr;
  uint32_t *cbits = nullptr;
};

// Takes every buffer of one build from get(key, bytes, &ptr): the
// context's keyed pool, or the allocation-free entry's static workspace.
template <typename K, typename Get>
int carve_build(Get &&get, const void *pts, int64_t n, int dim, int bits, bool lab,
                const pgvd_build_opts *opts, bool bucket_ok, BuildBufs<K> &b) {
  const int mbits = dim * bits;
  const int kb = (int)sizeof(K);
  const int passes = (mbits + 7) / 8;
  const int64_t nk = n + 1;
  void *p;
  // Each key buffer keeps 16 bytes before its keys: the distinct keys are
  // compacted in place from the sorted ones, one slot to the left when a
  // zero key is prepended.
  constexpr int64_t kPad = 16 / (int64_t)sizeof(K);
  CHK(get("keys_a", (size_t)(nk + kPad) * kb, &p));
  b.kA = static_cast<K *>(p) + kPad;
  CHK(get("keys_b", (size_t)(nk + kPad) * kb, &p));
  b.kB = static_cast<K *>(p) + kPad;
  if (lab) {
    CHK(get("vals_a", (size_t)n * 4, &p));
    b.vA = static_cast<uint32_t *>(p);
    CHK(get("vals_b", (size_t)n * 4, &p));
    b.vB = static_cast<uint32_t *>(p);
  }
  CHK(get("brt", (size_t)nk * brt_node_bytes<K>(), &p));
  b.brt = static_cast<BrtNodeT<K> *>(p);
  CHK(get("ps", (size_t)nk * 4, &p));
  b.ps = static_cast<uint32_t *>(p);
  CHK(get("parent", (size_t)nk * 4, &p));
  b.parent = static_cast<int32_t *>(p);
  CHK(get("defer", (size_t)nk * 4, &p));
  b.defer = static_cast<uint32_t *>(p);
  CHK(get("chunks", (size_t)((nk + kBlock - 1) / kBlock) * 4, &p));
  b.chunks = static_cast<uint32_t *>(p);
  const bool quantized = opts && opts->quantized_input;
  const bool custom = opts && opts->use_custom_bb;
  // float points (vectorised encode, which builds the bucket histogram) on
  // a context (the sort's levels are decided on the host): bucket sort;
  // otherwise the LSD passes
  b.bk_b = -1;
  if (bucket_ok && !quantized && encode_vectorised(static_cast<const float *>(pts), b.kA))
    b.bk_b = std::min(bucket_bits_for<K>(n), std::min(mbits, kBkMaxBits));
  // the LSD passes' look-back words stay for keys of at most 64 bits (the
  // bucket sort hands clustered keys to the passes)
  b.z = zero_layout(n, kb, (b.bk_b >= 0 && kb == 16) ? 0 : passes, lab);
  CHK(get("zero", b.z.total, &p));
  b.zb = static_cast<char *>(p);
  if (!quantized && !custom) {
    CHK(get("bbox_part", kBboxPartBytes, &p));
    b.bbox_part = static_cast<float *>(p);
  }
  if (b.bk_b >= 0) {
    CHK(get("bucket_ws", bucket_ws_bytes<K>(n), &p));
    bucket_carve<K>(p, n, b.bk);
  }
  return PGVD_OK;
}

// Enqueues the whole path on s over the carved buffers b, with its run
// state st and the octree output (capacity octree_cap nodes; node_conflict
// the same count, with labels).  ctx (nullable) only provides timing.
template <typename K>
int build_launch(pgvd_ctx *ctx, const BuildBufs<K> &b, RunState *st, int32_t *octree,
                 uint32_t *nconf, int64_t octree_cap, const void *pts, int64_t n, int dim,
                 int bits, const uint32_t *labels, const pgvd_build_opts *opts, hipStream_t s,
                 BuildOut<K> &out, bool defer_exact) {
  const int mbits = dim * bits;
  const int passes = (mbits + 7) / 8;
  const int prepend = ref_next_pow2((int)n) != n ? 1 : 0;   // zero-key padding (Q1)
  const int64_t nk = n + 1;
  const bool lab = labels != nullptr;
  K *kA = b.kA, *kB = b.kB;
  const ZeroLayout &z = b.z;
  char *zb = b.zb;
  uint32_t *hist = reinterpret_cast<uint32_t *>(zb + z.hist);
  uint32_t *sort_lb = reinterpret_cast<uint32_t *>(zb + z.sort_lb);
  uint64_t *lb_u = reinterpret_cast<uint64_t *>(zb + z.lb_unique);
  uint64_t *lb_s = reinterpret_cast<uint64_t *>(zb + z.lb_splits);
  uint64_t *lb_b = reinterpret_cast<uint64_t *>(zb + z.lb_brt);
  uint32_t *cbits = lab ? reinterpret_cast<uint32_t *>(zb + z.cbits) : nullptr;
  uint32_t *msd_hist = reinterpret_cast<uint32_t *>(zb + z.msd_hist);
  const int bk_b = b.bk_b;

  const bool quantized = opts && opts->quantized_input;
  const bool custom = opts && opts->use_custom_bb;
  const bool clamped = opts && opts->clamped;
  const bool bucket = bk_b >= 0;
  // the build's start (state reset, workspace zeroed) rides on the bbox pass
  // when there is one
  const bool bbox_pass = !quantized && !custom;
  if (!bbox_pass) {
    Stage t(ctx, "prep", s);
    launch_prep(st, reinterpret_cast<uint32_t *>(zb), z.total / 4, s);
  }
  CHK(launch_status("prep"));
  if (quantized) {
    Stage t(ctx, "encode", s);
    launch_encode_int<K>(static_cast<const int32_t *>(pts), n, dim, bits, kA, hist, passes, s);
  } else {
    if (bbox_pass) {
      Stage t(ctx, "bbox", s);
      launch_bbox(static_cast<const float *>(pts), n, dim, b.bbox_part, st, s,
                  reinterpret_cast<uint32_t *>(zb), (bucket ? z.sort_lb : z.total) / 4,
                  true);
    }
    const int64_t width = (int64_t)1 << bits;
    const float effw = (float)(clamped ? width : width - 1);
    Stage t(ctx, "encode", s);
    if (bucket)
      launch_encode<K>(static_cast<const float *>(pts), n, dim, bits, effw, custom ? 1 : 0,
                       custom ? opts->bb_min : nullptr, custom ? opts->bb_max : nullptr, kA,
                       nullptr, hist, 0, st, s, msd_hist, mbits - bk_b, 1 << bk_b);
    else
      launch_encode<K>(static_cast<const float *>(pts), n, dim, bits, effw, custom ? 1 : 0,
                       custom ? opts->bb_min : nullptr, custom ? opts->bb_max : nullptr, kA,
                       nullptr, hist, passes, st, s);
  }
  CHK(launch_status("encode"));
  int parity = 0;
  bool bucket_sorted = false;
  if (bucket) {
    // the bucket sort leaves keys and labels in (kA, vA)
    Stage t(ctx, "msd", s);
    BucketArgs<K> ba{kA, kB, b.vA, b.vB, labels, msd_hist, n, mbits, bk_b, prepend, b.bk};
    const int rc = launch_bucket_sort<K>(ba, st, ctx->st_host, ctx->sync_ev, s);
    if (rc == 2) return fail(PGVD_ERR_INTERNAL, "bucket sort: a segment outlived its levels");
    if (rc == 1) return launch_status("bucket sort") != PGVD_OK ? PGVD_ERR_HIP
                                                                  : fail(PGVD_ERR_HIP, "bucket sort");
    bucket_sorted = rc == 0;
    // rc 3: clustered keys of at most 64 bits; the encoded keys are untouched
    if (!bucket_sorted)
      launch_histogram<K>(kA, n, hist, passes, s, nullptr, sort_lb, (z.total - z.sort_lb) / 4);
  }
  if (!bucket_sorted) {
    Stage span(ctx, "sort", s, 2);
    PassTimer pt(ctx, n > 1 ? passes : 0);
    parity = launch_onesweep<K>(kA, kB, labels, b.vA, b.vB, n, passes, hist, sort_lb, st, s,
                                pt.events(), nullptr);
  }
  CHK(launch_status("sort"));
  K *sorted = parity ? kB : kA;
  K *uniq = sorted - prepend;   // in place (see unique_tile)
  // packed local splits go to the sort's other buffer, free after the sort
  uint32_t *lsx = reinterpret_cast<uint32_t *>(parity ? kA : kB);
  const uint32_t *sortedv =
      lab ? (bucket_sorted ? b.vA : (n <= 1 ? labels : (parity ? b.vB : b.vA))) : nullptr;
  {
    Stage t(ctx, "unique", s);
    launch_unique<K>(sorted, sortedv, n, prepend, uniq, cbits, lb_u, st, s,
                     bucket_sorted ? st : nullptr);
  }
  CHK(launch_status("unique"));
  {
    Stage t(ctx, "brt", s);
    // local-split totals per chunk of 256 nodes go to chunks
    launch_brt_ls<K>(uniq, nk, mbits, dim, b.brt, b.parent, lsx, b.chunks, b.defer, st, s);
  }
  CHK(launch_status("brt"));
  {
    // chunk bases and the octree size; the gather writes ps from them
    Stage t(ctx, "splits_scan", s);
    launch_chunk_scan(b.chunks, nk, lb_b, st, s);
  }
  CHK(launch_status("splits_scan"));
  if (!defer_exact) {
    // order-exact claims, local splits and scan; no-ops unless some BRT
    // split took the FP32-quirk value (with a host to wait for the build,
    // they run at its finish, and only when a split did)
    Stage t(ctx, "malformed", s);
    launch_claims_if_malformed<K>(b.brt, b.parent, nk, st, s);
    launch_splits_scan<K>(b.brt, b.parent, nullptr, b.ps, nk, dim, 1, lb_s, st, s);
  }
  CHK(launch_status("malformed path"));
  {
    Stage t(ctx, "octree", s);
    hipEvent_t gev[2] = {nullptr, nullptr};
    if (ctx && ctx->timing == 3) {
      gev[0] = take_event(ctx);
      gev[1] = take_event(ctx);
    }
    const bool timed = gev[0] && gev[1];
    launch_octree<K>(b.brt, b.parent, b.ps, lsx, b.chunks, uniq, cbits, nk, dim, mbits, octree,
                     nconf, octree_cap, defer_exact ? 2 : 0, b.defer, st, s,
                     timed ? gev : nullptr);
    if (timed) ctx->recs.push_back({"gather", gev[0], gev[1]});
  }
  CHK(launch_status("octree"));
  out.uniq = uniq;
  out.lsx = lsx;
  out.cbits = cbits;
  return PGVD_OK;
}

template <typename K>
int build_impl(pgvd_ctx *ctx, const void *pts, int64_t n, int dim, int bits,
               const uint32_t *labels, const pgvd_build_opts *opts, hipStream_t s) {
  const bool lab = labels != nullptr;
  BuildBufs<K> b;
  CHK(carve_build<K>([ctx](const char *key, size_t bytes, void **p) {
        return pool_get(ctx, key, bytes, p);
      }, pts, n, dim, bits, lab, opts, !(opts && opts->no_host_wait), b));
  int32_t *octree;
  uint32_t *nconf;
  int64_t want = 2 * (n + 1);
  if (want < 1024) want = 1024;
  if (ctx->octree_cap < want || ctx->octree_stride != (1 << dim) + 1 ||
      (lab && ctx->pool["node_conflict"].bytes < (size_t)ctx->octree_cap * 4))
    CHK(ensure_octree<K>(ctx, dim, want, lab, &octree, &nconf));
  octree = static_cast<int32_t *>(ctx->pool["octree"].p);
  nconf = lab ? static_cast<uint32_t *>(ctx->pool["node_conflict"].p) : nullptr;
  BuildOut<K> o;
  CHK(build_launch<K>(ctx, b, ctx->st, octree, nconf, ctx->octree_cap, pts, n, dim, bits, labels,
                      opts, s, o, true));
  HIPCHK(hipMemcpyAsync(ctx->st_host, ctx->st, sizeof(RunState), hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(ctx->done, s));
  LastBuild &L = ctx->last;
  L.valid = true;
  L.n = n;
  L.dim = dim;
  L.bits = bits;
  L.mbits = dim * bits;
  L.kb = (int)sizeof(K);
  L.labels = lab;
  L.keys_u = o.uniq;
  L.brt = b.brt;
  L.parent = b.parent;
  L.ps = b.ps;
  L.lsx = o.lsx;
  L.chunks = b.chunks;
  L.defer = b.defer;
  L.conflict_bits = o.cbits;
  L.lb_splits = reinterpret_cast<uint64_t *>(b.zb + b.z.lb_splits);
  L.exact_deferred = true;
  L.stream = s;
  return PGVD_OK;
}

// The order-exact path of a build whose tree turned out malformed (some BRT
// split took the FP32-quirk value): claims, order-exact local splits and
// scan, and the scatter fill, as build_launch would have run them.
template <typename K>
int exact_path(pgvd_ctx *ctx) {
  LastBuild &L = ctx->last;
  const int64_t nk = L.n + 1;
  int32_t *octree = static_cast<int32_t *>(ctx->pool["octree"].p);
  uint32_t *nconf = L.labels ? static_cast<uint32_t *>(ctx->pool["node_conflict"].p) : nullptr;
  const BrtNodeT<K> *brt = static_cast<const BrtNodeT<K> *>(L.brt);
  {
    Stage t(ctx, "malformed", L.stream);
    launch_claims_if_malformed<K>(brt, L.parent, nk, ctx->st, L.stream);
    launch_splits_scan<K>(brt, L.parent, nullptr, L.ps, nk, L.dim, 1, L.lb_splits, ctx->st,
                          L.stream);
    launch_octree<K>(brt, L.parent, L.ps, L.lsx, L.chunks, static_cast<const K *>(L.keys_u),
                     L.conflict_bits, nk, L.dim, L.mbits, octree, nconf, ctx->octree_cap, 0,
                     L.defer, ctx->st, L.stream);
  }
  CHK(launch_status("order-exact path"));
  return copy_state_and_wait(ctx, L.stream);
}

template <typename K>
int refill_octree(pgvd_ctx *ctx, int64_t nodes) {
  LastBuild &L = ctx->last;
  int32_t *octree;
  uint32_t *nconf;
  CHK(ensure_octree<K>(ctx, L.dim, nodes, L.labels, &octree, &nconf));
  // clear the capacity bit and refill from the kept BRT / prefix sums
  ctx->st_host->status &= ~(uint32_t)ST_CAPACITY;
  HIPCHK(hipMemcpyAsync(&ctx->st->status, &ctx->st_host->status, 4, hipMemcpyHostToDevice,
                        L.stream));
  launch_octree<K>(static_cast<const BrtNodeT<K> *>(L.brt), L.parent, L.ps, L.lsx, L.chunks,
                   static_cast<const K *>(L.keys_u), L.conflict_bits, L.n + 1, L.dim, L.mbits,
                   octree, nconf, ctx->octree_cap, 0, L.defer, ctx->st, L.stream);
  CHK(launch_status("octree refill"));
  return copy_state_and_wait(ctx, L.stream);
}

// ---------------------------------------------------------------------------
// stage helpers
// ---------------------------------------------------------------------------
template <typename K>
int radix_sort_impl(pgvd_ctx *ctx, K *keys, uint32_t *vals, int64_t n, int mbits, hipStream_t s) {
  const int passes = (mbits + 7) / 8;
  void *p;
  CHK(pool_get(ctx, "sort_alt_k", (size_t)n * sizeof(K), &p));
  K *alt = static_cast<K *>(p);
  uint32_t *altv = nullptr;
  if (vals) {
    CHK(pool_get(ctx, "sort_alt_v", (size_t)n * 4, &p));
    altv = static_cast<uint32_t *>(p);
  }
  const ZeroLayout z = zero_layout(n, (int)sizeof(K), passes, vals != nullptr);
  CHK(pool_get(ctx, "zero_stage", z.total, &p));
  char *zb = static_cast<char *>(p);
  uint32_t *hist = reinterpret_cast<uint32_t *>(zb + z.hist);
  launch_prep(ctx->st, reinterpret_cast<uint32_t *>(zb), z.total / 4, s);
  launch_histogram<K>(keys, n, hist, passes, s);
  int parity;
  {
    PassTimer pt(ctx, passes);
    parity = launch_onesweep<K>(keys, alt, vals, vals, altv, n, passes, hist,
                                reinterpret_cast<uint32_t *>(zb + z.sort_lb), ctx->st, s,
                                pt.events());
  }
  CHK(launch_status("onesweep"));
  if (parity) {
    HIPCHK(hipMemcpyAsync(keys, alt, (size_t)n * sizeof(K), hipMemcpyDeviceToDevice, s));
    if (vals) HIPCHK(hipMemcpyAsync(vals, altv, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
  }
  return PGVD_OK;
}

template <typename K>
int unique_impl(pgvd_ctx *ctx, const K *in, K *out, int64_t n, int64_t *m_out, hipStream_t s) {
  const ZeroLayout z = zero_layout(n, (int)sizeof(K), 0, false);
  void *p;
  CHK(pool_get(ctx, "zero_stage", z.total, &p));
  char *zb = static_cast<char *>(p);
  launch_prep(ctx->st, reinterpret_cast<uint32_t *>(zb), z.total / 4, s);
  {
    Stage t(ctx, "unique", s);
    launch_unique<K>(in, nullptr, n, 0, out, nullptr, reinterpret_cast<uint64_t *>(zb + z.lb_unique),
                     ctx->st, s);
  }
  CHK(launch_status("unique"));
  CHK(copy_state_and_wait(ctx, s));
  *m_out = ctx->st_host->m;
  return PGVD_OK;
}

template <typename K>
int brt_impl(pgvd_ctx *ctx, const K *keys, int64_t m, int mbits, BrtNodeT<K> *brt,
             int32_t *parent, hipStream_t s) {
  launch_prep(ctx->st, nullptr, 0, s);
  if (m > 1) {
    launch_fill_u32(reinterpret_cast<uint32_t *>(parent), (size_t)(m - 1), 0u, s);
    Stage t(ctx, "brt", s);
    launch_brt<K>(keys, m, m, mbits, brt, parent, ctx->st, s);
    launch_finalize_parents(parent, m, m, ctx->st, s);
  }
  CHK(launch_status("brt"));
  CHK(copy_state_and_wait(ctx, s));
  // equal neighbouring keys give the tree the reference builds from them
  return status_error(ctx->st_host->status & ~(uint32_t)ST_DUP_KEY);
}

template <typename K>
int local_splits_impl(pgvd_ctx *ctx, const BrtNodeT<K> *brt, int64_t m, int dim, uint32_t *ls,
                      hipStream_t s) {
  void *p;
  CHK(pool_get(ctx, "ls_tmp", (size_t)m * 8, &p));
  launch_fill_u32(static_cast<uint32_t *>(p), (size_t)m * 2, 0u, s);
  launch_local_splits<K>(brt, m, dim, ls, static_cast<uint64_t *>(p), s);
  return launch_status("local splits");
}

template <typename K>
int brt_to_octree_impl(pgvd_ctx *ctx, const BrtNodeT<K> *brt, const int32_t *parent, int64_t m,
                       int dim, int32_t *octree, int64_t capacity, int64_t *octree_size,
                       hipStream_t s) {
  void *p;
  CHK(pool_get(ctx, "stage_ls", (size_t)m * 4, &p));
  uint32_t *ls = static_cast<uint32_t *>(p);
  CHK(pool_get(ctx, "stage_ps", (size_t)m * 4, &p));
  uint32_t *ps = static_cast<uint32_t *>(p);
  const ZeroLayout z = zero_layout(m, (int)sizeof(K), 0, false);
  CHK(pool_get(ctx, "zero_stage", z.total, &p));
  char *zb = static_cast<char *>(p);
  launch_prep(ctx->st, reinterpret_cast<uint32_t *>(zb), z.total / 4, s);
  CHK(local_splits_impl<K>(ctx, brt, m, dim, ls, s));
  launch_scan_u32(ls, ps, m, reinterpret_cast<uint64_t *>(zb + z.lb_scan), ctx->st, s);
  launch_set_sizes(ctx->st, ps, m, s);
  CHK(launch_status("splits scan"));
  CHK(copy_state_and_wait(ctx, s));
  const int64_t O = ctx->st_host->octree_size;
  *octree_size = O;
  if (O > capacity) return fail(PGVD_ERR_CAPACITY, "octree buffer too small");
  {
    Stage t(ctx, "octree", s);
    // mbits is not needed without conflict flags; keys are not read.
    launch_octree<K>(brt, parent, ps, nullptr, nullptr, nullptr, nullptr, m, dim, 0, octree,
                     nullptr, capacity, 1, nullptr, ctx->st, s);
  }
  CHK(launch_status("octree"));
  CHK(copy_state_and_wait(ctx, s));
  return status_error(ctx->st_host->status);
}

// ---------------------------------------------------------------------------
// allocation-free whole path
// ---------------------------------------------------------------------------
// Workspace of pgvd_build_octree_noalloc: every intermediate of a build (run
// state, keys, BRT, prefix sums, look-back words) in one static device
// array.  Each GPU the library runs on holds its own copy until the process
// ends.  In host code `scratch` only names the symbol; its device address
// on the current GPU comes from hipGetSymbolAddress (noalloc_workspace).
constexpr size_t kNoAllocScratchBytes = PGVD_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) static unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The workspace's address on the current device, looked up on the first
// call there and kept.
int noalloc_workspace(void **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(PGVD_ERR_NO_DEVICE, "device index above 63");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(PGVD_ERR_HIP, "hipGetSymbolAddress returned no address for the workspace");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return PGVD_OK;
}

template <typename K>
int noalloc_impl(char *base, const void *pts, int64_t n, int dim, int bits,
                 const uint32_t *labels, const pgvd_build_opts *opts, int32_t *octree,
                 int64_t octree_cap, uint32_t *node_conflict, int32_t *state, hipStream_t s) {
  // bump carving in a fixed order; the total is checked before any launch
  size_t off = 0;
  auto get = [&](const char *, size_t bytes, void **p) {
    if (bytes == 0) bytes = 256;
    *p = base + off;
    off += align256(bytes);
    return (int)PGVD_OK;
  };
  void *p;
  CHK(get("state", sizeof(RunState), &p));
  RunState *st = static_cast<RunState *>(p);
  BuildBufs<K> b;
  // (no host waits here: the LSD passes sort)
  CHK(carve_build<K>(get, pts, n, dim, bits, labels != nullptr, opts, false, b));
  if (off > kNoAllocScratchBytes)
    return fail(PGVD_ERR_TOO_LARGE,
                "pgvd_build_octree_noalloc: this build needs " + std::to_string(off) +
                    " bytes of workspace; the allocation-free entry holds " +
                    std::to_string(kNoAllocScratchBytes));
  BuildOut<K> o;
  CHK(build_launch<K>(nullptr, b, st, octree, node_conflict, octree_cap, pts, n, dim, bits,
                      labels, opts, s, o, false));
  launch_build_state(st, state, s);
  return launch_status("build state");
}

}  // namespace

// ===========================================================================
// exported entry points
// ===========================================================================
extern "C" {

const char *pgvd_last_error(void) { return g_err.c_str(); }
const char *pgvd_version(void) { return "pgvd-mi355x 0.1 (gfx950)"; }

int pgvd_next_pow2(int n) { return ref_next_pow2(n); }

int pgvd_ctx_create(int device, pgvd_ctx **out) {
  if (!out) return fail(PGVD_ERR_INVALID, "null out pointer");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(PGVD_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(PGVD_ERR_NO_DEVICE, "device index out of range");
  pgvd_ctx *ctx = new pgvd_ctx();
  ctx->device = device;
  int rc = PGVD_OK;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess ||
      hipMalloc(&ctx->st, sizeof(RunState)) != hipSuccess ||
      hipHostMalloc(&ctx->st_host, sizeof(RunState), hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->done, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->sync_ev, hipEventDisableTiming) != hipSuccess) {
    rc = fail(PGVD_ERR_HIP, "context setup failed");
  }
  if (rc != PGVD_OK) {
    pgvd_ctx_destroy(ctx);
    return rc;
  }
  std::memset(ctx->st_host, 0, sizeof(RunState));
  *out = ctx;
  return PGVD_OK;
}

int pgvd_ctx_destroy(pgvd_ctx *ctx) {
  if (!ctx) return PGVD_OK;
  // Teardown is best effort: errors here cannot be acted on.
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &kv : ctx->pool)
    if (kv.second.p) (void)hipFree(kv.second.p);
  for (auto &r : ctx->recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : ctx->spare) (void)hipEventDestroy(e);
  if (ctx->done) (void)hipEventDestroy(ctx->done);
  if (ctx->sync_ev) (void)hipEventDestroy(ctx->sync_ev);
  if (ctx->st) (void)hipFree(ctx->st);
  if (ctx->st_host) (void)hipHostFree(ctx->st_host);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return PGVD_OK;
}

int pgvd_ctx_stream(pgvd_ctx *ctx, pgvd_stream *out) {
  if (!ctx || !out) return fail(PGVD_ERR_INVALID, "null argument");
  *out = ctx->stream;
  return PGVD_OK;
}

int pgvd_ctx_set_pool_limit(pgvd_ctx *ctx, size_t bytes) {
  if (!ctx) return fail(PGVD_ERR_INVALID, "null context");
  ctx->pool_limit = bytes;
  return PGVD_OK;
}

int pgvd_ctx_pool_bytes(pgvd_ctx *ctx, size_t *out) {
  if (!ctx || !out) return fail(PGVD_ERR_INVALID, "null argument");
  *out = ctx->pool_bytes;
  return PGVD_OK;
}

int pgvd_debug_fail_launch(const char *stage) {
  g_fail_launch = stage ? stage : "";
  return PGVD_OK;
}

int pgvd_debug_gather_queue(int entries) {
  if (entries < 0) return fail(PGVD_ERR_INVALID, "entries must be >= 0");
  if (pgvd::set_gather_queue_cap((uint32_t)entries) != 0)
    return fail(PGVD_ERR_HIP, "hipMemcpyToSymbol failed for the gather queue cap");
  return PGVD_OK;
}

int pgvd_synchronize(pgvd_ctx *ctx, pgvd_stream s) {
  if (!ctx) return fail(PGVD_ERR_INVALID, "null context");
  CHK(use_device(ctx));
  HIPCHK(hipStreamSynchronize(pick(ctx, s)));
  return PGVD_OK;
}

int pgvd_alloc(pgvd_ctx *ctx, size_t bytes, void **dptr) {
  if (!ctx || !dptr) return fail(PGVD_ERR_INVALID, "null argument");
  CHK(use_device(ctx));
  if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();   // not left for the next launch check to report
    return fail(PGVD_ERR_ALLOC, "hipMalloc failed");
  }
  return PGVD_OK;
}

int pgvd_free(pgvd_ctx *ctx, void *dptr) {
  if (!ctx) return fail(PGVD_ERR_INVALID, "null context");
  CHK(use_device(ctx));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipFree(dptr));
  return PGVD_OK;
}

int pgvd_memcpy_htod(pgvd_ctx *ctx, void *dst, const void *src, size_t bytes, pgvd_stream s) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(PGVD_ERR_INVALID, "null argument");
  CHK(use_device(ctx));
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, pick(ctx, s)));
  HIPCHK(hipStreamSynchronize(pick(ctx, s)));
  return PGVD_OK;
}

int pgvd_memcpy_dtoh(pgvd_ctx *ctx, void *dst, const void *src, size_t bytes, pgvd_stream s) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(PGVD_ERR_INVALID, "null argument");
  CHK(use_device(ctx));
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, pick(ctx, s)));
  HIPCHK(hipStreamSynchronize(pick(ctx, s)));
  return PGVD_OK;
}

int pgvd_copy_kernel(pgvd_ctx *ctx, void *dst, const void *src, size_t bytes, pgvd_stream s) {
  if (!ctx || !dst || !src) return fail(PGVD_ERR_INVALID, "null argument");
  if (((uintptr_t)dst | (uintptr_t)src) & 15) return fail(PGVD_ERR_INVALID, "copy needs 16-byte aligned pointers");
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  {
    Stage t(ctx, "copy", st);
    launch_copy(dst, src, bytes, st);
  }
  return launch_status("copy");
}

int pgvd_digest(pgvd_ctx *ctx, const void *p, size_t bytes, pgvd_stream s, uint64_t *out) {
  if (!ctx || !out || (!p && bytes)) return fail(PGVD_ERR_INVALID, "null argument");
  if ((bytes & 3) || ((uintptr_t)p & 3)) return fail(PGVD_ERR_INVALID, "digest needs 4-byte aligned whole words");
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  void *d = nullptr;
  CHK(pool_get(ctx, "digest", 8, &d));
  launch_digest(p, bytes / 4, static_cast<uint64_t *>(d), true, st);
  CHK(launch_status("digest"));
  HIPCHK(hipMemcpyAsync(out, d, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return PGVD_OK;
}

int pgvd_timing_enable(pgvd_ctx *ctx, int on) {
  if (!ctx) return fail(PGVD_ERR_INVALID, "null context");
  ctx->timing = (on == 2 || on == 3) ? on : (on != 0 ? 1 : 0);
  return PGVD_OK;
}

int pgvd_timing_reset(pgvd_ctx *ctx) {
  if (!ctx) return fail(PGVD_ERR_INVALID, "null context");
  CHK(use_device(ctx));
  HIPCHK(hipDeviceSynchronize());
  for (auto &r : ctx->recs) {
    ctx->spare.push_back(r.a);
    ctx->spare.push_back(r.b);
  }
  ctx->recs.clear();
  return PGVD_OK;
}

int pgvd_timing_read(pgvd_ctx *ctx, const char *stage, double *total_ms, int64_t *count) {
  if (!ctx || !stage || !total_ms || !count) return fail(PGVD_ERR_INVALID, "null argument");
  CHK(use_device(ctx));
  HIPCHK(hipDeviceSynchronize());
  double tot = 0;
  int64_t c = 0;
  for (auto &r : ctx->recs) {
    if (r.stage != stage) continue;
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
    tot += ms;
    ++c;
  }
  *total_ms = tot;
  *count = c;
  return PGVD_OK;
}

int pgvd_bbox(pgvd_ctx *ctx, const float *pts, int64_t n, int dim, float *bb_out, pgvd_stream s) {
  if (!ctx || !pts || !bb_out) return fail(PGVD_ERR_INVALID, "null argument");
  if (dim != 2 && dim != 3) return fail(PGVD_ERR_INVALID, "dim must be 2 or 3");
  CHK(check_n(n));
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  launch_prep(ctx->st, nullptr, 0, st);
  void *part;
  CHK(pool_get(ctx, "bbox_part", kBboxPartBytes, &part));
  {
    Stage t(ctx, "bbox", st);
    launch_bbox(pts, n, dim, static_cast<float *>(part), ctx->st, st);
  }
  CHK(launch_status("bbox"));
  CHK(copy_state_and_wait(ctx, st));
  for (int k = 0; k < dim; ++k) {
    bb_out[k] = fdec_host(ctx->st_host->bb_min[k]);
    bb_out[dim + k] = fdec_host(ctx->st_host->bb_max[k]);
  }
  return PGVD_OK;
}

int pgvd_quantize(pgvd_ctx *ctx, const float *pts, int64_t n, int dim, int bits,
                  const float *custom_bb, int clamped, int64_t *qpts_out, int64_t *n_out,
                  pgvd_stream s) {
  if (!ctx || !pts || !qpts_out || !n_out) return fail(PGVD_ERR_INVALID, "null argument");
  CHK(check_dims(dim, bits));
  if (bits > 62) return fail(PGVD_ERR_INVALID, "bits must be <= 62 for int64 coordinates");
  CHK(check_n(n));
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  launch_prep(ctx->st, nullptr, 0, st);
  if (!custom_bb) {
    void *part;
    CHK(pool_get(ctx, "bbox_part", kBboxPartBytes, &part));
    launch_bbox(pts, n, dim, static_cast<float *>(part), ctx->st, st);
  }
  const int64_t width = (int64_t)1 << bits;
  const float effw = (float)(clamped ? width : width - 1);
  {
    Stage t(ctx, "encode", st);
    launch_encode<uint64_t>(pts, n, dim, bits, effw, custom_bb ? 1 : 0, custom_bb,
                            custom_bb ? custom_bb + dim : nullptr, nullptr, qpts_out, nullptr, 0,
                            ctx->st, st);
  }
  CHK(launch_status("quantize"));
  CHK(copy_state_and_wait(ctx, st));
  CHK(status_error(ctx->st_host->status));
  float mn[3], mx[3];
  for (int k = 0; k < dim; ++k) {
    mn[k] = custom_bb ? custom_bb[k] : fdec_host(ctx->st_host->bb_min[k]);
    mx[k] = custom_bb ? custom_bb[dim + k] : fdec_host(ctx->st_host->bb_max[k]);
  }
  float dw = mx[0] - mn[0];
  for (int k = 1; k < dim; ++k) {
    const float sz = mx[k] - mn[k];
    dw = (sz > dw) ? sz : dw;
  }
  *n_out = (dw == 0.0f) ? 1 : n;
  return PGVD_OK;
}

int pgvd_points_to_morton(pgvd_ctx *ctx, const void *pts, int coord_bytes, int64_t n, int dim,
                          int bits, void *keys_out, int key_bytes, pgvd_stream s) {
  if (!ctx || !pts || !keys_out) return fail(PGVD_ERR_INVALID, "null argument");
  if (coord_bytes != 4 && coord_bytes != 8) return fail(PGVD_ERR_INVALID, "coord_bytes must be 4 or 8");
  CHK(check_dims(dim, bits));
  CHK(check_n(n));
  if (key_bytes < key_bytes_for(dim * bits)) return fail(PGVD_ERR_INVALID, "key_bytes too small for bits*dim");
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  const int64_t padded = ref_next_pow2((int)n);
  Stage t(ctx, "encode", st);
  if (key_bytes == 4) launch_points_to_morton<uint32_t>(pts, coord_bytes, n, padded, dim, bits, static_cast<uint32_t *>(keys_out), st);
  else if (key_bytes == 8) launch_points_to_morton<uint64_t>(pts, coord_bytes, n, padded, dim, bits, static_cast<uint64_t *>(keys_out), st);
  else if (key_bytes == 16) launch_points_to_morton<Key128>(pts, coord_bytes, n, padded, dim, bits, static_cast<Key128 *>(keys_out), st);
  else return fail(PGVD_ERR_INVALID, "key_bytes must be 4, 8 or 16");
  return launch_status("points_to_morton");
}

int pgvd_find_leaf(pgvd_ctx *ctx, const int32_t *pts, int64_t q, int dim, int bits,
                   const int32_t *octree, int64_t octree_size, int32_t *parent, int32_t *octant,
                   int32_t *data, int32_t *level, uint32_t *status, pgvd_stream s) {
  if (!ctx || !pts || !octree || !parent || !octant || !data || !level)
    return fail(PGVD_ERR_INVALID, "null argument");
  if (dim != 2 && dim != 3) return fail(PGVD_ERR_INVALID, "dim must be 2 or 3");
  if (bits < 1 || bits > 31) return fail(PGVD_ERR_INVALID, "bits must be in 1..31 (int32 grid points)");
  if (q < 0) return fail(PGVD_ERR_INVALID, "q must be >= 0");
  if (octree_size < 1) return fail(PGVD_ERR_INVALID, "octree_size must be >= 1");
  if (q == 0) return PGVD_OK;
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  const int D = find_leaf_table_depth(q, dim, bits);
  void *table = nullptr;
  if (D > 0) CHK(pool_get(ctx, "find_leaf_table", size_t(8) << (D * dim), &table));
  Stage t(ctx, "find_leaf", st);
  launch_find_leaf(pts, q, dim, bits, octree, octree_size, table, D, parent, octant, data, level,
                   status, st);
  return launch_status("find_leaf");
}

int pgvd_cell_walk(pgvd_ctx *ctx, const float *segs, int64_t n, int bits, const int32_t *octree,
                   int64_t octree_size, int64_t *offsets, int32_t *cells, int64_t capacity,
                   int64_t *total_out, int32_t *status, pgvd_stream s) {
  if (!ctx || !octree || !offsets || !total_out || (n > 0 && !segs))
    return fail(PGVD_ERR_INVALID, "null argument");
  if (bits < 1 || bits > 30) return fail(PGVD_ERR_INVALID, "bits must be in 1..30 (2D)");
  if (n < 0 || n >= (int64_t(1) << 31)) return fail(PGVD_ERR_INVALID, "n must be in [0, 2^31)");
  if (octree_size < 1) return fail(PGVD_ERR_INVALID, "octree_size must be >= 1");
  if (capacity < 0) return fail(PGVD_ERR_INVALID, "capacity must be >= 0");
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  if (n == 0) {
    const int64_t zero = 0;
    // on the caller's stream, ordered with its work on the buffer
    HIPCHK(hipMemcpyAsync(offsets, &zero, 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    *total_out = 0;
    return PGVD_OK;
  }
  void *counts = nullptr, *ws = nullptr, *scratch = nullptr;
  CHK(pool_get(ctx, "walk_counts", 4 * static_cast<size_t>(n), &counts));
  CHK(pool_get(ctx, "walk_ws", 8 * static_cast<size_t>(walk_scan_ws_words(n)), &ws));
  // With a cell buffer, the first walk keeps each segment's first 32 cells
  // (up to 2 GB of scratch, or PGVD_WALK_SCRATCH_BYTES from the
  // environment), so that only longer walks run twice; past that, a count
  // walk and a write walk.
  constexpr int kLog2K = 5;
  const size_t spec_bytes = (static_cast<size_t>(n) << kLog2K) * 20;
  size_t spec_limit = size_t(2) << 30;
  if (const char *e = std::getenv("PGVD_WALK_SCRATCH_BYTES")) spec_limit = std::strtoull(e, nullptr, 10);
  if (cells && spec_bytes <= spec_limit)
    CHK(pool_get(ctx, "walk_scratch", spec_bytes, &scratch));
  int64_t total = 0;
  {
    Stage t(ctx, "cell_walk_count", st);
    launch_cell_walk_count(segs, n, bits, octree, octree_size, static_cast<uint32_t *>(counts),
                           status, offsets, static_cast<uint64_t *>(ws),
                           static_cast<int32_t *>(scratch), kLog2K, st);
    CHK(launch_status("cell_walk_count"));
  }
  HIPCHK(hipMemcpyAsync(&total, offsets + n, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *total_out = total;
  if (!cells) return PGVD_OK;
  if (total > capacity) return fail(PGVD_ERR_CAPACITY, "cell buffer too small (*total_out cells needed)");
  Stage t(ctx, "cell_walk", st);
  launch_cell_walk_write(segs, n, bits, octree, octree_size, offsets, cells,
                         static_cast<const int32_t *>(scratch), static_cast<uint32_t *>(counts),
                         kLog2K, st);
  return launch_status("cell_walk");
}

int pgvd_octree_format_text(pgvd_ctx *ctx, const int32_t *octree, int64_t O, int dim, char *text,
                            int64_t capacity, int64_t *len_out, pgvd_stream s) {
  if (!ctx || !len_out || (O > 0 && !octree)) return fail(PGVD_ERR_INVALID, "null argument");
  if (dim != 2 && dim != 3) return fail(PGVD_ERR_INVALID, "dim must be 2 or 3");
  if (O < 0 || O > INT32_MAX) return fail(PGVD_ERR_INVALID, "O must be in [0, 2^31)");
  if (capacity < 0) return fail(PGVD_ERR_INVALID, "capacity must be >= 0");
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  char hdr[24];
  const int hl = std::snprintf(hdr, sizeof hdr, "%lld ", static_cast<long long>(O));
  if (O == 0) {
    *len_out = hl;
    if (!text) return PGVD_OK;
    if (capacity < hl) return fail(PGVD_ERR_CAPACITY, "text buffer too small (*len_out bytes needed)");
    HIPCHK(hipMemcpyAsync(text, hdr, hl, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return PGVD_OK;
  }
  void *ws = nullptr;
  CHK(pool_get(ctx, "text_ws", 8 * text_ws_words(format_text_tiles(O)), &ws));
  uint64_t total = 0;
  {
    Stage t(ctx, "format_text", st);
    launch_format_text(octree, O, dim, hl, reinterpret_cast<uint8_t *>(text), capacity,
                       static_cast<uint64_t *>(ws), st);
    CHK(launch_status("format_text"));
  }
  HIPCHK(hipMemcpyAsync(&total, static_cast<uint64_t *>(ws) + 1, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *len_out = static_cast<int64_t>(total);
  if (text && static_cast<int64_t>(total) > capacity)
    return fail(PGVD_ERR_CAPACITY, "text buffer too small (*len_out bytes needed)");
  return PGVD_OK;
}

namespace {
inline bool text_space(char c) {
  return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f';
}
}  // namespace

int pgvd_octree_parse_text(pgvd_ctx *ctx, const char *text, int64_t len, int dim, int32_t *octree,
                           int64_t cap, int64_t *O_out, pgvd_stream s) {
  if (!ctx || !text || !O_out || (cap > 0 && !octree)) return fail(PGVD_ERR_INVALID, "null argument");
  if (dim != 2 && dim != 3) return fail(PGVD_ERR_INVALID, "dim must be 2 or 3");
  if (len < 1 || len >= (int64_t(1) << 42)) return fail(PGVD_ERR_INVALID, "len must be in [1, 2^42)");
  if (cap < 0) return fail(PGVD_ERR_INVALID, "cap must be >= 0");
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  // the node count, the first token, is read here
  std::vector<char> head(static_cast<size_t>(std::min<int64_t>(len, 4096)));
  HIPCHK(hipMemcpyAsync(head.data(), text, head.size(), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  size_t i = 0;
  while (i < head.size() && text_space(head[i])) ++i;
  bool neg = false;
  if (i < head.size() && (head[i] == '+' || head[i] == '-')) neg = head[i++] == '-';
  int64_t N = 0;
  int nd = 0;
  bool digit = false;
  while (i < head.size() && !text_space(head[i])) {
    if (head[i] < '0' || head[i] > '9') return fail(PGVD_ERR_INVALID, "the node count is not a number");
    digit = true;
    if (nd || head[i] != '0') {   // leading zeros skipped
      if (nd >= 10) return fail(PGVD_ERR_INVALID, "the node count is not a number");
      N = N * 10 + (head[i] - '0');
      ++nd;
    }
    ++i;
  }
  if (!digit || (i == head.size() && static_cast<int64_t>(i) < len))
    return fail(PGVD_ERR_INVALID, "no node count in the first 4096 bytes");
  // a negative count: the reference's resize throws ("-0" is 0)
  if (neg && N) return fail(PGVD_ERR_INVALID, "negative node count");
  if (N > INT32_MAX) return fail(PGVD_ERR_INVALID, "node count above 2^31 - 1");
  *O_out = N;
  if (N > cap) return fail(PGVD_ERR_CAPACITY, "octree buffer too small (*O_out nodes needed)");
  if (N == 0) return PGVD_OK;
  const int64_t tiles = parse_text_tiles(len);
  void *ws = nullptr;
  CHK(pool_get(ctx, "text_ws", 8 * text_ws_words(tiles), &ws));
  uint64_t res[2] = {0, 0};
  {
    Stage t(ctx, "parse_text", st);
    launch_parse_text(reinterpret_cast<const uint8_t *>(text), len, dim, N, octree,
                      static_cast<uint64_t *>(ws), st);
    CHK(launch_status("parse_text"));
  }
  HIPCHK(hipMemcpyAsync(res, static_cast<uint64_t *>(ws) + 1, 16, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const uint64_t need = static_cast<uint64_t>(N) * ((1u << dim) + 1) + 1;
  if (res[1] != ~0ull)
    return fail(PGVD_ERR_INVALID, "malformed integer at token " + std::to_string(res[1] - 1));
  if (res[0] < need)
    return fail(PGVD_ERR_INVALID, "text ends after " + std::to_string(res[0]) + " of " +
                                      std::to_string(need) + " tokens");
  return PGVD_OK;
}

namespace {
// Binary octree file: this header, then O records of 2^dim + 1 int32
// (children, leaf), little-endian, as the device holds them.
struct OctreeFileHeader {
  char magic[8];        // "PGVDOCT1"
  uint32_t dim;
  uint32_t record_bytes;
  int64_t O;
  uint64_t reserved;
};
static_assert(sizeof(OctreeFileHeader) == 32, "32-byte header");
constexpr char kOctMagic[8] = {'P', 'G', 'V', 'D', 'O', 'C', 'T', '1'};

struct File {
  FILE *f;
  explicit File(FILE *p) : f(p) {}
  ~File() {
    if (f) std::fclose(f);
  }
};
}  // namespace

int pgvd_octree_save(pgvd_ctx *ctx, const int32_t *octree, int64_t O, int dim, const char *path,
                     int format) {
  if (!ctx || !path || (O > 0 && !octree)) return fail(PGVD_ERR_INVALID, "null argument");
  if (dim != 2 && dim != 3) return fail(PGVD_ERR_INVALID, "dim must be 2 or 3");
  if (format != PGVD_OCTREE_TEXT && format != PGVD_OCTREE_BINARY)
    return fail(PGVD_ERR_INVALID, "format must be PGVD_OCTREE_TEXT or PGVD_OCTREE_BINARY");
  if (O < 0 || O > INT32_MAX) return fail(PGVD_ERR_INVALID, "O must be in [0, 2^31)");
  CHK(use_device(ctx));
  hipStream_t st = ctx->stream;
  const size_t rec = size_t(4) * ((1u << dim) + 1);
  std::vector<char> host;
  if (format == PGVD_OCTREE_TEXT) {
    int64_t len = 0;
    CHK(pgvd_octree_format_text(ctx, octree, O, dim, nullptr, 0, &len, nullptr));
    void *buf = nullptr;
    CHK(pool_get(ctx, "text_buf", static_cast<size_t>(len), &buf));
    CHK(pgvd_octree_format_text(ctx, octree, O, dim, static_cast<char *>(buf), len, &len, nullptr));
    host.resize(static_cast<size_t>(len));
    HIPCHK(hipMemcpyAsync(host.data(), buf, host.size(), hipMemcpyDeviceToHost, st));
  } else {
    OctreeFileHeader h{};
    std::memcpy(h.magic, kOctMagic, 8);
    h.dim = static_cast<uint32_t>(dim);
    h.record_bytes = static_cast<uint32_t>(rec);
    h.O = O;
    host.resize(sizeof h + rec * static_cast<size_t>(O));
    std::memcpy(host.data(), &h, sizeof h);
    if (O) HIPCHK(hipMemcpyAsync(host.data() + sizeof h, octree, rec * O, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  File f(std::fopen(path, "wb"));
  if (!f.f) return fail(PGVD_ERR_INVALID, std::string("cannot open for writing: ") + path);
  if (!host.empty() && std::fwrite(host.data(), 1, host.size(), f.f) != host.size())
    return fail(PGVD_ERR_INVALID, std::string("short write: ") + path);
  if (std::fclose(f.f) != 0) {
    f.f = nullptr;
    return fail(PGVD_ERR_INVALID, std::string("cannot close: ") + path);
  }
  f.f = nullptr;
  return PGVD_OK;
}

int pgvd_octree_load(pgvd_ctx *ctx, const char *path, int dim, int32_t *octree, int64_t cap,
                     int64_t *O_out) {
  if (!ctx || !path || !O_out || (cap > 0 && !octree)) return fail(PGVD_ERR_INVALID, "null argument");
  if (dim != 2 && dim != 3) return fail(PGVD_ERR_INVALID, "dim must be 2 or 3");
  if (cap < 0) return fail(PGVD_ERR_INVALID, "cap must be >= 0");
  CHK(use_device(ctx));
  hipStream_t st = ctx->stream;
  File f(std::fopen(path, "rb"));
  if (!f.f) return fail(PGVD_ERR_INVALID, std::string("cannot open: ") + path);
  std::vector<char> host;
  char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof chunk, f.f)) > 0) host.insert(host.end(), chunk, chunk + got);
  if (std::ferror(f.f)) return fail(PGVD_ERR_INVALID, std::string("read error: ") + path);
  const size_t rec = size_t(4) * ((1u << dim) + 1);
  if (host.size() >= 8 && std::memcmp(host.data(), kOctMagic, 8) == 0) {
    OctreeFileHeader h;
    if (host.size() < sizeof h) return fail(PGVD_ERR_INVALID, "truncated octree header");
    std::memcpy(&h, host.data(), sizeof h);
    if (h.dim != static_cast<uint32_t>(dim) || h.record_bytes != rec)
      return fail(PGVD_ERR_INVALID, "octree file holds another dimension");
    if (h.O < 0 || host.size() != sizeof h + rec * static_cast<size_t>(h.O))
      return fail(PGVD_ERR_INVALID, "octree file size does not match its node count");
    *O_out = h.O;
    if (h.O > cap) return fail(PGVD_ERR_CAPACITY, "octree buffer too small (*O_out nodes needed)");
    if (h.O) HIPCHK(hipMemcpyAsync(octree, host.data() + sizeof h, rec * h.O, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return PGVD_OK;
  }
  if (host.empty()) return fail(PGVD_ERR_INVALID, std::string("empty file: ") + path);
  void *buf = nullptr;
  CHK(pool_get(ctx, "text_buf", host.size(), &buf));
  HIPCHK(hipMemcpyAsync(buf, host.data(), host.size(), hipMemcpyHostToDevice, st));
  return pgvd_octree_parse_text(ctx, static_cast<const char *>(buf), static_cast<int64_t>(host.size()),
                                dim, octree, cap, O_out, nullptr);
}

int pgvd_radix_sort(pgvd_ctx *ctx, void *keys, uint32_t *vals, int64_t n, int key_bytes, int mbits,
                    pgvd_stream s) {
  if (!ctx || !keys) return fail(PGVD_ERR_INVALID, "null argument");
  if (n < 0 || n >= (int64_t(1) << 30) - 1) return fail(PGVD_ERR_TOO_LARGE, "n must be in [0, 2^30-1)");
  if (mbits < 1 || mbits > key_bytes * 8) return fail(PGVD_ERR_INVALID, "mbits out of range for key_bytes");
  if (n <= 1) return PGVD_OK;
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  if (key_bytes == 4) return radix_sort_impl<uint32_t>(ctx, static_cast<uint32_t *>(keys), vals, n, mbits, st);
  if (key_bytes == 8) return radix_sort_impl<uint64_t>(ctx, static_cast<uint64_t *>(keys), vals, n, mbits, st);
  if (key_bytes == 16) return radix_sort_impl<Key128>(ctx, static_cast<Key128 *>(keys), vals, n, mbits, st);
  return fail(PGVD_ERR_INVALID, "key_bytes must be 4, 8 or 16");
}

int pgvd_unique_sorted(pgvd_ctx *ctx, const void *keys_in, void *keys_out, int64_t n,
                       int key_bytes, int64_t *m_out, pgvd_stream s) {
  if (!ctx || !keys_in || !keys_out || !m_out) return fail(PGVD_ERR_INVALID, "null argument");
  if (n < 0 || n >= (int64_t(1) << 30) - 1) return fail(PGVD_ERR_TOO_LARGE, "n must be in [0, 2^30-1)");
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  if (key_bytes == 4) return unique_impl<uint32_t>(ctx, static_cast<const uint32_t *>(keys_in), static_cast<uint32_t *>(keys_out), n, m_out, st);
  if (key_bytes == 8) return unique_impl<uint64_t>(ctx, static_cast<const uint64_t *>(keys_in), static_cast<uint64_t *>(keys_out), n, m_out, st);
  if (key_bytes == 16) return unique_impl<Key128>(ctx, static_cast<const Key128 *>(keys_in), static_cast<Key128 *>(keys_out), n, m_out, st);
  return fail(PGVD_ERR_INVALID, "key_bytes must be 4, 8 or 16");
}

int pgvd_stream_scan(pgvd_ctx *ctx, const uint32_t *in, uint32_t *out, int64_t n, pgvd_stream s) {
  if (!ctx || !in || !out) return fail(PGVD_ERR_INVALID, "null argument");
  if (n < 0 || n >= (int64_t(1) << 30) - 1) return fail(PGVD_ERR_TOO_LARGE, "n must be in [0, 2^30-1)");
  if (n == 0) return PGVD_OK;
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  const ZeroLayout z = zero_layout(n, 4, 0, false);
  void *p;
  CHK(pool_get(ctx, "zero_stage", z.total, &p));
  char *zb = static_cast<char *>(p);
  launch_prep(ctx->st, reinterpret_cast<uint32_t *>(zb), z.total / 4, st);
  {
    Stage t(ctx, "scan", st);
    launch_scan_u32(in, out, n, reinterpret_cast<uint64_t *>(zb + z.lb_scan), ctx->st, st);
  }
  return launch_status("scan");
}

#define PGVD_BY_KEY(kb, CALL)                                                   \
  do {                                                                          \
    if ((kb) == 4) { using K = uint32_t; CALL; }                                \
    else if ((kb) == 8) { using K = uint64_t; CALL; }                           \
    else if ((kb) == 16) { using K = Key128; CALL; }                            \
    else return fail(PGVD_ERR_INVALID, "key_bytes must be 4, 8 or 16");        \
  } while (0)

int pgvd_bit_predicate(pgvd_ctx *ctx, const void *keys, int key_bytes, uint32_t *pred, int index,
                       int compared, int64_t n, pgvd_stream s) {
  if (!ctx || (n > 0 && (!keys || !pred))) return fail(PGVD_ERR_INVALID, "null argument");
  if (n <= 0) return PGVD_OK;
  CHK(use_device(ctx));
  PGVD_BY_KEY(key_bytes, launch_bit_predicate<K>(static_cast<const K *>(keys), pred, index,
                                                 (uint32_t)(compared != 0), n, pick(ctx, s)));
  return launch_status("bit predicate");
}

int pgvd_unique_predicate(pgvd_ctx *ctx, const void *keys, int key_bytes, uint32_t *pred,
                          int64_t n, pgvd_stream s) {
  if (!ctx || (n > 0 && (!keys || !pred))) return fail(PGVD_ERR_INVALID, "null argument");
  if (n <= 0) return PGVD_OK;
  CHK(use_device(ctx));
  PGVD_BY_KEY(key_bytes, launch_unique_predicate<K>(static_cast<const K *>(keys), pred, n,
                                                    pick(ctx, s)));
  return launch_status("unique predicate");
}

int pgvd_single_compact(pgvd_ctx *ctx, const void *in, void *out, int key_bytes,
                        const uint32_t *pred, const uint32_t *addr, int64_t n, pgvd_stream s) {
  if (!ctx || (n > 0 && (!in || !out || !pred || !addr))) return fail(PGVD_ERR_INVALID, "null argument");
  if (n <= 0) return PGVD_OK;
  CHK(use_device(ctx));
  PGVD_BY_KEY(key_bytes, launch_single_compact<K>(static_cast<const K *>(in), static_cast<K *>(out),
                                                  pred, addr, n, pick(ctx, s)));
  return launch_status("single compact");
}

int pgvd_double_compact(pgvd_ctx *ctx, const void *in, void *out, int key_bytes,
                        const uint32_t *pred, const uint32_t *addr, int64_t n, pgvd_stream s) {
  if (!ctx || (n > 0 && (!in || !out || !pred || !addr))) return fail(PGVD_ERR_INVALID, "null argument");
  if (n <= 0) return PGVD_OK;
  CHK(use_device(ctx));
  PGVD_BY_KEY(key_bytes, launch_double_compact<K>(static_cast<const K *>(in), static_cast<K *>(out),
                                                  pred, addr, n, pick(ctx, s)));
  return launch_status("double compact");
}

int pgvd_init_octree(pgvd_ctx *ctx, void *octree, int64_t nodes, int dim, pgvd_stream s) {
  if (!ctx || (nodes > 0 && !octree)) return fail(PGVD_ERR_INVALID, "null argument");
  if (dim != 2 && dim != 3) return fail(PGVD_ERR_INVALID, "dim must be 2 or 3");
  if (nodes <= 0) return PGVD_OK;
  CHK(use_device(ctx));
  launch_init_octree(static_cast<int32_t *>(octree), nodes, dim, pick(ctx, s));
  return launch_status("init octree");
}

int pgvd_build_brt(pgvd_ctx *ctx, const void *keys, int64_t m, int key_bytes, int mbits,
                   void *brt_out, int32_t *parent_out, pgvd_stream s) {
  if (!ctx || !keys || (m > 1 && (!brt_out || !parent_out))) return fail(PGVD_ERR_INVALID, "null argument");
  if (m < 1) return fail(PGVD_ERR_EMPTY, "no keys");
  if (m >= (int64_t(1) << 30) - 1) return fail(PGVD_ERR_TOO_LARGE, "m must be below 2^30-1");
  if (mbits < 1 || mbits > key_bytes * 8) return fail(PGVD_ERR_INVALID, "mbits out of range for key_bytes");
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  if (key_bytes == 4) return brt_impl<uint32_t>(ctx, static_cast<const uint32_t *>(keys), m, mbits, static_cast<BrtNodeT<uint32_t> *>(brt_out), parent_out, st);
  if (key_bytes == 8) return brt_impl<uint64_t>(ctx, static_cast<const uint64_t *>(keys), m, mbits, static_cast<BrtNodeT<uint64_t> *>(brt_out), parent_out, st);
  if (key_bytes == 16) return brt_impl<Key128>(ctx, static_cast<const Key128 *>(keys), m, mbits, static_cast<BrtNodeT<Key128> *>(brt_out), parent_out, st);
  return fail(PGVD_ERR_INVALID, "key_bytes must be 4, 8 or 16");
}

int pgvd_local_splits(pgvd_ctx *ctx, const void *brt, int64_t m, int key_bytes, int dim,
                      uint32_t *ls_out, pgvd_stream s) {
  if (!ctx || !ls_out || (m > 1 && !brt)) return fail(PGVD_ERR_INVALID, "null argument");
  if (dim != 2 && dim != 3) return fail(PGVD_ERR_INVALID, "dim must be 2 or 3");
  if (m < 1) return fail(PGVD_ERR_EMPTY, "no keys");
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  if (key_bytes == 4) return local_splits_impl<uint32_t>(ctx, static_cast<const BrtNodeT<uint32_t> *>(brt), m, dim, ls_out, st);
  if (key_bytes == 8) return local_splits_impl<uint64_t>(ctx, static_cast<const BrtNodeT<uint64_t> *>(brt), m, dim, ls_out, st);
  if (key_bytes == 16) return local_splits_impl<Key128>(ctx, static_cast<const BrtNodeT<Key128> *>(brt), m, dim, ls_out, st);
  return fail(PGVD_ERR_INVALID, "key_bytes must be 4, 8 or 16");
}

int pgvd_brt_to_octree(pgvd_ctx *ctx, const void *brt, const int32_t *parent, int64_t m,
                       int key_bytes, int dim, void *octree_out, int64_t capacity,
                       int64_t *octree_size, pgvd_stream s) {
  if (!ctx || !octree_out || !octree_size || (m > 1 && (!brt || !parent)))
    return fail(PGVD_ERR_INVALID, "null argument");
  if (dim != 2 && dim != 3) return fail(PGVD_ERR_INVALID, "dim must be 2 or 3");
  if (m < 1) return fail(PGVD_ERR_EMPTY, "no keys");
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  int32_t *o = static_cast<int32_t *>(octree_out);
  if (key_bytes == 4) return brt_to_octree_impl<uint32_t>(ctx, static_cast<const BrtNodeT<uint32_t> *>(brt), parent, m, dim, o, capacity, octree_size, st);
  if (key_bytes == 8) return brt_to_octree_impl<uint64_t>(ctx, static_cast<const BrtNodeT<uint64_t> *>(brt), parent, m, dim, o, capacity, octree_size, st);
  if (key_bytes == 16) return brt_to_octree_impl<Key128>(ctx, static_cast<const BrtNodeT<Key128> *>(brt), parent, m, dim, o, capacity, octree_size, st);
  return fail(PGVD_ERR_INVALID, "key_bytes must be 4, 8 or 16");
}

int pgvd_build_octree(pgvd_ctx *ctx, const void *pts, int64_t n, int dim, int bits,
                      const uint32_t *labels, const pgvd_build_opts *opts, pgvd_stream s) {
  if (!ctx || !pts) return fail(PGVD_ERR_INVALID, "null argument");
  CHK(check_n(n));
  CHK(check_dims(dim, bits));
  if (opts && opts->quantized_input && bits > 31) return fail(PGVD_ERR_INVALID, "int32 grid input allows bits <= 31");
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  ctx->last.valid = false;
  const int kb = key_bytes_for(dim * bits);
  if (kb == 4) return build_impl<uint32_t>(ctx, pts, n, dim, bits, labels, opts, st);
  if (kb == 8) return build_impl<uint64_t>(ctx, pts, n, dim, bits, labels, opts, st);
  return build_impl<Key128>(ctx, pts, n, dim, bits, labels, opts, st);
}

int pgvd_build_octree_noalloc(const void *pts, int64_t n, int dim, int bits,
                              const uint32_t *labels, const pgvd_build_opts *opts,
                              int32_t *octree, int64_t octree_cap, uint32_t *node_conflict,
                              pgvd_build_state *state, pgvd_stream s) {
  if (!pts || !octree || !state) return fail(PGVD_ERR_INVALID, "null argument");
  if (labels && !node_conflict)
    return fail(PGVD_ERR_INVALID, "labels need a node_conflict buffer");
  if (octree_cap < 1) return fail(PGVD_ERR_INVALID, "octree_cap must be at least 1");
  CHK(check_n(n));
  CHK(check_dims(dim, bits));
  if (opts && opts->quantized_input && bits > 31) return fail(PGVD_ERR_INVALID, "int32 grid input allows bits <= 31");
  void *base = nullptr;
  CHK(noalloc_workspace(&base));
  hipStream_t st = static_cast<hipStream_t>(s);
  int32_t *out4 = reinterpret_cast<int32_t *>(state);
  char *ws = static_cast<char *>(base);
  const int kb = key_bytes_for(dim * bits);
  if (kb == 4) return noalloc_impl<uint32_t>(ws, pts, n, dim, bits, labels, opts, octree, octree_cap, node_conflict, out4, st);
  if (kb == 8) return noalloc_impl<uint64_t>(ws, pts, n, dim, bits, labels, opts, octree, octree_cap, node_conflict, out4, st);
  return noalloc_impl<Key128>(ws, pts, n, dim, bits, labels, opts, octree, octree_cap, node_conflict, out4, st);
}

int pgvd_build_octree_finish(pgvd_ctx *ctx, pgvd_octree_result *out) {
  if (!ctx || !out) return fail(PGVD_ERR_INVALID, "null argument");
  if (!ctx->last.valid) return fail(PGVD_ERR_INVALID, "no build in flight on this context");
  CHK(use_device(ctx));
  HIPCHK(hipEventSynchronize(ctx->done));
  LastBuild &L = ctx->last;
  if (L.exact_deferred && ctx->st_host->malformed) {
    L.exact_deferred = false;
    if (L.kb == 4) CHK(exact_path<uint32_t>(ctx));
    else if (L.kb == 8) CHK(exact_path<uint64_t>(ctx));
    else CHK(exact_path<Key128>(ctx));
  }
  for (int attempt = 0; attempt < 2 && (ctx->st_host->status & ST_CAPACITY); ++attempt) {
    const int64_t need = (int64_t)ctx->st_host->octree_size;
    if (L.kb == 4) CHK(refill_octree<uint32_t>(ctx, need + need / 4 + 64));
    else if (L.kb == 8) CHK(refill_octree<uint64_t>(ctx, need + need / 4 + 64));
    else CHK(refill_octree<Key128>(ctx, need + need / 4 + 64));
  }
  const RunState &R = *ctx->st_host;
  if (R.status & ST_CAPACITY) return fail(PGVD_ERR_INTERNAL, "octree refill did not fit");
  CHK(status_error(R.status));
  std::memset(out, 0, sizeof(*out));
  out->n = L.n;
  out->m = R.m;
  out->octree_size = R.octree_size;
  out->dim = L.dim;
  out->bits = L.bits;
  out->mbits = L.mbits;
  out->key_bytes = L.kb;
  out->malformed = R.malformed ? 1 : 0;
  out->bucket_sort = R.msd_ok ? 1 : 0;
  out->keys = L.keys_u;
  out->brt = L.brt;
  out->parent = L.parent;
  out->prefix_sums = L.ps;
  out->octree = static_cast<const int32_t *>(ctx->pool["octree"].p);
  out->key_conflict = L.conflict_bits;
  out->node_conflict = L.labels ? static_cast<const uint32_t *>(ctx->pool["node_conflict"].p) : nullptr;
  return PGVD_OK;
}

}  // extern "C"
