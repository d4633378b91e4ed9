// batch.hip -- the batched form of the C ABI (pgvd_batch_*).
//
// A plan holds B independent point sets of uneven size.  Every per-key
// array (keys, BRT, prefix sums, parent links) gives instance s the slots
// [kofs[s], kofs[s] + n_s + 1); every stage is one launch over tiles that
// never straddle two instances, described by host-built tile tables, so the
// whole batch costs the same dozen launches as one plain build.  The launch
// sequence is captured once into a HIP graph and replayed.  Each instance
// goes through exactly the plain pipeline's steps (padding zero key, stable
// per-instance sort, unique, BRT, local splits, gather), so its output is
// bit-identical to pgvd_build_octree on that instance alone.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "capi_internal.hpp"

using namespace pgvd;
using namespace pgvd_capi;

struct pgvd_batch {
  pgvd_ctx *ctx = nullptr;
  hipStream_t s = nullptr;
  int dim = 0, bits = 0, mbits = 0, kb = 0, passes = 0;
  bool quantized = false, clamped = false, use_graph = true;
  int64_t B = 0, total_pts = 0, total_slots = 0;
  int sort_ipt = 16, brt_rows = 4;
  int digit_bits = 8, sort_waves = 3;
  std::vector<SegInfo> h_seg;
  std::vector<int64_t> h_m, h_O, h_koff, h_ooff;
  // device
  SegInfo *seg = nullptr;
  uint32_t *kofs = nullptr;
  SegState *ss = nullptr;
  SegState *ss_host = nullptr;      // pinned
  RunState *st = nullptr;
  RunState *st_host = nullptr;      // pinned
  uint32_t *tables = nullptr;
  BatchDev dev;
  void *keys_a = nullptr, *keys_b = nullptr, *brt = nullptr;
  uint32_t *ps = nullptr, *defer = nullptr;
  int32_t *parent = nullptr;
  char *zero = nullptr;
  size_t zero_bytes = 0, off_hist = 0, off_sort = 0, off_lbu = 0, off_lbl = 0;
  int32_t *octree = nullptr;
  int64_t oct_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void *graph_pts = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int parity = 0;
  bool ran = false;
};

namespace {

template <typename T>
int dalloc(T **p, size_t bytes) {
  void *q = nullptr;
  if (bytes == 0) bytes = 256;
  if (hipMalloc(&q, bytes) != hipSuccess) {
    (void)hipGetLastError();   // not left for the next launch check to report
    return fail(PGVD_ERR_ALLOC, "hipMalloc failed for a batch buffer (" + std::to_string(bytes) + " bytes)");
  }
  *p = static_cast<T *>(q);
  return PGVD_OK;
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

// Tiles of `tile` units over `units(s)` units per instance.
struct HostTable {
  std::vector<uint32_t> seg, first;
};

template <typename F>
HostTable make_table(const std::vector<SegInfo> &si, int64_t tile, F units) {
  HostTable t;
  t.first.resize(si.size() + 1);
  uint32_t n = 0;
  for (size_t s = 0; s < si.size(); ++s) {
    t.first[s] = n;
    const int64_t u = units(si[s]);
    const int64_t k = (u + tile - 1) / tile;
    for (int64_t j = 0; j < k; ++j) t.seg.push_back((uint32_t)s);
    n += (uint32_t)k;
  }
  t.first[si.size()] = n;
  return t;
}

void release_graph(pgvd_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_pts = nullptr;
}

// (Re)builds the five tile tables for the plan's sort_ipt / brt_rows, and the
// look-back workspace they size.
int build_tables(pgvd_batch *b) {
  const std::vector<SegInfo> &si = b->h_seg;
  const HostTable chunk = make_table(si, kEncodeChunk, [](const SegInfo &x) { return (int64_t)x.n; });
  const HostTable sort = make_table(si, (int64_t)b->sort_ipt * kBlock, [](const SegInfo &x) { return (int64_t)x.n; });
  const HostTable scan = make_table(si, kScanTile, [](const SegInfo &x) { return (int64_t)x.n + 1; });
  const HostTable brt = make_table(si, (int64_t)b->brt_rows * kBlock, [](const SegInfo &x) { return (int64_t)x.n + 1; });
  const HostTable node = make_table(si, kBlock, [](const SegInfo &x) { return (int64_t)x.n + 1; });
  std::vector<uint32_t> all;
  size_t off[10];
  const HostTable *ts[5] = {&chunk, &sort, &scan, &brt, &node};
  for (int i = 0; i < 5; ++i) {
    off[2 * i] = all.size();
    all.insert(all.end(), ts[i]->seg.begin(), ts[i]->seg.end());
    off[2 * i + 1] = all.size();
    all.insert(all.end(), ts[i]->first.begin(), ts[i]->first.end());
  }
  dfree(b->tables);
  CHK(dalloc(&b->tables, all.size() * 4));
  HIPCHK(hipMemcpy(b->tables, all.data(), all.size() * 4, hipMemcpyHostToDevice));
  BatchDev &d = b->dev;
  TileTable *tt[5] = {&d.chunk, &d.sort, &d.scan, &d.brt, &d.node};
  int64_t *cnt[5] = {&d.n_chunk, &d.n_sort, &d.n_scan, &d.n_brt, &d.n_node};
  for (int i = 0; i < 5; ++i) {
    tt[i]->seg = b->tables + off[2 * i];
    tt[i]->first = b->tables + off[2 * i + 1];
    *cnt[i] = (int64_t)ts[i]->seg.size();
  }
  d.sort_ipt = b->sort_ipt;
  d.brt_rows = b->brt_rows;
  d.digit_bits = b->digit_bits;
  d.sort_waves = b->sort_waves;
  b->passes = (b->mbits + b->digit_bits - 1) / b->digit_bits;
  // zero-initialised workspace: digit histograms, sort look-back (all
  // passes), unique and local-split scan look-back
  size_t o = 0;
  b->off_hist = o;
  o += align256((size_t)b->B * b->passes * 256 * 4);
  b->off_sort = o;
  o += align256((size_t)b->passes * (size_t)d.n_sort * 256 * 4);
  b->off_lbu = o;
  o += align256((size_t)d.n_scan * 8);
  b->off_lbl = o;
  o += align256((size_t)d.n_scan * 8);
  if (o > b->zero_bytes) {
    dfree(b->zero);
    CHK(dalloc(&b->zero, o));
    b->zero_bytes = o;
  }
  release_graph(b);
  return PGVD_OK;
}

int ensure_octree(pgvd_batch *b, int64_t nodes) {
  if (nodes <= b->oct_cap) return PGVD_OK;
  const int stride = (1 << b->dim) + 1;
  dfree(b->octree);
  CHK(dalloc(&b->octree, (size_t)nodes * stride * 4));
  b->oct_cap = nodes;
  release_graph(b);
  return PGVD_OK;
}

template <typename K>
void enqueue_typed(pgvd_batch *b, const void *pts) {
  hipStream_t s = b->s;
  const BatchDev &d = b->dev;
  K *kA = static_cast<K *>(b->keys_a), *kB = static_cast<K *>(b->keys_b);
  launch_prep(b->st, reinterpret_cast<uint32_t *>(b->zero), b->zero_bytes / 4, s);
  launch_seg_init(b->ss, b->B, s);
  const int64_t width = (int64_t)1 << b->bits;
  const float effw = (float)(b->clamped ? width : width - 1);
  if (!b->quantized) launch_bbox_seg(static_cast<const float *>(pts), b->dim, d, s);
  uint32_t *bhist = reinterpret_cast<uint32_t *>(b->zero + b->off_hist);
  launch_encode_seg<K>(pts, b->quantized ? 1 : 0, b->dim, b->bits, effw, d, kA, bhist, b->passes,
                       b->st, s);
  b->parity = launch_onesweep_seg<K>(kA, kB, d, b->passes, bhist,
                                     reinterpret_cast<uint32_t *>(b->zero + b->off_sort), b->st, s);
  K *sorted = b->parity ? kB : kA;
  K *uniq = b->parity ? kA : kB;
  launch_unique_seg<K>(sorted, uniq, d, reinterpret_cast<uint64_t *>(b->zero + b->off_lbu), b->st,
                       s);
  uint32_t *lsx = reinterpret_cast<uint32_t *>(sorted);   // free after unique
  launch_brt_seg<K>(uniq, b->mbits, b->dim, d, static_cast<BrtNodeT<K> *>(b->brt), b->parent,
                    lsx, b->defer, b->st, s);
  launch_ls_scan_seg(lsx, b->ps, d, reinterpret_cast<uint64_t *>(b->zero + b->off_lbl), b->st, s);
  launch_oct_offsets(d, b->st, s);
  launch_gather_seg<K>(static_cast<const BrtNodeT<K> *>(b->brt), b->ps, lsx, b->dim, d,
                       b->octree, b->oct_cap, b->defer, b->st, s);
}

void enqueue(pgvd_batch *b, const void *pts) {
  if (b->kb == 4) enqueue_typed<uint32_t>(b, pts);
  else if (b->kb == 8) enqueue_typed<uint64_t>(b, pts);
  else enqueue_typed<Key128>(b, pts);
}

int run_impl(pgvd_batch *b, const void *pts) {
  CHK(use_device(b->ctx));
  HIPCHK(hipEventRecord(b->ev0, b->s));
  if (b->use_graph) {
    if (!b->exec || b->graph_pts != pts) {
      release_graph(b);
      HIPCHK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
      enqueue(b, pts);
      const hipError_t le = hipGetLastError();
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(b->s, &g);
      if (le != hipSuccess || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(PGVD_ERR_HIP, std::string("batch graph capture: ") +
                                      hipGetErrorString(le != hipSuccess ? le : ce));
      }
      b->graph = g;
      HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_pts = pts;
    }
    HIPCHK(hipGraphLaunch(b->exec, b->s));
  } else {
    enqueue(b, pts);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) return fail(PGVD_ERR_HIP, std::string("batch launch: ") + hipGetErrorString(le));
  }
  HIPCHK(hipEventRecord(b->ev1, b->s));
  HIPCHK(hipMemcpyAsync(b->st_host, b->st, sizeof(RunState), hipMemcpyDeviceToHost, b->s));
  HIPCHK(hipMemcpyAsync(b->ss_host, b->ss, (size_t)b->B * sizeof(SegState),
                        hipMemcpyDeviceToHost, b->s));
  b->ran = true;
  return PGVD_OK;
}

template <typename K>
void regather_typed(pgvd_batch *b) {
  K *sorted = static_cast<K *>(b->parity ? b->keys_b : b->keys_a);
  launch_gather_seg<K>(static_cast<const BrtNodeT<K> *>(b->brt), b->ps,
                       reinterpret_cast<const uint32_t *>(sorted), b->dim, b->dev, b->octree,
                       b->oct_cap, b->defer, b->st, b->s);
}

int finish_impl(pgvd_batch *b) {
  if (!b->ran) return fail(PGVD_ERR_INVALID, "pgvd_batch_finish before pgvd_batch_run");
  CHK(use_device(b->ctx));
  HIPCHK(hipStreamSynchronize(b->s));
  uint32_t stt = b->st_host->status;
  if (stt & ST_CAPACITY) {
    // grow to the reported total and fill again from the kept BRT / prefix
    // sums; the gather appended nothing to its overflow list when it failed
    const int64_t need = (int64_t)b->st_host->octree_size;
    CHK(ensure_octree(b, need + need / 8 + 1024));
    b->st_host->status &= ~(uint32_t)ST_CAPACITY;
    HIPCHK(hipMemcpyAsync(&b->st->status, &b->st_host->status, 4, hipMemcpyHostToDevice, b->s));
    if (b->kb == 4) regather_typed<uint32_t>(b);
    else if (b->kb == 8) regather_typed<uint64_t>(b);
    else regather_typed<Key128>(b);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) return fail(PGVD_ERR_HIP, std::string("batch refill: ") + hipGetErrorString(le));
    HIPCHK(hipMemcpyAsync(b->st_host, b->st, sizeof(RunState), hipMemcpyDeviceToHost, b->s));
    HIPCHK(hipStreamSynchronize(b->s));
    stt = b->st_host->status;
  }
  if (stt & ST_NEG_COORD) return fail(PGVD_ERR_BAD_COORD, "bad coordinate: a point quantized below zero");
  if (stt & ST_DUP_KEY) return fail(PGVD_ERR_DUP_KEY, "BRT input has equal neighbouring keys");
  if (stt & (ST_STACK | ST_WALK)) return fail(PGVD_ERR_INTERNAL, "octree walk exceeded its bound");
  if (stt & ST_TOO_LARGE) return fail(PGVD_ERR_TOO_LARGE, "batch octree exceeds 2^32 nodes");
  if (stt & ST_CAPACITY) return fail(PGVD_ERR_CAPACITY, "octree buffer too small after refill");
  if (b->st_host->malformed)
    return fail(PGVD_ERR_INTERNAL, "an instance's BRT took an FP32-quirk split (instances are limited to 2^23 points)");
  for (int64_t i = 0; i < b->B; ++i) {
    b->h_m[i] = b->ss_host[i].m;
    b->h_O[i] = b->ss_host[i].octree_size;
    b->h_ooff[i] = b->ss_host[i].oct_off;
  }
  return PGVD_OK;
}

}  // namespace

extern "C" {

int pgvd_batch_create(pgvd_ctx *ctx, const int64_t *sizes, int64_t count, int dim, int bits,
                      const pgvd_batch_opts *opts, pgvd_batch **out) {
  if (!ctx || !sizes || !out) return fail(PGVD_ERR_INVALID, "null argument");
  *out = nullptr;
  if (count <= 0) return fail(PGVD_ERR_EMPTY, "a batch needs at least one instance");
  if (dim != 2 && dim != 3) return fail(PGVD_ERR_INVALID, "dim must be 2 or 3");
  if (bits < 1 || bits * dim > 128) return fail(PGVD_ERR_INVALID, "bits must satisfy 1 <= bits*dim <= 128");
  int64_t total = 0, slots = 0;
  for (int64_t i = 0; i < count; ++i) {
    if (sizes[i] <= 0) return fail(PGVD_ERR_EMPTY, "Zero points not supported (instance " + std::to_string(i) + ")");
    if (sizes[i] > kMaxSegPoints)
      return fail(PGVD_ERR_TOO_LARGE, "batched instances hold at most 2^23 points; use pgvd_build_octree");
    total += sizes[i];
    slots += sizes[i] + 1;
  }
  if (slots >= (int64_t(1) << 31)) return fail(PGVD_ERR_TOO_LARGE, "batch exceeds 2^31 key slots");
  CHK(use_device(ctx));
  pgvd_batch *b = new pgvd_batch();
  b->ctx = ctx;
  b->s = ctx->stream;
  b->dim = dim;
  b->bits = bits;
  b->mbits = dim * bits;
  b->kb = key_bytes_for(b->mbits);
  b->passes = (b->mbits + 7) / 8;
  b->quantized = opts && opts->quantized_input;
  b->clamped = opts && opts->clamped;
  b->use_graph = !opts || opts->use_graph;
  b->B = count;
  b->total_pts = total;
  b->total_slots = slots;
  b->sort_ipt = b->kb == 16 ? 7 : 15;
  b->brt_rows = 4;
  if (opts && opts->sort_ipt) {
    if (!onesweep_seg_ipt_ok(b->kb, opts->sort_ipt)) { delete b; return fail(PGVD_ERR_INVALID, "unsupported sort_ipt for this key width"); }
    b->sort_ipt = opts->sort_ipt;
  }
  if (opts && opts->brt_rows) {
    if (!brt_seg_rows_ok(b->kb, opts->brt_rows)) { delete b; return fail(PGVD_ERR_INVALID, "unsupported brt_rows for this key width"); }
    b->brt_rows = opts->brt_rows;
  }
  if (opts && opts->digit_bits) {
    if (!onesweep_seg_digit_ok(b->kb, opts->digit_bits)) { delete b; return fail(PGVD_ERR_INVALID, "unsupported digit_bits for this key width"); }
    b->digit_bits = opts->digit_bits;
  }
  if (opts && opts->sort_waves) {
    if (!onesweep_seg_waves_ok(opts->sort_waves)) { delete b; return fail(PGVD_ERR_INVALID, "sort_waves must be 2, 3 or 4"); }
    b->sort_waves = opts->sort_waves;
  }
  b->passes = (b->mbits + b->digit_bits - 1) / b->digit_bits;
  b->h_seg.resize(count);
  b->h_koff.resize(count);
  b->h_m.assign(count, 0);
  b->h_O.assign(count, 0);
  b->h_ooff.assign(count, 0);
  std::vector<uint32_t> kofs(count + 1);
  int64_t p = 0, k = 0;
  for (int64_t i = 0; i < count; ++i) {
    SegInfo si{};
    si.pofs = (uint64_t)p;
    si.n = (uint32_t)sizes[i];
    si.kofs = (uint32_t)k;
    si.prepend = ref_next_pow2((int)sizes[i]) != sizes[i] ? 1u : 0u;   // padding (Q1)
    b->h_seg[i] = si;
    b->h_koff[i] = k;
    kofs[i] = (uint32_t)k;
    p += sizes[i];
    k += sizes[i] + 1;
  }
  kofs[count] = (uint32_t)k;
  auto bail = [&](int rc) {
    pgvd_batch_destroy(b);
    return rc;
  };
  int rc;
  if ((rc = dalloc(&b->seg, count * sizeof(SegInfo)))) return bail(rc);
  if ((rc = dalloc(&b->kofs, (count + 1) * 4))) return bail(rc);
  if ((rc = dalloc(&b->ss, count * sizeof(SegState)))) return bail(rc);
  if ((rc = dalloc(&b->st, sizeof(RunState)))) return bail(rc);
  if (hipHostMalloc((void **)&b->ss_host, count * sizeof(SegState)) != hipSuccess ||
      hipHostMalloc((void **)&b->st_host, sizeof(RunState)) != hipSuccess)
    return bail(fail(PGVD_ERR_ALLOC, "hipHostMalloc failed"));
  if (hipMemcpy(b->seg, b->h_seg.data(), count * sizeof(SegInfo), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->kofs, kofs.data(), (count + 1) * 4, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(PGVD_ERR_HIP, "batch table upload failed"));
  const size_t kbytes = (size_t)slots * b->kb;
  if ((rc = dalloc(&b->keys_a, kbytes))) return bail(rc);
  if ((rc = dalloc(&b->keys_b, kbytes))) return bail(rc);
  const size_t nb = b->kb == 4 ? sizeof(BrtNodeT<uint32_t>) : (b->kb == 8 ? sizeof(BrtNodeT<uint64_t>) : sizeof(BrtNodeT<Key128>));
  if ((rc = dalloc(&b->brt, (size_t)slots * nb))) return bail(rc);
  if ((rc = dalloc(&b->ps, (size_t)slots * 4))) return bail(rc);
  if ((rc = dalloc(&b->parent, (size_t)slots * 4))) return bail(rc);
  if ((rc = dalloc(&b->defer, (size_t)slots * 4))) return bail(rc);
  b->dev.seg = b->seg;
  b->dev.kofs = b->kofs;
  b->dev.ss = b->ss;
  b->dev.B = count;
  if ((rc = build_tables(b))) return bail(rc);
  if ((rc = ensure_octree(b, std::max<int64_t>(slots, 1024)))) return bail(rc);
  if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
    return bail(fail(PGVD_ERR_HIP, "hipEventCreate failed"));
  *out = b;
  return PGVD_OK;
}

int pgvd_batch_destroy(pgvd_batch *b) {
  if (!b) return PGVD_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->s);
  release_graph(b);
  dfree(b->seg);
  dfree(b->kofs);
  dfree(b->ss);
  dfree(b->st);
  if (b->ss_host) (void)hipHostFree(b->ss_host);
  if (b->st_host) (void)hipHostFree(b->st_host);
  dfree(b->tables);
  dfree(b->keys_a);
  dfree(b->keys_b);
  dfree(b->brt);
  dfree(b->ps);
  dfree(b->parent);
  dfree(b->defer);
  dfree(b->zero);
  dfree(b->octree);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  delete b;
  return PGVD_OK;
}

int pgvd_batch_run(pgvd_batch *b, const void *pts) {
  if (!b || !pts) return fail(PGVD_ERR_INVALID, "null argument");
  return run_impl(b, pts);
}

int pgvd_batch_finish(pgvd_batch *b, pgvd_batch_result *out) {
  if (!b || !out) return fail(PGVD_ERR_INVALID, "null argument");
  CHK(finish_impl(b));
  out->count = b->B;
  out->dim = b->dim;
  out->bits = b->bits;
  out->mbits = b->mbits;
  out->key_bytes = b->kb;
  out->octree_stride = (1 << b->dim) + 1;
  out->total_octree = b->st_host->octree_size;
  out->m = b->h_m.data();
  out->octree_size = b->h_O.data();
  out->key_offset = b->h_koff.data();
  out->octree_offset = b->h_ooff.data();
  out->keys = b->parity ? b->keys_a : b->keys_b;
  out->brt = b->brt;
  out->parent = b->parent;
  out->prefix_sums = b->ps;
  out->octree = b->octree;
  return PGVD_OK;
}

int pgvd_batch_last_ms(pgvd_batch *b, float *ms) {
  if (!b || !ms) return fail(PGVD_ERR_INVALID, "null argument");
  if (!b->ran) return fail(PGVD_ERR_INVALID, "no run yet");
  HIPCHK(hipEventSynchronize(b->ev1));
  HIPCHK(hipEventElapsedTime(ms, b->ev0, b->ev1));
  return PGVD_OK;
}

int pgvd_batch_tune(pgvd_batch *b, const void *pts, int *sort_ipt, int *brt_rows,
                    int *digit_bits, int *sort_waves) {
  if (!b || !pts) return fail(PGVD_ERR_INVALID, "null argument");
  CHK(use_device(b->ctx));
  const bool graph = b->use_graph;
  b->use_graph = false;
  // the fastest of three runs (after one that warms up) of the current setting
  auto time_it = [&](float *t) -> int {
    CHK(build_tables(b));
    *t = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      CHK(run_impl(b, pts));
      CHK(finish_impl(b));
      float ms = 0;
      CHK(pgvd_batch_last_ms(b, &ms));
      if (rep > 0) *t = std::min(*t, ms);
    }
    return PGVD_OK;
  };
  // Coordinate descent: sort tile and BRT window together, then the digit
  // width, then the sort kernel's waves per SIMD, each keeping the best.
  const std::vector<int> ipts = b->kb == 16 ? std::vector<int>{4, 7, 8} : std::vector<int>{8, 12, 15, 16};
  const std::vector<int> rows = {2, 4};
  const std::vector<int> digits = b->kb == 16 ? std::vector<int>{8} : std::vector<int>{6, 7, 8};
  const std::vector<int> waves = {2, 3, 4};
  float best = 1e30f, t = 0;
  int bi = b->sort_ipt, br = b->brt_rows, bd = b->digit_bits, bw = b->sort_waves;
  int rc = PGVD_OK;
  for (int ipt : ipts) {
    for (int r : rows) {
      b->sort_ipt = ipt;
      b->brt_rows = r;
      if ((rc = time_it(&t))) break;
      if (t < best) { best = t; bi = ipt; br = r; }
    }
    if (rc) break;
  }
  b->sort_ipt = bi;
  b->brt_rows = br;
  for (size_t k = 0; k < digits.size() && rc == PGVD_OK; ++k) {
    if (digits[k] == bd) continue;
    b->digit_bits = digits[k];
    if ((rc = time_it(&t))) break;
    if (t < best) { best = t; bd = digits[k]; }
  }
  b->digit_bits = bd;
  for (size_t k = 0; k < waves.size() && rc == PGVD_OK; ++k) {
    if (waves[k] == bw) continue;
    b->sort_waves = waves[k];
    if ((rc = time_it(&t))) break;
    if (t < best) { best = t; bw = waves[k]; }
  }
  b->use_graph = graph;
  b->sort_ipt = bi;
  b->brt_rows = br;
  b->digit_bits = bd;
  b->sort_waves = bw;
  CHK(build_tables(b));
  if (rc) return rc;
  if (sort_ipt) *sort_ipt = bi;
  if (brt_rows) *brt_rows = br;
  if (digit_bits) *digit_bits = bd;
  if (sort_waves) *sort_waves = bw;
  return PGVD_OK;
}

}  // extern "C"
