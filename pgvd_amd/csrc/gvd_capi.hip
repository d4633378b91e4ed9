// gvd_capi.hip -- C ABI of the GVD refinement step (include/pgvd.h):
// FindMultiCells' segments, walks and sampling, Geom::FitBoxes, and the
// Karras iteration of Octree2::build (viewer/Octree2.cpp:127-190).  The
// kernels are in gvd.hip and walk.hip; every intermediate stays on the
// device and lives in the context's keyed pool.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/pgvd.h"
#include "capi_internal.hpp"
#include "kernels.hpp"

using namespace pgvd;
using namespace pgvd_capi;

static_assert(sizeof(pgvd_cell_intersections) == 160, "CellIntersections record is 40 words");

namespace {

const char *fit_error_text(int code) {
  switch (code) {
    case 1: return "FitBoxes doesn't support intersections";
    case 2: return "FitBoxes: no end but not intersecting";
    case 3: return "FitBoxes: closest points between A and B must include one endpoint";
    case 4: return "FitBoxes: no convergence within 4096 rounds";
    default: return "FitBoxes failed";
  }
}

// The walk of pgvd_cell_walk into pooled offsets and cells; *total_out
// cells.
int walk_pooled(pgvd_ctx *ctx, const float *segs, int64_t n, int bits, const int32_t *octree,
                int64_t O, int32_t *status, int64_t **offsets_out, int32_t **cells_out,
                int64_t *total_out, hipStream_t st) {
  void *offsets = nullptr, *counts = nullptr, *ws = nullptr, *scratch = nullptr, *cells = nullptr;
  CHK(pool_get(ctx, "mc_offsets", 8 * static_cast<size_t>(n + 1), &offsets));
  CHK(pool_get(ctx, "walk_counts", 4 * static_cast<size_t>(n), &counts));
  CHK(pool_get(ctx, "walk_ws", 8 * static_cast<size_t>(walk_scan_ws_words(n)), &ws));
  constexpr int kLog2K = 5;
  const size_t spec_bytes = (static_cast<size_t>(n) << kLog2K) * 20;
  if (spec_bytes <= (size_t(2) << 30)) CHK(pool_get(ctx, "walk_scratch", spec_bytes, &scratch));
  {
    Stage t(ctx, "gvd_walk", st);
    launch_cell_walk_count(segs, n, bits, octree, O, static_cast<uint32_t *>(counts), status,
                           static_cast<int64_t *>(offsets), static_cast<uint64_t *>(ws),
                           static_cast<int32_t *>(scratch), kLog2K, st);
  }
  HIPCHK(hipGetLastError());
  int64_t total = 0;
  HIPCHK(hipMemcpyAsync(&total, static_cast<int64_t *>(offsets) + n, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (total >= int64_t(0xFFFFFFFF)) return fail(PGVD_ERR_TOO_LARGE, "more than 2^32 - 2 visited cells");
  CHK(pool_get(ctx, "mc_cells", 20 * static_cast<size_t>(std::max<int64_t>(total, 1)), &cells));
  {
    Stage t(ctx, "gvd_walk_write", st);
    launch_cell_walk_write(segs, n, bits, octree, O, static_cast<const int64_t *>(offsets),
                           static_cast<int32_t *>(cells), static_cast<const int32_t *>(scratch),
                           static_cast<uint32_t *>(counts), kLog2K, st);
  }
  HIPCHK(hipGetLastError());
  *offsets_out = static_cast<int64_t *>(offsets);
  *cells_out = static_cast<int32_t *>(cells);
  *total_out = total;
  return PGVD_OK;
}

int multi_cells_impl(pgvd_ctx *ctx, const float *segs, const int32_t *labels, int64_t n, int bits,
                     const int32_t *octree, int64_t O, int32_t *ci, int32_t *walk_status,
                     hipStream_t st) {
  int64_t *offsets = nullptr;
  int32_t *cells = nullptr;
  int64_t total = 0;
  if (n > 0) CHK(walk_pooled(ctx, segs, n, bits, octree, O, walk_status, &offsets, &cells, &total, st));
  void *parts = nullptr, *vlabel = nullptr, *first = nullptr, *second = nullptr, *best = nullptr;
  CHK(pool_get(ctx, "mc_parts", 16 * static_cast<size_t>(std::max<int64_t>(total, 1)), &parts));
  CHK(pool_get(ctx, "mc_vlabel", 4 * static_cast<size_t>(std::max<int64_t>(total, 1)), &vlabel));
  CHK(pool_get(ctx, "mc_first", 16 * static_cast<size_t>(O), &first));
  CHK(pool_get(ctx, "mc_second", 16 * static_cast<size_t>(O), &second));
  CHK(pool_get(ctx, "mc_best", 64 * static_cast<size_t>(O), &best));
  {
    Stage t(ctx, "gvd_multi_cells", st);
    launch_multi_cells(segs, labels, n, offsets, cells, total, O, static_cast<float4 *>(parts),
                       static_cast<int32_t *>(vlabel), static_cast<uint32_t *>(first),
                       static_cast<uint32_t *>(second), static_cast<unsigned long long *>(best),
                       ci, st);
  }
  HIPCHK(hipGetLastError());
  return PGVD_OK;
}

// FitBoxes over items (pairs or CellIntersections slots) in three steps:
// fit_count enqueues the counting pass and the scans of the counts;
// fit_totals reads the totals, the first error (~0, or item << 8 | code)
// and, optionally, one more device word back and waits; fit_write enqueues
// the writing pass into buffers that hold the totals.
struct FitWork {
  uint32_t *ns = nullptr, *nb = nullptr;
  int64_t *so = nullptr, *bo = nullptr;
  unsigned long long *err = nullptr;
  int64_t items = 0;
};

int fit_count(pgvd_ctx *ctx, const float *src, bool from_ci, int64_t items, float min_d,
              int64_t *sample_offsets, int64_t *box_offsets, int32_t *status, FitWork *w,
              hipStream_t st) {
  void *ns = nullptr, *nb = nullptr, *soff = nullptr, *boff = nullptr, *ws = nullptr, *err = nullptr;
  const size_t m = static_cast<size_t>(std::max<int64_t>(items, 1));
  CHK(pool_get(ctx, "fit_ns", 4 * m, &ns));
  CHK(pool_get(ctx, "fit_nb", 4 * m, &nb));
  if (!sample_offsets) CHK(pool_get(ctx, "fit_soff", 8 * (m + 1), &soff));
  if (!box_offsets) CHK(pool_get(ctx, "fit_boff", 8 * (m + 1), &boff));
  CHK(pool_get(ctx, "fit_ws", 8 * static_cast<size_t>(walk_scan_ws_words(items)), &ws));
  CHK(pool_get(ctx, "fit_err", 8, &err));
  w->ns = static_cast<uint32_t *>(ns);
  w->nb = static_cast<uint32_t *>(nb);
  w->so = sample_offsets ? sample_offsets : static_cast<int64_t *>(soff);
  w->bo = box_offsets ? box_offsets : static_cast<int64_t *>(boff);
  w->err = static_cast<unsigned long long *>(err);
  w->items = items;
  {
    Stage t(ctx, "gvd_fit_count", st);
    launch_fit_count(src, from_ci, items, min_d, w->ns, w->nb, status, w->err, st);
  }
  HIPCHK(hipGetLastError());
  if (items > 0) {
    launch_offsets_scan(w->ns, items, w->so, static_cast<uint64_t *>(ws), st);
    launch_offsets_scan(w->nb, items, w->bo, static_cast<uint64_t *>(ws), st);
  } else {
    const int64_t zero = 0;
    HIPCHK(hipMemcpyAsync(w->so, &zero, 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(w->bo, &zero, 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));      // `zero` is a stack value
  }
  HIPCHK(hipGetLastError());
  return PGVD_OK;
}

int fit_totals(const FitWork &w, int64_t *samples_out, int64_t *boxes_out,
               unsigned long long *first_err, const uint32_t *extra_word, uint32_t *extra_out,
               hipStream_t st) {
  int64_t totals[2] = {0, 0};
  unsigned long long e = ~0ull;
  uint32_t x = 0;
  HIPCHK(hipMemcpyAsync(&totals[0], w.so + w.items, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&totals[1], w.bo + w.items, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&e, w.err, 8, hipMemcpyDeviceToHost, st));
  if (extra_word) HIPCHK(hipMemcpyAsync(&x, extra_word, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *samples_out = totals[0];
  *boxes_out = totals[1];
  *first_err = e;
  if (extra_out) *extra_out = x;
  return PGVD_OK;
}

int fit_write(pgvd_ctx *ctx, const FitWork &w, const float *src, bool from_ci, float min_d,
              int32_t *extra, float *samples, float *origins, float *lengths, hipStream_t st) {
  if (w.items == 0) return PGVD_OK;
  Stage t(ctx, "gvd_fit_write", st);
  launch_fit_write(src, from_ci, w.items, min_d, w.so, w.bo, w.ns, samples, origins, lengths,
                   extra, st);
  HIPCHK(hipGetLastError());
  return PGVD_OK;
}

// count, totals, capacity check, write
int fit_impl(pgvd_ctx *ctx, const float *src, bool from_ci, int64_t items, float min_d,
             int32_t *extra, float *samples, float *origins, float *lengths, int64_t capacity,
             int64_t *sample_offsets, int64_t *box_offsets, int32_t *status,
             int64_t *samples_out, int64_t *boxes_out, unsigned long long *first_err,
             hipStream_t st) {
  FitWork w;
  CHK(fit_count(ctx, src, from_ci, items, min_d, sample_offsets, box_offsets, status, &w, st));
  CHK(fit_totals(w, samples_out, boxes_out, first_err, nullptr, nullptr, st));
  if (*samples_out > capacity || *boxes_out > capacity)
    return fail(PGVD_ERR_CAPACITY, "FitBoxes output exceeds the capacity (*samples_out, *boxes_out)");
  if (*samples_out > 0 || *boxes_out > 0)
    CHK(fit_write(ctx, w, src, from_ci, min_d, extra, samples, origins, lengths, st));
  return PGVD_OK;
}

int fit_error(unsigned long long e, const char *what) {
  if (e == ~0ull) return PGVD_OK;
  char msg[200];
  std::snprintf(msg, sizeof msg, "%s %llu: %s", what, e >> 8, fit_error_text((int)(e & 0xFF)));
  return fail(PGVD_ERR_GEOMETRY, msg);
}

int check_segments_args(int64_t P, const int64_t *counts, int64_t *nv, int64_t *nseg) {
  if (P < 0) return fail(PGVD_ERR_INVALID, "P must be >= 0");
  if (P > 0 && !counts) return fail(PGVD_ERR_INVALID, "null counts");
  int64_t v = 0, sgs = 0;
  for (int64_t j = 0; j < P; ++j) {
    if (counts[j] < 0) return fail(PGVD_ERR_INVALID, "negative vertex count");
    v += counts[j];
    sgs += counts[j] > 0 ? counts[j] - 1 : 0;
  }
  if (v >= (int64_t(1) << 31)) return fail(PGVD_ERR_TOO_LARGE, "2^31 or more vertices");
  *nv = v;
  *nseg = sgs;
  return PGVD_OK;
}

int segments_impl(pgvd_ctx *ctx, const float *verts, const int64_t *counts, int64_t P,
                  const float bb_min[2], float bbw, int bits, float *segs, int32_t *labels,
                  int64_t nv, hipStream_t st) {
  std::vector<int64_t> off(2 * static_cast<size_t>(P) + 1);
  int64_t v = 0, sgs = 0;
  for (int64_t j = 0; j < P; ++j) {
    off[j] = v;
    off[P + 1 + j] = sgs;
    v += counts[j];
    sgs += counts[j] > 0 ? counts[j] - 1 : 0;
  }
  off[P] = v;
  void *doff = nullptr, *bad = nullptr;
  CHK(pool_get(ctx, "seg_offsets", 8 * off.size(), &doff));
  CHK(pool_get(ctx, "seg_bad", 4, &bad));
  HIPCHK(hipMemcpyAsync(doff, off.data(), 8 * off.size(), hipMemcpyHostToDevice, st));
  const uint32_t zero = 0;
  HIPCHK(hipMemcpyAsync(bad, &zero, 4, hipMemcpyHostToDevice, st));
  launch_poly_segments(verts, nv, static_cast<int64_t *>(doff), static_cast<int64_t *>(doff) + P + 1,
                       P, bb_min, bbw, bits, segs, labels, static_cast<uint32_t *>(bad), st);
  HIPCHK(hipGetLastError());
  uint32_t b = 0;
  HIPCHK(hipMemcpyAsync(&b, bad, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));       // also keeps off alive until the copy ran
  if (b) return fail(PGVD_ERR_BAD_COORD, "obj2Oct cannot have coord less than zero");
  return PGVD_OK;
}

}  // namespace

extern "C" {

int pgvd_polyline_segments(pgvd_ctx *ctx, const float *verts, const int64_t *counts, int64_t P,
                           const float bb_min[2], float bbw, int bits, float *segs,
                           int32_t *labels, int64_t *nseg_out, pgvd_stream s) {
  if (!ctx || !bb_min || !nseg_out) return fail(PGVD_ERR_INVALID, "null argument");
  if (bits < 1 || bits > 30) return fail(PGVD_ERR_INVALID, "bits must be in 1..30 (2D)");
  int64_t nv = 0, nseg = 0;
  CHK(check_segments_args(P, counts, &nv, &nseg));
  if (nv > 0 && !verts) return fail(PGVD_ERR_INVALID, "null verts");
  if (nseg > 0 && (!segs || !labels)) return fail(PGVD_ERR_INVALID, "null segs or labels");
  *nseg_out = nseg;
  if (nv == 0) return PGVD_OK;
  CHK(use_device(ctx));
  return segments_impl(ctx, verts, counts, P, bb_min, bbw, bits, segs, labels, nv, pick(ctx, s));
}

int pgvd_multi_cells(pgvd_ctx *ctx, const float *segs, const int32_t *labels, int64_t n, int bits,
                     const int32_t *octree, int64_t octree_size, pgvd_cell_intersections *ci,
                     int32_t *walk_status, pgvd_stream s) {
  if (!ctx || !octree || !ci || (n > 0 && (!segs || !labels)))
    return fail(PGVD_ERR_INVALID, "null argument");
  if (bits < 1 || bits > 30) return fail(PGVD_ERR_INVALID, "bits must be in 1..30 (2D)");
  if (n < 0 || n >= (int64_t(1) << 31)) return fail(PGVD_ERR_INVALID, "n must be in [0, 2^31)");
  if (octree_size < 1 || octree_size >= (int64_t(1) << 29))
    return fail(PGVD_ERR_INVALID, "octree_size must be in [1, 2^29)");
  CHK(use_device(ctx));
  return multi_cells_impl(ctx, segs, labels, n, bits, octree, octree_size,
                          reinterpret_cast<int32_t *>(ci), walk_status, pick(ctx, s));
}

int pgvd_fit_boxes(pgvd_ctx *ctx, const float *pairs, int64_t n, float min_d, float *samples,
                   float *origins, float *lengths, int64_t capacity, int64_t *sample_offsets,
                   int64_t *box_offsets, int64_t *samples_out, int64_t *boxes_out, int32_t *status,
                   pgvd_stream s) {
  if (!ctx || !samples_out || !boxes_out || (n > 0 && !pairs)) return fail(PGVD_ERR_INVALID, "null argument");
  if (n < 0 || n >= (int64_t(1) << 31)) return fail(PGVD_ERR_INVALID, "n must be in [0, 2^31)");
  if (capacity < 0) return fail(PGVD_ERR_INVALID, "capacity must be >= 0");
  if (capacity > 0 && (!samples || !origins || !lengths))
    return fail(PGVD_ERR_INVALID, "null output buffer");
  CHK(use_device(ctx));
  unsigned long long e = ~0ull;
  CHK(fit_impl(ctx, pairs, false, n, min_d, nullptr, samples, origins, lengths, capacity,
               sample_offsets, box_offsets, status, samples_out, boxes_out, &e, pick(ctx, s)));
  HIPCHK(hipStreamSynchronize(pick(ctx, s)));
  return fit_error(e, "pair");
}

int pgvd_find_multi_cells(pgvd_ctx *ctx, const pgvd_cell_intersections *ci, int64_t octree_size,
                          float min_d, int32_t *extra, float *samples, float *origins,
                          float *lengths, int64_t capacity, int64_t *samples_out,
                          int64_t *boxes_out, pgvd_stream s) {
  if (!ctx || !ci || !samples_out || !boxes_out) return fail(PGVD_ERR_INVALID, "null argument");
  if (octree_size < 0 || octree_size >= (int64_t(1) << 29))
    return fail(PGVD_ERR_INVALID, "octree_size must be in [0, 2^29)");
  if (capacity < 0) return fail(PGVD_ERR_INVALID, "capacity must be >= 0");
  CHK(use_device(ctx));
  void *dummy = nullptr;
  // the write pass writes every output; absent ones go to a pooled sink
  if (capacity > 0 && (!samples || !origins || !lengths)) {
    CHK(pool_get(ctx, "fit_sink", 16 * static_cast<size_t>(capacity), &dummy));
    float *sink = static_cast<float *>(dummy);
    if (!samples) samples = sink;
    if (!origins) origins = sink + 2 * capacity;
    if (!lengths) lengths = sink;
  }
  unsigned long long e = ~0ull;
  CHK(fit_impl(ctx, reinterpret_cast<const float *>(ci), true, 4 * octree_size, min_d, extra,
               samples, origins, lengths, capacity, nullptr, nullptr, nullptr, samples_out,
               boxes_out, &e, pick(ctx, s)));
  HIPCHK(hipStreamSynchronize(pick(ctx, s)));
  return fit_error(e, "node*4+octant");
}

int pgvd_gvd_refine(pgvd_ctx *ctx, const float *verts, const int64_t *counts, int64_t P, int bits,
                    const float *custom_bb, int max_iterations, float min_d,
                    int64_t *extra_per_iteration, pgvd_gvd_result *out, pgvd_stream s) {
  if (!ctx || !out) return fail(PGVD_ERR_INVALID, "null argument");
  if (bits < 1 || bits > 30) return fail(PGVD_ERR_INVALID, "bits must be in 1..30 (2D)");
  if (max_iterations < 1) return fail(PGVD_ERR_INVALID, "max_iterations must be >= 1");
  int64_t nv = 0, nseg = 0;
  CHK(check_segments_args(P, counts, &nv, &nseg));
  if (nv == 0) return fail(PGVD_ERR_EMPTY, "no vertices");
  if (!verts) return fail(PGVD_ERR_INVALID, "null verts");
  CHK(use_device(ctx));
  hipStream_t st = pick(ctx, s);
  std::memset(out, 0, sizeof *out);

  // Octree2.cpp:166-172: the box of every vertex, or the caller's; the
  // quantization of Karras::Quantize (Karras.cpp:69-99) takes the vertices'
  // own box (Octree2.cpp:175 passes no customBB).
  float vb[4];
  CHK(pgvd_bbox(ctx, verts, nv, 2, vb, st));
  const float *bb = custom_bb ? custom_bb : vb;
  const float bbw = std::max(bb[2] - bb[0], bb[3] - bb[1]);   // BoundingBox::max_size
  void *q64 = nullptr;
  CHK(pool_get(ctx, "gvd_q64", 16 * static_cast<size_t>(nv), &q64));
  int64_t nq = 0;
  CHK(pgvd_quantize(ctx, verts, nv, 2, bits, nullptr, 0, static_cast<int64_t *>(q64), &nq, st));

  // The points of each build: grown by the extra points, kept in one of two
  // pooled buffers.
  const char *qkey[2] = {"gvd_q_a", "gvd_q_b"};
  int qcur = 0;
  int64_t qcap = std::max<int64_t>(2 * nq, 1024);
  void *q = nullptr;
  CHK(pool_get(ctx, qkey[qcur], 8 * static_cast<size_t>(qcap), &q));
  launch_narrow_i64(static_cast<int64_t *>(q64), 2 * nq, static_cast<int32_t *>(q), st);
  HIPCHK(hipGetLastError());
  if (nq == 1) HIPCHK(hipMemsetAsync(q, 0, 8, st));     // Karras.cpp:84-88: one point {0,0}

  void *segs = nullptr, *labels = nullptr, *wstat = nullptr, *flag = nullptr;
  CHK(pool_get(ctx, "gvd_segs", 16 * static_cast<size_t>(std::max<int64_t>(nseg, 1)), &segs));
  CHK(pool_get(ctx, "gvd_labels", 4 * static_cast<size_t>(std::max<int64_t>(nseg, 1)), &labels));
  CHK(pool_get(ctx, "gvd_wstat", 4 * static_cast<size_t>(std::max<int64_t>(nseg, 1)), &wstat));
  CHK(pool_get(ctx, "gvd_flag", 4, &flag));
  CHK(segments_impl(ctx, verts, counts, P, bb, bbw, bits, static_cast<float *>(segs),
                    static_cast<int32_t *>(labels), nv, st));

  int it = 0;
  int64_t n_extra = 0, n_boxes = 0, O = 0;
  const int32_t *octree = nullptr;
  void *ci = nullptr, *extra = nullptr, *samples = nullptr, *origins = nullptr, *lengths = nullptr;
  int64_t ecap = 0;
  for (;;) {
    // Octree2.cpp:178-181: the extra points join the points
    if (n_extra > 0) {
      if (nq + n_extra > qcap) {
        qcap = 2 * (nq + n_extra);
        void *nqbuf = nullptr;
        CHK(pool_get(ctx, qkey[1 - qcur], 8 * static_cast<size_t>(qcap), &nqbuf));
        HIPCHK(hipMemcpyAsync(nqbuf, q, 8 * static_cast<size_t>(nq), hipMemcpyDeviceToDevice, st));
        q = nqbuf;
        qcur = 1 - qcur;
      }
      HIPCHK(hipMemcpyAsync(static_cast<int32_t *>(q) + 2 * nq, extra,
                            8 * static_cast<size_t>(n_extra), hipMemcpyDeviceToDevice, st));
      nq += n_extra;
    }
    n_extra = 0;
    n_boxes = 0;
    O = 0;
    octree = nullptr;
    if (nq > 1) {
      // :182-183, BuildOctreeInParallel from the integer points
      pgvd_build_opts opts;
      std::memset(&opts, 0, sizeof opts);
      opts.quantized_input = 1;
      CHK(pgvd_build_octree(ctx, q, nq, 2, bits, nullptr, &opts, st));
      pgvd_octree_result r;
      CHK(pgvd_build_octree_finish(ctx, &r));
      O = r.octree_size;
      octree = r.octree;
      // :186, FindMultiCells
      CHK(pool_get(ctx, "gvd_ci", 160 * static_cast<size_t>(O), &ci));
      CHK(multi_cells_impl(ctx, static_cast<float *>(segs), static_cast<int32_t *>(labels), nseg,
                           bits, octree, O, static_cast<int32_t *>(ci),
                           static_cast<int32_t *>(wstat), st));
      launch_any_nonzero(static_cast<int32_t *>(wstat), nseg, static_cast<uint32_t *>(flag), st);
      HIPCHK(hipGetLastError());
      // the sampling: counts, then the points into buffers that hold them
      FitWork w;
      CHK(fit_count(ctx, static_cast<float *>(ci), true, 4 * O, min_d, nullptr, nullptr, nullptr,
                    &w, st));
      unsigned long long e = ~0ull;
      uint32_t walk_bad = 0;
      CHK(fit_totals(w, &n_extra, &n_boxes, &e, static_cast<uint32_t *>(flag), &walk_bad, st));
      if (walk_bad) return fail(PGVD_ERR_GEOMETRY, "a cell walk passed 10000 cells (Infinite loop)");
      CHK(fit_error(e, "node*4+octant"));
      ecap = std::max<int64_t>(std::max(n_extra, n_boxes), 1);
      CHK(pool_get(ctx, "gvd_extra", 8 * static_cast<size_t>(ecap), &extra));
      CHK(pool_get(ctx, "gvd_samples", 8 * static_cast<size_t>(ecap), &samples));
      CHK(pool_get(ctx, "gvd_origins", 8 * static_cast<size_t>(ecap), &origins));
      CHK(pool_get(ctx, "gvd_lengths", 4 * static_cast<size_t>(ecap), &lengths));
      if (n_extra > 0 || n_boxes > 0)
        CHK(fit_write(ctx, w, static_cast<float *>(ci), true, min_d, static_cast<int32_t *>(extra),
                      static_cast<float *>(samples), static_cast<float *>(origins),
                      static_cast<float *>(lengths), st));
    }
    if (extra_per_iteration) extra_per_iteration[it] = n_extra;
    ++it;
    if (!(it < max_iterations && n_extra > 0)) break;     // :188
  }
  HIPCHK(hipStreamSynchronize(st));
  out->iterations = it;
  out->n_qpoints = nq;
  out->octree_size = O;
  out->n_segments = nseg;
  out->n_extra = n_extra;
  out->n_boxes = n_boxes;
  out->qpoints = static_cast<const int32_t *>(q);
  out->octree = octree;
  out->ci = static_cast<const pgvd_cell_intersections *>(ci);
  out->extra = static_cast<const int32_t *>(extra);
  out->samples = static_cast<const float *>(samples);
  out->origins = static_cast<const float *>(origins);
  out->lengths = static_cast<const float *>(lengths);
  return PGVD_OK;
}

}  // extern "C"
