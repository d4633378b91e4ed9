// kernels.hpp -- host-side launchers for the pgvd HIP kernels and the
// device-resident run state they share.  Internal to libpgvd.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace pgvd {

// Error bits collected on the device during a run (RunState::status).
enum : uint32_t {
  ST_NEG_COORD = 1u << 0,    // a coordinate quantized below zero
  ST_CAPACITY = 1u << 1,     // octree larger than the output capacity
  ST_STACK = 1u << 2,        // gather stack overflow (malformed tree)
  ST_TOO_LARGE = 1u << 3,    // a count exceeded the 30-bit look-back field
  ST_DUP_KEY = 1u << 4,      // BRT input had equal neighbouring keys
  ST_WALK = 1u << 5,         // parent walk did not terminate within bound
};

constexpr int kMaxPasses = 16;    // 128-bit keys / 8-bit digits
constexpr int kNumCounters = 24;  // tile counters: sort passes, unique, scans
// Counter slots after the sort passes' kMaxPasses.
enum {
  CTR_UNIQUE = kMaxPasses,
  CTR_SPLITS = kMaxPasses + 1,
  CTR_SCAN = kMaxPasses + 2,
  CTR_LS = kMaxPasses + 3,
  CTR_BRT_DEFER = kMaxPasses + 4,   // BRT nodes left to the global-memory pass
  CTR_OCT_DEFER = kMaxPasses + 5,   // octree gathers left to the global pass
};

// Per-run scalars that kernels read and write on the device.  Cleared by the
// prep kernel at the start of every run.
struct RunState {
  uint32_t bb_min[3];        // fenc-encoded bbox minimum
  uint32_t bb_max[3];        // fenc-encoded bbox maximum
  uint32_t status;           // ST_* bits
  uint32_t malformed;        // 1 if some BRT node took an FP32-quirk split
  uint32_t m;                // unique key count (incl. the padding zero)
  uint32_t octree_size;      // O = ps[m-1]
  uint32_t msd_ok;           // 1: the bucket sort (bucket.hip) sorted the keys
  uint32_t msd_dups;         // 1: the bucket sort saw equal keys (or a zero key to merge)
  uint32_t counters[kNumCounters];
  // bucket sort (bucket.hip): the current level's segments and their keys,
  // the next level's, the LDS sort items so far, the current level's tiles
  uint32_t bk_segs, bk_keys, bk_next_segs, bk_next_keys, bk_items, bk_tiles;
  uint32_t bk_lsd;           // 1: too clustered, the LSD passes sort instead
};

// Fixed-layout BRT node in device memory (the reference's BrtNode,
// C/BrtNode.h:10-21, minus `parent`, which is a separate int32 array).
template <typename K>
struct BrtNodeT {
  typename StoredKey<K>::type lcp;
  int32_t left;
  uint16_t lcp_length;
  uint8_t left_leaf;
  uint8_t right_leaf;
};

// Launch configuration constants.
constexpr int kBlock = 256;

// ---- batched form ----
// One instance of a batch: its points pts[pofs .. pofs+n) (DIM floats or
// ints each) and its key slots [kofs, kofs+n+1) in every per-key array
// (keys, BRT, prefix sums, parent links).  Set once per batch plan.
struct SegInfo {
  uint64_t pofs;       // first point
  uint32_t n;          // points (1 <= n <= kMaxSegPoints)
  uint32_t kofs;       // first key slot
  uint32_t prepend;    // 1 if the zero-key padding applies (nextPow2(n) != n)
  uint32_t pad_;
};
// Per-instance results written on the device.
struct SegState {
  uint32_t bb_min[3];  // fenc-encoded bbox
  uint32_t bb_max[3];
  uint32_t m;          // unique keys (incl. the padding zero)
  uint32_t octree_size;
  uint32_t oct_off;    // first octree node of the instance in the batch output
  uint32_t pad_[3];
};
// Tiles of one stage over the instances: tile t belongs to instance seg[t];
// instance s owns tiles [first[s], first[s+1]).
struct TileTable {
  const uint32_t *seg;
  const uint32_t *first;
};
// Largest instance the batched form takes: ranges stay <= 2^23 keys, where
// the BRT's FP32 split step is exact, so every instance's radix tree is well
// formed.
constexpr int64_t kMaxSegPoints = int64_t(1) << 23;
// Points per bbox / encode chunk.
constexpr int kEncodeChunk = 65536;

// ---- morton.hip ----
void launch_prep(RunState *st, uint32_t *zero_words, size_t nzero, hipStream_t s);
// partials: scratch of kBboxPartBytes (one 8-float record per workgroup).
constexpr int kBboxParts = 2048;
constexpr size_t kBboxPartBytes = (size_t)kBboxParts * 8 * sizeof(float);
// prep: also the build's start (launch_prep's work), in the same launch
// where the points allow.
void launch_bbox(const float *pts, int64_t n, int dim, float *partials, RunState *st,
                 hipStream_t s, uint32_t *zero_words = nullptr, size_t nzero = 0,
                 bool prep = false);
// Quantize + Morton encode (+ optional digit histograms for `passes` passes).
// bb: if custom, (min, max) given; else taken from st (launch_bbox).
// msd_hist (nullable): instead of the digit histograms (passes is taken as
// 0), the histogram of key >> msd_shift over msd_nb buckets (the bucket
// sort's), in the vectorised form only.
template <typename K>
void launch_encode(const float *pts, int64_t n, int dim, int bits, float effw,
                   int custom_bb, const float *bbmin, const float *bbmax,
                   K *keys, int64_t *qout, uint32_t *hist, int passes, RunState *st,
                   hipStream_t s, uint32_t *msd_hist = nullptr, int msd_shift = 0,
                   int msd_nb = 0);
// true when launch_encode takes the vectorised form (and so fills msd_hist)
bool encode_vectorised(const float *pts, const void *keys);
// Morton encode from integer points (int32 or int64 coordinates).
template <typename K>
void launch_points_to_morton(const void *pts, int coord_bytes, int64_t n, int64_t padded_n,
                             int dim, int bits, K *keys, hipStream_t s);
// skip_if_msd (nullable): does nothing when skip_if_msd->msd_ok.  Otherwise
// also zeroes zero_words[0 .. nzero) (16-byte aligned, nzero a multiple of
// 4): the passes' look-back words, which the build's start leaves alone
// beside the bucket sort.
template <typename K>
void launch_histogram(const K *keys, int64_t n, uint32_t *hist, int passes, hipStream_t s,
                      const RunState *skip_if_msd = nullptr, uint32_t *zero_words = nullptr,
                      size_t nzero = 0);
// Morton encode + digit histograms from int32 grid points (no padding).
template <typename K>
void launch_encode_int(const int32_t *pts, int64_t n, int dim, int bits, K *keys,
                       uint32_t *hist, int passes, hipStream_t s);

// ---- stages.hip (the reference's elementary stage kernels) ----
template <typename K>
void launch_bit_predicate(const K *in, uint32_t *pred, int index, uint32_t compared, int64_t n,
                          hipStream_t s);
template <typename K>
void launch_unique_predicate(const K *in, uint32_t *pred, int64_t n, hipStream_t s);
template <typename K>
void launch_single_compact(const K *in, K *out, const uint32_t *pred, const uint32_t *addr,
                           int64_t n, hipStream_t s);
template <typename K>
void launch_double_compact(const K *in, K *out, const uint32_t *pred, const uint32_t *addr,
                           int64_t n, hipStream_t s);
void launch_init_octree(int32_t *octree, int64_t nodes, int dim, hipStream_t s);

// ---- util.hip ----
void launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s);
void launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s);
void launch_set_sizes(RunState *st, const uint32_t *ps, int64_t m, hipStream_t s);
// out4 (device, 4 int32): pgvd_status of the run, m, octree size, malformed.
void launch_build_state(const RunState *st, int32_t *out4, hipStream_t s);
// *out (+)= digest of `words` 32-bit words at p: sum of mix64(w_i | i << 32).
void launch_digest(const void *p, size_t words, uint64_t *out, bool zero_first, hipStream_t s);
// *out = digest of count[0] * per_item words at p (at most n_max), the count
// read on the device.
void launch_digest_count(const void *p, const uint32_t *count, uint32_t per_item, size_t n_max,
                         uint64_t *out, hipStream_t s);

// ---- sort.hip ----
// Onesweep LSD radix sort of n keys (+ optional u32 payload) over `passes`
// 8-bit digits.  Returns 0 if the result is in (keys, vals), 1 if in
// (keys_alt, vals_alt).  hist holds the digit histograms (passes x 256),
// lookback `passes` x tiles x 256 words, cleared beforehand.
// Pass 0 reads payload from vals_in (may equal vals); later passes ping-pong
// vals/vals_alt in step with the keys.
// pass_events (nullable): 2*passes events recorded around each pass.
// skip_if_msd (nullable): the passes do nothing when skip_if_msd->msd_ok.
template <typename K>
int launch_onesweep(K *keys, K *keys_alt, const uint32_t *vals_in, uint32_t *vals,
                    uint32_t *vals_alt, int64_t n, int passes, const uint32_t *hist,
                    uint32_t *lookback, RunState *st, hipStream_t s,
                    hipEvent_t *pass_events = nullptr, const RunState *skip_if_msd = nullptr);

// ---- query.hip ----
// OctreeUtils::FindLeaf for q int32 grid points (see pgvd_find_leaf).  With
// D > 0 the launch first fills `table` (8 << (D * dim) bytes) with the
// octree's cells at depth D; find_leaf_table_depth picks D (0: no table).
#ifndef PGVD_FINDLEAF_TABLE_BITS
#define PGVD_FINDLEAF_TABLE_BITS 21
#endif
constexpr int kFindLeafTableBits = PGVD_FINDLEAF_TABLE_BITS;
int find_leaf_table_depth(int64_t q, int dim, int bits);
void launch_find_leaf(const int32_t *pts, int64_t q, int dim, int bits, const int32_t *octree,
                      int64_t octree_size, void *table, int D, int32_t *parent, int32_t *octant,
                      int32_t *data, int32_t *level, uint32_t *status, hipStream_t s);

// ---- walk.hip ----
// CellWalk (viewer/Octree2.cpp:265-313) for n 2D segments: the count pass
// writes counts (n), status (nullable) and offsets (n + 1, exclusive scan);
// ws holds walk_scan_ws_words(n) words.  The write pass puts each segment's
// cells (5 int32 each) at cells + 5 * offsets[i].  With scratch (n << log2k
// cells), the count pass also keeps each segment's first 2^log2k cells
// there, and the write pass moves them and walks again only the segments
// that had more.
int64_t walk_scan_ws_words(int64_t n);
void launch_cell_walk_count(const float *segs, int64_t n, int bits, const int32_t *octree,
                            int64_t size, uint32_t *counts, int32_t *status, int64_t *offsets,
                            uint64_t *ws, int32_t *scratch, int log2k, hipStream_t s);
void launch_cell_walk_write(const float *segs, int64_t n, int bits, const int32_t *octree,
                            int64_t size, const int64_t *offsets, int32_t *cells,
                            const int32_t *scratch, uint32_t *counts, int log2k, hipStream_t s);
// offsets[0] = 0, offsets[i + 1] = counts[0] + ... + counts[i] (n + 1
// entries); ws holds walk_scan_ws_words(n) words.
void launch_offsets_scan(const uint32_t *counts, int64_t n, int64_t *offsets, uint64_t *ws,
                         hipStream_t s);

// ---- gvd.hip ----
// obj2Oct of every vertex and the segments of every polygon (vertices
// [voff[j], voff[j+1]), segments from soff[j]); bad: 1 if a coordinate fell
// below zero.
void launch_poly_segments(const float *verts, int64_t nv, const int64_t *voff, const int64_t *soff,
                          int64_t P, const float bbmin[2], float bbw, int bits, float *segs,
                          int32_t *labels, uint32_t *bad, hipStream_t s);
// MCCallback over a finished walk (offsets, cells; `total` cells) into the
// CellIntersections records ci (40 words per node).  parts/vlabel: total
// entries; first/second: 4*O; best: 8*O.
void launch_multi_cells(const float *segs, const int32_t *labels, int64_t n,
                        const int64_t *offsets, const int32_t *cells, int64_t total, int64_t O,
                        float4 *parts, int32_t *vlabel, uint32_t *first, uint32_t *second,
                        unsigned long long *best, int32_t *ci, hipStream_t s);
// FitBoxes over items: pairs (8 floats each) or CellIntersections slots
// (from_ci, items = 4*O).  Count pass: ns/nb per item, status (nullable),
// *err = min(item << 8 | code) over failing items (~0 if none).
void launch_fit_count(const float *src, bool from_ci, int64_t items, float min_d, uint32_t *ns,
                      uint32_t *nb, int32_t *status, unsigned long long *err, hipStream_t s);
void launch_fit_write(const float *src, bool from_ci, int64_t items, float min_d,
                      const int64_t *soff, const int64_t *boff, const uint32_t *ns, float *samples,
                      float *origins, float *lengths, int32_t *extra, hipStream_t s);
// out[i] = (int32_t)in[i]
void launch_narrow_i64(const int64_t *in, int64_t n, int32_t *out, hipStream_t s);
// *flag = 1 if some v[i] != 0, else 0
void launch_any_nonzero(const int32_t *v, int64_t n, uint32_t *flag, hipStream_t s);

// ---- textio.hip ----
// The octree's text form (operator<< / operator>>, C/OctNode.h:58-98).  ws
// holds text_ws_words(tiles) words: word 1 receives the total (bytes
// formatted, or tokens seen), word 2 the first malformed record token + 1
// (all ones: none).
int64_t text_ws_words(int64_t tiles);
int64_t format_text_tiles(int64_t O);
int64_t parse_text_tiles(int64_t len);
void launch_format_text(const int32_t *octree, int64_t O, int dim, int hdr_len, uint8_t *text,
                        int64_t capacity, uint64_t *ws, hipStream_t s);
void launch_parse_text(const uint8_t *text, int64_t len, int dim, int64_t N, int32_t *octree,
                       uint64_t *ws, hipStream_t s);

// ---- bucket.hip ----
// The whole path's sort for 32/64/128-bit keys with optional u32 labels:
// level 0 splits by the top b bits (b <= kBkMaxBits) and sorts every bucket
// of at most bucket_cap<K>() keys in LDS; larger buckets go through levels
// that split them below their highest varying bit.  The result (keys, and
// labels in the same order) is in (kA, vA); not stable among equal keys.
constexpr int kBkMaxBits = 14;
constexpr int kBkMaxBuckets = 1 << kBkMaxBits;
constexpr int kBkCursor1 = kBkMaxBuckets;                       // first scatter's cursors
constexpr int kBkTileStart = kBkMaxBuckets + 128 * 32;          // second scatter's tile starts
                                                                // (after 128 cursors 32 words apart)
constexpr int kBkCursorWords = kBkTileStart + 128 + 1;
// A segment left to a level: keys [start, start + count) of one buffer,
// all sharing the bits above their highest varying bit (OR ^ AND).
struct BkSeg {
  uint32_t start, count;
  uint32_t tile_base;   // first tile of the level's scatter
  uint32_t shift;       // digit = (key >> shift) & (2^nbits - 1)
  uint32_t nbits;       // 0: every key equal (copied)
  uint32_t pad_[3];
  unsigned long long or_lo, or_hi, and_lo, and_hi;
};
// An LDS sort item: keys [start, start + count) of buffer `parity` (0: A,
// 1: B), all sharing their bits from `sbits` up (sbits 0: all equal, copied).
struct BkItem {
  uint32_t start, count, parity, sbits;
};
struct BucketMem {
  uint32_t *start = nullptr, *cursor = nullptr;
  BkSeg *segs[2] = {nullptr, nullptr};
  uint32_t *hrows[2] = {nullptr, nullptr};
  uint32_t *lcursor = nullptr;
  BkItem *items = nullptr;
  uint32_t max_segs = 0, max_items = 0;
};
template <typename K>
struct BucketArgs {
  K *kA, *kB;                 // keys in kA (encode's output); kB the other buffer
  uint32_t *vA, *vB;          // label buffers (with labels)
  const uint32_t *labels;     // the keys' labels in input order (nullable)
  const uint32_t *hist;       // histogram of the top b bits (2^b counters)
  int64_t n;
  int mbits, b, prepend;
  BucketMem mem;
};
inline constexpr size_t align256_dev(size_t x) { return (x + 255) & ~(size_t)255; }
template <typename K> int bucket_cap();
template <typename K> int bucket_bits_for(int64_t n);
template <typename K> size_t bucket_ws_bytes(int64_t n);
template <typename K> void bucket_carve(void *ws, int64_t n, BucketMem &m);
// Enqueues the sort on s.  Waits on the host (ev, st_host: pinned) for the
// level-0 plan and each level's scan while the device works on.  0: sorted;
// 3: keys of at most 64 bits too clustered (a quarter of them in buckets
// above CAP): nothing moved, the caller sorts by the LSD passes; else a
// launch failed (1) or a segment outlived the levels (2).
template <typename K>
int launch_bucket_sort(const BucketArgs<K> &a, RunState *st, RunState *st_host, hipEvent_t ev,
                       hipStream_t s);

// Bucket histogram bins the vectorised encode kernel holds in LDS beside
// eight workgroups per CU (it shares them with the LSD digit histograms:
// kMaxPasses x 256); up to kBkMaxBits it takes a larger LDS histogram.
constexpr int kEncodeMaxBucketBits = 12;
// Words between the first scatter's cursors (one per L2 line: every scatter
// tile claims on these few words).
constexpr int kBkCStride = 32;
int64_t onesweep_tiles(int key_bytes, bool payload, int64_t n);
int onesweep_tile_size(int key_bytes, bool payload);

// ---- scan.hip ----
constexpr int kScanTile = 4096;
inline int64_t scan_tiles(int64_t n) { return (n + kScanTile - 1) / kScanTile; }
// Unique over [zero?] + keys[0..n): writes out[0..m) and st->m.  With labels,
// sets bit r of conflict_bits for every run r whose labels are not all equal.
// skip_if_msd (nullable): when skip_if_msd->msd_ok and not ->msd_dups (the
// bucket sort found every key distinct and none to merge with the prepended
// zero), the keys are already in place: only out[0] = 0 (prepend_zero) and
// st->m = prepend_zero + n are written.
template <typename K>
void launch_unique(const K *keys, const uint32_t *labels, int64_t n, int prepend_zero, K *out,
                   uint32_t *conflict_bits, uint64_t *lookback, RunState *st, hipStream_t s,
                   const RunState *skip_if_msd = nullptr);
// Local splits (from the BRT) scanned inclusively into ps[0..m); rewrites
// parent[] from claim form (claimer+1) to the reference form; writes ls[]
// when non-null; st->octree_size = ps[m-1].
// only_if_malformed: the kernel returns at once unless st->malformed.
template <typename K>
void launch_splits_scan(const BrtNodeT<K> *brt, int32_t *parent, uint32_t *ls, uint32_t *ps,
                        int64_t max_m, int dim, int only_if_malformed, uint64_t *lookback,
                        RunState *st, hipStream_t s);
// Plain inclusive scan of n u32 values (StreamScan_p).
void launch_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, uint64_t *lookback,
                     RunState *st, hipStream_t s);

// ---- brt.hip ----
// Karras BRT over keys[0..m), m = st->m (or m_host >= 0).  parent[] must be
// zero on entry; receives claimer+1 per claimed node (max over claimers).
template <typename K>
void launch_brt(const K *keys, int64_t max_m, int64_t m_host, int mbits, BrtNodeT<K> *brt,
                int32_t *parent, RunState *st, hipStream_t s);
// Rewrite parent[] from claim form to the reference form (BRT stage API).
void launch_finalize_parents(int32_t *parent, int64_t max_m, int64_t m_host, RunState *st,
                             hipStream_t s);
// Pipeline: every BRT node with its local split and attach octant packed as
// lsx[c] = ls | octant << 8 (lsx[0..m)), parent links as plain stores
// (well-formed tree; see brt.hip), and chunk_tot[k] = the local splits of
// nodes [256k, 256k+256) summed.  defer[] holds the miss words of the
// overflow pass: (max_m + 63) / 64 64-bit words, 8-byte aligned.
template <typename K>
void launch_brt_ls(const K *keys, int64_t max_m, int mbits, int dim, BrtNodeT<K> *brt,
                   int32_t *parent, uint32_t *lsx, uint32_t *chunk_tot, uint32_t *defer,
                   RunState *st, hipStream_t s);
// chunk_tot[0..ceil(m/256)) -> exclusive prefix sums in place;
// st->octree_size = total (the octree size).
void launch_chunk_scan(uint32_t *chunk_tot, int64_t max_m, uint64_t *lookback, RunState *st,
                       hipStream_t s);
// When st->malformed: parent[] <- claim form (zero, then atomicMax claims).
template <typename K>
void launch_claims_if_malformed(const BrtNodeT<K> *brt, int32_t *parent, int64_t max_m,
                                RunState *st, hipStream_t s);

// ---- octree.hip ----
// Test hook of the octree window pass (octree.hip): its per-wave queues hold
// at most cap entries (0: their full size).  Returns 0 on success.
int set_gather_queue_cap(uint32_t cap);

// Octree fill.  parent[] in the reference form.  force_scatter 1 selects the
// order-exact scatter path unconditionally (stage API, where the BRT comes
// from the caller; lsx and defer may then be null); 0 lets st->malformed
// select it on the device; 2 launches the gather path only (the caller
// runs the order-exact path itself when the tree turns out malformed).
// The gather path reads lsx from
// launch_brt_ls and uses defer[] for its miss words ((max_m + 63) / 64
// 64-bit words, 8-byte aligned).
// The gather path also writes ps[0..m) from chunk_excl (launch_chunk_scan)
// and lsx.  gather_ev (nullable): two events recorded around the gather
// kernel's launch.
template <typename K>
void launch_octree(const BrtNodeT<K> *brt, const int32_t *parent, uint32_t *ps,
                   const uint32_t *lsx, const uint32_t *chunk_excl, const K *keys,
                   const uint32_t *conflict_bits, int64_t max_m, int dim, int mbits,
                   int32_t *octree, uint32_t *node_conflict, int64_t capacity,
                   int force_scatter, uint32_t *defer, RunState *st, hipStream_t s,
                   hipEvent_t *gather_ev = nullptr);
// Local splits the reference's way (each node writes its internal children's
// entries, highest claimer last): ComputeLocalSplits_p for a caller's BRT.
template <typename K>
void launch_local_splits(const BrtNodeT<K> *brt, int64_t m, int dim, uint32_t *ls,
                         uint64_t *tmp, hipStream_t s);

// ---- batched stages: one launch per stage for every instance of a batch ----
// Device-side description of a batch plan.  Tile tables:
//   chunk: kEncodeChunk points      sort: sort_ipt*256 keys
//   scan:  kScanTile key slots      brt:  brt_rows*256 key slots
//   node:  256 key slots
// (key-slot tables cover n+1 slots per instance).
struct BatchDev {
  const SegInfo *seg = nullptr;
  const uint32_t *kofs = nullptr;   // B+1 entries: first key slot per instance, then the total
  SegState *ss = nullptr;
  int64_t B = 0;
  int sort_ipt = 16, brt_rows = 4;
  int digit_bits = 8;   // radix digit width of the sort passes (6, 7 or 8)
  int sort_waves = 3;   // waves per SIMD the sort kernel is compiled for (2, 3 or 4)
  TileTable chunk{}, sort{}, scan{}, brt{}, node{};
  int64_t n_chunk = 0, n_sort = 0, n_scan = 0, n_brt = 0, n_node = 0;
};
// Sort tile configurations the batched sort is built for (keys per thread),
// radix digit widths (6, 7, 8 bits; 128-bit keys: 8, so that at most 16
// passes run) and waves per SIMD (2, 3, 4).
bool onesweep_seg_ipt_ok(int key_bytes, int ipt);
bool onesweep_seg_digit_ok(int key_bytes, int digit_bits);
bool onesweep_seg_waves_ok(int waves);
// BRT window configurations (nodes per thread).
bool brt_seg_rows_ok(int key_bytes, int rows);

void launch_seg_init(SegState *ss, int64_t B, hipStream_t s);                   // morton.hip
void launch_bbox_seg(const float *pts, int dim, const BatchDev &b, hipStream_t s);
// Quantize (float points, per-instance bbox) or take int32 grid points, Morton
// encode into keys[kofs..), per-instance digit histograms bhist[s][pass][256].
template <typename K>
void launch_encode_seg(const void *pts, int quantized, int dim, int bits, float effw,
                       const BatchDev &b, K *keys, uint32_t *bhist, int passes, RunState *st,
                       hipStream_t s);
// Per-instance onesweep sort; returns 1 if the result is in keys_alt.
template <typename K>
int launch_onesweep_seg(K *keys, K *keys_alt, const BatchDev &b, int passes,
                        const uint32_t *bhist, uint32_t *lookback, RunState *st, hipStream_t s);
// Per-instance unique (with its zero-key padding) into out[kofs..); ss[s].m.
template <typename K>
void launch_unique_seg(const K *keys, K *out, const BatchDev &b, uint64_t *lookback,
                       RunState *st, hipStream_t s);
// Per-instance BRT + packed local splits (as launch_brt_ls); defer holds
// up to one key-slot index per key slot.
template <typename K>
void launch_brt_seg(const K *keys, int mbits, int dim, const BatchDev &b, BrtNodeT<K> *brt,
                    int32_t *parent, uint32_t *lsx, uint32_t *defer, RunState *st,
                    hipStream_t s);
// Per-instance inclusive scan of the local splits; ss[s].octree_size.
void launch_ls_scan_seg(const uint32_t *lsx, uint32_t *ps, const BatchDev &b,
                        uint64_t *lookback, RunState *st, hipStream_t s);
// ss[s].oct_off = exclusive scan of the octree sizes; st->octree_size = total.
void launch_oct_offsets(const BatchDev &b, RunState *st, hipStream_t s);
// Per-instance octree gather into octree[oct_off[s]..) (instance-local child
// indices); flags ST_CAPACITY if the total exceeds capacity.
template <typename K>
void launch_gather_seg(const BrtNodeT<K> *brt, const uint32_t *ps, const uint32_t *lsx, int dim,
                       const BatchDev &b, int32_t *octree, int64_t capacity, uint32_t *defer,
                       RunState *st, hipStream_t s);

}  // namespace pgvd
