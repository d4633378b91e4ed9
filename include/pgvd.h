/*
 * pgvd.h -- C ABI of libpgvd.so, the MI355X (gfx950) octree-construction
 * path of edwardsjohnmartin/pgvd.
 *
 * Every compute entry point replaces one function of the reference's
 * OpenCL driver layer (namespace Kernels, opencl/Kernels.h:26-50) or of its
 * facade (Karras.h:23-36); the comment above each names it.  The reference
 * passes cl::Buffer handles and returns an OR-accumulated cl_int; this ABI
 * takes plain device pointers and sizes and returns a pgvd_status code
 * (0 = success) with a message in pgvd_last_error().  Stream arguments are
 * hipStream_t values passed as void*; NULL means the context's own stream.
 *
 * Keys are native words chosen by mbits (= bits * dim): uint32_t for
 * mbits <= 32, uint64_t for <= 64, and 16-byte (lo, hi) pairs for <= 128.
 * They carry the same numeric value as the reference's BigUnsigned
 * (C/BigUnsigned.h:8-16); key_bytes arguments are 4, 8 or 16.
 *
 * Sizes: n < 2^30 points.  dim is 2 or 3; bits * dim <= 128 (3D: bits <= 42).
 */
#ifndef PGVD_H
#define PGVD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pgvd_ctx pgvd_ctx;
typedef void *pgvd_stream;

typedef enum pgvd_status {
  PGVD_OK = 0,
  PGVD_ERR_INVALID = 1,     /* bad argument: null pointer, dim, bits, key_bytes */
  PGVD_ERR_EMPTY = 2,       /* zero points (opencl/Kernels.cpp:389-391 throws) */
  PGVD_ERR_BAD_COORD = 3,   /* coordinate quantized below zero (Karras.cpp:56-63 throws) */
  PGVD_ERR_NO_DEVICE = 4,   /* no HIP device / device index out of range */
  PGVD_ERR_ALLOC = 5,       /* device or pinned allocation failed */
  PGVD_ERR_HIP = 6,         /* a HIP runtime call or kernel launch failed */
  PGVD_ERR_CAPACITY = 7,    /* caller's output buffer too small; see *required */
  PGVD_ERR_TOO_LARGE = 8,   /* n >= 2^30 */
  PGVD_ERR_DUP_KEY = 9,     /* BRT input has equal neighbouring keys */
  PGVD_ERR_INTERNAL = 10,
  PGVD_NOT_READY = 11,      /* pgvd_server_poll: the request is not complete yet */
  PGVD_ERR_GEOMETRY = 12    /* the GVD refinement met a case where the reference
                               throws: FitBoxes' logic_errors (opencl/Geom.cpp:
                               492-505) or a cell walk past 10000 cells
                               (viewer/Octree2.cpp:290) */
} pgvd_status;

/* Message for the last failing call on this thread. */
const char *pgvd_last_error(void);
const char *pgvd_version(void);

/* ---- context, memory, streams (the role of CLFW, CLFW/src/clfw.cpp:82-265) ----
 * A context owns one HIP stream (default priority) and a keyed pool of
 * device buffers reused across calls (CLFW::get's keyed reuse,
 * CLFW/src/clfw.cpp:212-226). */
int pgvd_ctx_create(int device, pgvd_ctx **out);
int pgvd_ctx_destroy(pgvd_ctx *ctx);
int pgvd_ctx_stream(pgvd_ctx *ctx, pgvd_stream *out);
/* Most device bytes the context's pool may hold (0, the default: no limit).
 * A call whose pool buffers would pass it fails with PGVD_ERR_ALLOC before
 * allocating, as when the device is out of memory; the buffers already
 * held stay, and the context stays usable.  Lowering the limit frees
 * nothing.  No counterpart in the reference (for callers that share a
 * device between contexts). */
int pgvd_ctx_set_pool_limit(pgvd_ctx *ctx, size_t bytes);
/* Device bytes the context's pool holds now. */
int pgvd_ctx_pool_bytes(pgvd_ctx *ctx, size_t *out);
/* Test hook, no counterpart in the reference: the next launch check of the
 * named stage on this thread ("encode", "sort", "unique", "brt",
 * "splits_scan", "octree", ...; the stage names of the message a failing
 * call leaves) reports PGVD_ERR_HIP as if its launch had failed, and the
 * call returns at once, with work already enqueued left to run.  NULL or ""
 * clears it.  For tests of the error paths. */
int pgvd_debug_fail_launch(const char *stage);
/* Test hook, no counterpart in the reference: the octree gather's per-wave
 * queues of breadth-first work hold at most `entries` entries (0 restores
 * their full size), so that tests can force the path where a full queue
 * sends its node to the global-memory pass.  Applies to every later build
 * on the current device.  Results do not change, only where they are
 * computed. */
int pgvd_debug_gather_queue(int entries);
int pgvd_synchronize(pgvd_ctx *ctx, pgvd_stream s);
int pgvd_alloc(pgvd_ctx *ctx, size_t bytes, void **dptr);
int pgvd_free(pgvd_ctx *ctx, void *dptr);
int pgvd_memcpy_htod(pgvd_ctx *ctx, void *dst, const void *src, size_t bytes, pgvd_stream s);
int pgvd_memcpy_dtoh(pgvd_ctx *ctx, void *dst, const void *src, size_t bytes, pgvd_stream s);
/* Device-to-device copy through a plain copy kernel (also the calibration
 * kernel for HBM byte counters). */
int pgvd_copy_kernel(pgvd_ctx *ctx, void *dst, const void *src, size_t bytes, pgvd_stream s);
/* Position-dependent 64-bit digest of a device range (bytes a multiple of 4,
 * 4-byte aligned): sum over its 32-bit words w_i of splitmix64(w_i | i<<32).
 * Waits for the result.  Compares outputs on the device without a download;
 * no counterpart in the reference. */
int pgvd_digest(pgvd_ctx *ctx, const void *p, size_t bytes, pgvd_stream s, uint64_t *out);

/* Kernel timing with HIP events on the launch stream.  on = 1 records an
 * event pair around every stage ("prep", "bbox", "encode", "msd" (the
 * bucket sort), each "sort_pass", "unique", "brt", "splits_scan",
 * "malformed", "octree"); on = 2 records one pair around all radix-sort
 * passes of a build ("sort") and nothing else; on = 3 one pair around the
 * octree gather kernel of a build ("gather") and nothing else (each event
 * adds ~10 us of queue time).
 * pgvd_timing_read returns the summed milliseconds and record count of one
 * stage since the last reset; it synchronizes. */
int pgvd_timing_enable(pgvd_ctx *ctx, int on);
int pgvd_timing_reset(pgvd_ctx *ctx);
int pgvd_timing_read(pgvd_ctx *ctx, const char *stage, double *total_ms, int64_t *count);

/* ---- stage entry points (device pointers) ---- */

/* Kernels::nextPow2 (opencl/Kernels.cpp:20). */
int pgvd_next_pow2(int n);

/* Bounding box of n points (BoundingBox.h:166-175 over the points, as in
 * Karras.cpp:75-83).  bb_out (host, 2*dim floats): min then max.  Returns
 * after the result is on the host. */
int pgvd_bbox(pgvd_ctx *ctx, const float *pts, int64_t n, int dim, float *bb_out, pgvd_stream s);

/* Karras::Quantize (Karras.cpp:69-99): FP32 quantize to a 2^bits grid.
 * custom_bb (host, 2*dim floats or NULL) plays customBB; clamped as the
 * reference.  qpts_out: device, n*dim int64.  *n_out (host) receives n, or
 * 1 when the box has zero size (Karras.cpp:84-88).  Synchronizes.
 * PGVD_ERR_BAD_COORD where the reference throws. */
int pgvd_quantize(pgvd_ctx *ctx, const float *pts, int64_t n, int dim, int bits,
                  const float *custom_bb, int clamped, int64_t *qpts_out, int64_t *n_out,
                  pgvd_stream s);

/* Kernels::PointsToMorton_p (opencl/Kernels.cpp:47-60, kernel
 * kernels.cl:2-22): keys_out (device) receives next_pow2(n) keys, the
 * Morton keys of the n points followed by zero keys.  pts: device, n*dim
 * integers of coord_bytes (4 or 8). */
int pgvd_points_to_morton(pgvd_ctx *ctx, const void *pts, int coord_bytes, int64_t n, int dim,
                          int bits, void *keys_out, int key_bytes, pgvd_stream s);

/* OctreeUtils::FindLeaf (OctreeUtils.cpp:18-71), batched: for each of q
 * grid points (pts: device, q*dim int32; a coordinate's bits from `bits` up
 * are ignored, as xyz2z ignores them) descend the octree (device,
 * octree_size nodes of 2^dim children + leaf mask, as pgvd_build_octree
 * writes it) from node 0 by the point's Morton digits, the octant masked to
 * dim bits (the reference leaves that mask commented out, :37-41), until the
 * octant's leaf bit is set.  Per query (device arrays of q):
 *   parent  the node holding the leaf octant (the OctCell's parent_idx);
 *   octant  that octant;
 *   data    children[octant] of that node (-1: an empty cell);
 *   level   the digits descended: the cell's width is 2^(bits - level), its
 *           origin the point with its low bits - level bits cleared;
 *   status  (nullable) 0; 1 when no leaf was reached (the reference
 *           throws); 2 when a child index leaves the octree (not one the
 *           build wrote).
 * 1 <= bits <= 31.  A batch of at least 2^22 queries in 3D or 2^21 in 2D
 * (fewer when bits is small) first tabulates the octree's cells at depth 7
 * (3D) or 10 (2D) in a 16 MB (8 MB) buffer the context owns: such calls on
 * one context on different streams must not overlap. */
int pgvd_find_leaf(pgvd_ctx *ctx, const int32_t *pts, int64_t q, int dim, int bits,
                   const int32_t *octree, int64_t octree_size, int32_t *parent, int32_t *octant,
                   int32_t *data, int32_t *level, uint32_t *status, pgvd_stream s);

/* CellWalk (viewer/Octree2.cpp:265-313), batched: for each of n 2D
 * segments (segs: device, ax, ay, bx, by as float in grid units, as
 * obj2Oct gives them) the leaf cells of a 2D octree the walk visits, in
 * order.  A step finds the leaf of the current point (truncated to int, the
 * octant masked to two bits as in pgvd_find_leaf), intersects the segment
 * with the cell's edges (OctreeUtils::FindIntersections,
 * OctreeUtils.cpp:80-172, the same float operations and epsilons) and moves
 * to the farthest intersection.  Outputs (device):
 *   offsets  n + 1 int64: segment i's cells are [offsets[i], offsets[i+1]);
 *   cells    5 int32 per cell: origin x, origin y, width, parent node,
 *            octant; NULL to count only;
 *   status   (nullable) n int32: 0, or 1 when a walk passed 10000 cells
 *            (the reference throws; the first 10000 are kept).
 * *total_out (host) receives the number of cells; PGVD_ERR_CAPACITY when
 * it exceeds capacity (offsets and status are then valid).  With a cell
 * buffer, one walk counts and keeps each segment's first 32 cells in a
 * scratch buffer the context owns (up to 2 GB; past that, a count walk and
 * a write walk), and only segments with more cells walk again.  Returns
 * after the count is read back; the rest is asynchronous on the stream.
 * 1 <= bits <= 30.  Workspace buffers belong to the context: calls on one
 * context must not overlap on different streams. */
int pgvd_cell_walk(pgvd_ctx *ctx, const float *segs, int64_t n, int bits, const int32_t *octree,
                   int64_t octree_size, int64_t *offsets, int32_t *cells, int64_t capacity,
                   int64_t *total_out, int32_t *status, pgvd_stream s);

/* The octree's text form: operator<< for std::vector<OctNode>
 * (C/OctNode.h:58-87, commented out in the reference), formatted on the
 * device.  The text is the node count and a space, then for every node each
 * of its 2^dim children and its leaf mask in decimal, each followed by one
 * space.  octree: device, O records; text: device, capacity bytes, or NULL
 * to measure.  *len_out (host) receives the text's length;
 * PGVD_ERR_CAPACITY if it exceeds capacity (the buffer's contents are then
 * unspecified).  Synchronous: returns when *len_out is known. */
int pgvd_octree_format_text(pgvd_ctx *ctx, const int32_t *octree, int64_t O, int dim, char *text,
                            int64_t capacity, int64_t *len_out, pgvd_stream s);

/* operator>> for std::vector<OctNode> (C/OctNode.h:69-98), parsed on the
 * device: the node count N, then N records of 2^dim + 1 integers separated
 * by any whitespace, the leaf mask cast to unsigned char as the reference
 * reads it; tokens after the last record are ignored.  text: device, len
 * bytes; octree: device, room for cap records.  *O_out (host) receives N;
 * PGVD_ERR_CAPACITY if N > cap.  A token is an optional sign and digits
 * (leading zeros skipped, as operator>> skips them) whose value is within
 * int32.  PGVD_ERR_INVALID when the count is not in the first 4096 bytes or
 * is negative ("-0" is zero), a record token is not such a number (where
 * operator>> sets failbit), or the text ends early.  One difference: a
 * token must end at whitespace, so a last record token with other bytes
 * after its digits ("15x") is rejected where operator>> reads its numeric
 * prefix and stops.  Synchronous. */
int pgvd_octree_parse_text(pgvd_ctx *ctx, const char *text, int64_t len, int dim, int32_t *octree,
                           int64_t cap, int64_t *O_out, pgvd_stream s);

/* Octree files, replacing the reference's (commented-out) stream operators
 * on files.  PGVD_OCTREE_TEXT is the text form above; PGVD_OCTREE_BINARY is
 * a 32-byte header ("PGVDOCT1", uint32 dim, uint32 record bytes, int64 node
 * count, 8 zero bytes) followed by the records as the device holds them
 * (little-endian int32).  Load tells the two apart by the header.  Both run
 * on the context's stream and return when done; the octree pointer is
 * device memory. */
enum { PGVD_OCTREE_TEXT = 0, PGVD_OCTREE_BINARY = 1 };
int pgvd_octree_save(pgvd_ctx *ctx, const int32_t *octree, int64_t O, int dim, const char *path,
                     int format);
int pgvd_octree_load(pgvd_ctx *ctx, const char *path, int dim, int32_t *octree, int64_t cap,
                     int64_t *O_out);

/* Kernels::RadixSortBigUnsigned (opencl/Kernels.cpp:219-247): ascending
 * stable sort of n keys in place, over the low mbits bits; vals (nullable,
 * u32 per key) is permuted with the keys. */
int pgvd_radix_sort(pgvd_ctx *ctx, void *keys, uint32_t *vals, int64_t n, int key_bytes,
                    int mbits, pgvd_stream s);

/* Kernels::UniqueSorted (opencl/Kernels.cpp:198-217): drop adjacent equal
 * keys of a sorted array.  keys_out (device) has room for n keys; *m_out
 * (host) receives the count.  Synchronizes (the reference does a blocking
 * read of the count, :214). */
int pgvd_unique_sorted(pgvd_ctx *ctx, const void *keys_in, void *keys_out, int64_t n,
                       int key_bytes, int64_t *m_out, pgvd_stream s);

/* Kernels::StreamScan_p / StreamScan_s (opencl/Kernels.cpp:102-159):
 * inclusive scan of n u32 values, modulo 2^32. */
int pgvd_stream_scan(pgvd_ctx *ctx, const uint32_t *in, uint32_t *out, int64_t n, pgvd_stream s);

/* The elementary steps of the reference's bit-serial sort and compaction
 * (the whole-path entry does not use them):
 *   Kernels::BitPredicate (opencl/Kernels.cpp:77-89, C/ParallelAlgorithms.c:13-18):
 *     pred[i] = (bit `index` of key i == compared)
 *   Kernels::UniquePredicate (opencl/Kernels.cpp:91-100, C/ParallelAlgorithms.c:22-31):
 *     pred[0] = 1, pred[i] = (key i != key i-1)
 *   Kernels::SingleCompact (opencl/Kernels.cpp:161-172, C/ParallelAlgorithms.c:73-81):
 *     out[addr[i]-1] = in[i] where pred[i] == 1
 *   Kernels::DoubleCompact (opencl/Kernels.cpp:174-196, C/ParallelAlgorithms.c:54-71):
 *     keys with pred 1 to addr[i]-1, the rest after them in order
 *   Kernels::InitOctree (opencl/Kernels.cpp:315-331, C/BuildOctree.c:134-139):
 *     children = -1, leaf = all octant bits, for `nodes` nodes
 * addr is the inclusive scan of pred (pgvd_stream_scan). */
int pgvd_bit_predicate(pgvd_ctx *ctx, const void *keys, int key_bytes, uint32_t *pred, int index,
                       int compared, int64_t n, pgvd_stream s);
int pgvd_unique_predicate(pgvd_ctx *ctx, const void *keys, int key_bytes, uint32_t *pred,
                          int64_t n, pgvd_stream s);
int pgvd_single_compact(pgvd_ctx *ctx, const void *in, void *out, int key_bytes,
                        const uint32_t *pred, const uint32_t *addr, int64_t n, pgvd_stream s);
int pgvd_double_compact(pgvd_ctx *ctx, const void *in, void *out, int key_bytes,
                        const uint32_t *pred, const uint32_t *addr, int64_t n, pgvd_stream s);
int pgvd_init_octree(pgvd_ctx *ctx, void *octree, int64_t nodes, int dim, pgvd_stream s);

/* BRT node records (C/BrtNode.h:10-21 without `parent`, which is a separate
 * int32 array with the reference's values: -1 for the root, 0 for a node no
 * other node claims). */
typedef struct pgvd_brt_node32 {
  uint32_t lcp;
  int32_t left;
  uint16_t lcp_length;
  uint8_t left_leaf, right_leaf;
} pgvd_brt_node32;
typedef struct pgvd_brt_node64 {
  uint64_t lcp;
  int32_t left;
  uint16_t lcp_length;
  uint8_t left_leaf, right_leaf;
} pgvd_brt_node64;
typedef struct pgvd_brt_node128 {
  uint64_t lcp_lo, lcp_hi;
  int32_t left;
  uint16_t lcp_length;
  uint8_t left_leaf, right_leaf;
} pgvd_brt_node128;

/* ---- the reference's byte layouts (host-side conversion) ----
 * BigUnsigned (C/BigUnsigned.h:8-16): int len, then 11 8-bit blocks, least
 * significant first, with no zero block at len - 1 (zapLeadingZeros,
 * C/BigUnsigned.c:22-26); blocks at and above len are not part of the
 * value.  BrtNode (C/BrtNode.h:10-21): 32 bytes, lcp a BigUnsigned.
 * Callers that read keys or BRT nodes back into the reference's structs
 * (tests/KernelTests.cpp:74,130,190 read BigUnsigned buffers) convert the
 * library's native records with these.  All pointers are host memory;
 * nothing runs on the device. */
typedef struct pgvd_big_unsigned {
  int32_t len;
  uint8_t blk[11];
  uint8_t pad_;
} pgvd_big_unsigned;
typedef struct pgvd_ref_brt_node {
  int32_t left;
  uint8_t left_leaf, right_leaf, pad_[2];
  pgvd_big_unsigned lcp;
  int32_t lcp_length;
  int32_t parent;
} pgvd_ref_brt_node;
/* n native keys (key_bytes 4, 8 or 16) -> BigUnsigned; unused blocks and
 * padding are written as zero.  PGVD_ERR_TOO_LARGE for a key above 88 bits
 * (BIG_INTEGER_SIZE 11). */
int pgvd_keys_to_big_unsigned(const void *keys, int key_bytes, int64_t n, pgvd_big_unsigned *out);
/* n BigUnsigned -> native keys.  PGVD_ERR_INVALID when a len is outside
 * 0..11; PGVD_ERR_TOO_LARGE when a value does not fit key_bytes. */
int pgvd_big_unsigned_to_keys(const pgvd_big_unsigned *in, int64_t n, int key_bytes, void *keys);
/* compareBU (C/BigUnsigned.c:132-153): -1, 0 or 1. */
int pgvd_compare_big_unsigned(const pgvd_big_unsigned *x, const pgvd_big_unsigned *y);
/* count BRT records (pgvd_brt_node32/64/128 by key_bytes) and their parent
 * links -> BrtNode; and back. */
int pgvd_brt_to_ref_nodes(const void *brt, const int32_t *parent, int64_t count, int key_bytes,
                          pgvd_ref_brt_node *out);
int pgvd_ref_nodes_to_brt(const pgvd_ref_brt_node *in, int64_t count, int key_bytes, void *brt,
                          int32_t *parent);

/* Kernels::BuildBinaryRadixTree_p (opencl/Kernels.cpp:249-264, kernel
 * C/BuildBRT.c:90-181): m sorted keys -> m-1 nodes (brt_out, record type by
 * key_bytes) and parent_out (m-1 int32: a child claimed twice keeps its
 * later claimer, as the reference's node loop leaves it).  Equal
 * neighbouring keys give the tree the reference computes from them (where
 * three equal keys end the array, the reference also writes one parent
 * field past its nodes; that write is dropped). */
int pgvd_build_brt(pgvd_ctx *ctx, const void *keys, int64_t m, int key_bytes, int mbits,
                   void *brt_out, int32_t *parent_out, pgvd_stream s);

/* Kernels::ComputeLocalSplits_p (opencl/Kernels.cpp:275-301, kernel
 * C/BuildOctree.c:17-27): ls_out (device, m u32). */
int pgvd_local_splits(pgvd_ctx *ctx, const void *brt, int64_t m, int key_bytes, int dim,
                      uint32_t *ls_out, pgvd_stream s);

/* Octree node (C/OctNode.h:16-26): children[1<<dim] int32 then leaf int32;
 * 20 bytes in 2D, 36 in 3D.  leaf bit o is set when children[o] == -1 (the
 * reference's 2D definition, extended to 8 octants in 3D). */

/* Kernels::BinaryRadixToOctree_p (opencl/Kernels.cpp:333-368): local
 * splits, inclusive scan and octree fill (C/BuildOctree.c:73-139) for a BRT
 * of m keys.  octree_out (device) holds `capacity` nodes; *octree_size
 * (host) receives the node count; PGVD_ERR_CAPACITY if it exceeds
 * capacity.  Synchronizes (the reference reads the size back, :347). */
int pgvd_brt_to_octree(pgvd_ctx *ctx, const void *brt, const int32_t *parent, int64_t m,
                       int key_bytes, int dim, void *octree_out, int64_t capacity,
                       int64_t *octree_size, pgvd_stream s);

/* ---- whole path ---- */

typedef struct pgvd_build_opts {
  int use_custom_bb;        /* Karras::Quantize customBB (Karras.cpp:76-77)      */
  float bb_min[3], bb_max[3];
  int clamped;              /* Karras::Quantize clamped (Karras.cpp:46-49)       */
  int quantized_input;      /* 1: pts are int32 grid coordinates (n*dim),
                               as Kernels::BuildOctree_p takes vector<intn>     */
  int no_host_wait;         /* 1: the call never waits for the device: the sort
                               takes the LSD radix passes instead of the bucket
                               sort, whose levels are decided on the host while
                               the device works (for callers that enqueue builds
                               on several streams from one thread)             */
} pgvd_build_opts;

typedef struct pgvd_octree_result {
  int64_t n;                /* input points                                      */
  int64_t m;                /* unique keys (including the padding zero key)      */
  int64_t octree_size;      /* octree nodes                                      */
  int dim, bits, mbits, key_bytes;
  int malformed;            /* 1 if some BRT split took the FP32-quirk value     */
  int bucket_sort;          /* 1 if the bucket sort ordered the keys, 0 the
                               LSD radix passes                               */
  /* device pointers owned by the context, valid until its next build:       */
  const void *keys;         /* m sorted unique keys                              */
  const void *brt;          /* m-1 pgvd_brt_node{32,64,128}                      */
  const int32_t *parent;    /* m-1 parent links                                  */
  const uint32_t *prefix_sums;   /* m inclusive prefix sums of local splits      */
  const int32_t *octree;    /* octree_size nodes                                 */
  const uint32_t *key_conflict;  /* bit r: run r has >1 label (labels given)     */
  const uint32_t *node_conflict; /* per node: octant mask of conflicting leaves  */
} pgvd_octree_result;

/* Kernels::BuildOctree_p (opencl/Kernels.cpp:413-429) preceded by the
 * bounding box and quantization of Karras::Quantize (Karras.cpp:69-99):
 * float points (device, n*dim) -> octree, all on the device.  labels
 * (device, n u32, nullable) are object ids carried through the sort for the
 * conflict-cell pass.  Enqueues the whole path on the stream and returns;
 * pgvd_build_octree_finish waits and reports the sizes.  With float points
 * the sort is a bucket sort whose levels the host decides from small counts
 * the device reports mid-build: the call then returns once the device has
 * reached the sort's last level (the work queued behind keeps it busy), not
 * at once, unless opts->no_host_wait. */
int pgvd_build_octree(pgvd_ctx *ctx, const void *pts, int64_t n, int dim, int bits,
                      const uint32_t *labels, const pgvd_build_opts *opts, pgvd_stream s);

/* Waits for the last pgvd_build_octree on ctx, grows the octree buffer and
 * refills it if it was too small, and fills *out.  Errors the device
 * detected (PGVD_ERR_BAD_COORD, PGVD_ERR_DUP_KEY) are returned here. */
int pgvd_build_octree_finish(pgvd_ctx *ctx, pgvd_octree_result *out);

/* Allocation-free whole path.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It computes what
 * pgvd_build_octree computes, but every intermediate (run state, keys, BRT,
 * prefix sums, look-back words) lives in a static device workspace of
 * PGVD_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on
 * until the process ends (about 2^20 3D points with 64-bit keys; a larger
 * build fails with PGVD_ERR_TOO_LARGE before anything is enqueued).  No
 * context: it runs on the current HIP device and stream s (NULL: the null
 * stream), calls no hipMalloc or hipFree, and returns once the work is
 * enqueued.  Outputs (device): octree (octree_cap nodes, the records of
 * pgvd_build_octree), node_conflict (octree_cap words, required with
 * labels) and *state, written last on the stream:
 *   status       pgvd_status of the build: PGVD_ERR_BAD_COORD as the whole
 *                path reports it; PGVD_ERR_CAPACITY when the octree has more
 *                than octree_cap nodes (then nothing was written to octree;
 *                octree_size says how many to provide);
 *   m            unique keys (including the padding zero key);
 *   octree_size  octree nodes;
 *   malformed    1 if some BRT split took the FP32-quirk value.
 * Every call on one GPU shares that GPU's workspace, which keeps the last
 * call's contents: calls that overlap on one GPU (other streams or threads)
 * race and must be serialised by the caller. */
#define PGVD_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
typedef struct pgvd_build_state {
  int32_t status;
  uint32_t m;
  uint32_t octree_size;
  uint32_t malformed;
} pgvd_build_state;
int pgvd_build_octree_noalloc(const void *pts, int64_t n, int dim, int bits,
                              const uint32_t *labels, const pgvd_build_opts *opts,
                              int32_t *octree, int64_t octree_cap, uint32_t *node_conflict,
                              pgvd_build_state *state, pgvd_stream s);

/* ---- GVD refinement (2D) ------------------------------------------------
 * The step after the octree in the reference's viewer: Octree2::build(const
 * Polylines&) (viewer/Octree2.cpp:127-190) builds the octree of the
 * polylines' vertices, finds the cells two polylines pass through
 * (FindMultiCells, :430-506, whose call at :186 the reference keeps
 * commented out) and adds the samples Geom::FitBoxes places between the two
 * (opencl/Geom.cpp:467-716) as extra points of the next build.  Every
 * output is bit-exact with the reference's float and double arithmetic.
 * Workspace buffers belong to the context: calls on one context must not
 * overlap on different streams. */

/* CellIntersections (CellIntersections.h:9-55) of one octree node: per
 * octant o the first two labels whose segments reached the cell, at slots
 * 2o and 2o+1 (-1 where empty), and the longest part of each label's
 * segments inside the cell (a strictly longer part replaced the kept one).
 * 160 bytes. */
typedef struct pgvd_cell_intersections {
  int32_t label[8];
  float seg[8][4];          /* ax, ay, bx, by in grid units; 0 where empty */
} pgvd_cell_intersections;

/* FindMultiCells' segments (viewer/Octree2.cpp:442-450): Octree2::obj2Oct
 * (:215-225) of every vertex (verts: device, nv = sum(counts) points, x y
 * float, polygon after polygon; counts: host, P vertex counts), grid
 * coordinates ((v - bb_min) / bbw * 2^bits in float; a coordinate at or past
 * 2^bits becomes 2^bits - 0.0001 rounded to float), then vertex i to i+1
 * of every polygon, labelled with the polygon's index.  segs (device,
 * 4 floats per segment) and labels (device int32) receive the
 * sum(max(counts - 1, 0)) segments; *nseg_out (host) their number.
 * PGVD_ERR_BAD_COORD when a coordinate falls below zero (obj2Oct throws).
 * 1 <= bits <= 30.  Synchronizes. */
int pgvd_polyline_segments(pgvd_ctx *ctx, const float *verts, const int64_t *counts, int64_t P,
                           const float bb_min[2], float bbw, int bits, float *segs,
                           int32_t *labels, int64_t *nseg_out, pgvd_stream s);

/* FindMultiCells' walks (viewer/Octree2.cpp:430-455): each of n segments
 * (device, ax ay bx by in grid units, in the order the reference walks
 * them: polygon by polygon) is walked through the 2D octree (device,
 * octree_size nodes as pgvd_build_octree writes them) as pgvd_cell_walk
 * walks it, and MCCallback (:366-428) keeps, for every visited cell, the
 * segment's part inside it (the two edge intersections, or the one and the
 * endpoint whose truncated coordinates lie in the cell's half-open box)
 * under labels[i] (device int32) in ci (device, octree_size records).
 * walk_status (device, n int32, nullable): 1 where a walk passed 10000
 * cells (the reference throws; the first 10000 are kept).  Returns after
 * the walk's cell count is read back; the rest is asynchronous. */
int pgvd_multi_cells(pgvd_ctx *ctx, const float *segs, const int32_t *labels, int64_t n, int bits,
                     const int32_t *octree, int64_t octree_size, pgvd_cell_intersections *ci,
                     int32_t *walk_status, pgvd_stream s);

/* Geom::FitBoxes(A, B, min_d, &samples, &origins, &lengths) for n pairs
 * (pairs: device, 8 floats each, A's ax ay bx by then B's).  Outputs
 * (device): the samples (2 floats each), box origins (2 floats each) and
 * widths of all pairs, pair after pair, while they fit `capacity` entries;
 * sample_offsets and box_offsets (n + 1 int64 each, nullable): where each
 * pair's entries start; status (n int32, nullable): 0, or the reference's
 * logic_error as 1 ("doesn't support intersections"), 2 ("no end but not
 * intersecting"), 3 ("must include one endpoint"), or 4 when the loop ran
 * 4096 rounds (the reference's has no bound).  *samples_out and *boxes_out
 * (host) receive the totals; PGVD_ERR_CAPACITY when either exceeds
 * capacity (nothing is then written), PGVD_ERR_GEOMETRY when some pair
 * failed (everything is written; the message names the first).
 * Synchronizes. */
int pgvd_fit_boxes(pgvd_ctx *ctx, const float *pairs, int64_t n, float min_d, float *samples,
                   float *origins, float *lengths, int64_t capacity, int64_t *sample_offsets,
                   int64_t *box_offsets, int64_t *samples_out, int64_t *boxes_out, int32_t *status,
                   pgvd_stream s);

/* FindMultiCells' sampling (viewer/Octree2.cpp:457-505): for every node
 * and octant of ci (device, octree_size records) that holds two labels
 * whose parts do not overlap on one line (Geom::multi_intersection,
 * opencl/Geom.cpp:109-151), FitBoxes with min_d; the samples truncated to
 * int (convert_intn) are the extra points, in node, octant and sample
 * order.  Outputs (device, nullable each): extra (2 int32 per point),
 * samples, origins (2 floats each) and lengths, up to capacity entries.
 * *samples_out and *boxes_out (host) receive the totals;
 * PGVD_ERR_CAPACITY or PGVD_ERR_GEOMETRY as pgvd_fit_boxes.  Synchronizes. */
int pgvd_find_multi_cells(pgvd_ctx *ctx, const pgvd_cell_intersections *ci, int64_t octree_size,
                          float min_d, int32_t *extra, float *samples, float *origins,
                          float *lengths, int64_t capacity, int64_t *samples_out,
                          int64_t *boxes_out, pgvd_stream s);

typedef struct pgvd_gvd_result {
  int iterations;           /* octree builds run                                    */
  int64_t n_qpoints;        /* points of the last build: the quantized vertices and
                               every extra point added before it                 */
  int64_t octree_size;      /* nodes of the last octree (0: fewer than 2 points)  */
  int64_t n_segments;
  int64_t n_extra;          /* extra points found on the last octree (not added)  */
  int64_t n_boxes;
  /* device pointers owned by the context, valid until its next call:          */
  const int32_t *qpoints;   /* n_qpoints x 2                                        */
  const int32_t *octree;    /* octree_size nodes of 5 int32                         */
  const pgvd_cell_intersections *ci;  /* octree_size records                       */
  const int32_t *extra;     /* n_extra x 2                                          */
  const float *samples;     /* n_extra x 2 (the extra points before truncation)     */
  const float *origins;     /* n_boxes x 2                                          */
  const float *lengths;     /* n_boxes                                              */
} pgvd_gvd_result;

/* Octree2::build(const Polylines&) (viewer/Octree2.cpp:127-190) with its
 * FindMultiCells call in place, on the device: the vertices (device, nv =
 * sum(counts) 2D points, polygon after polygon; counts on the host) are
 * quantized with Karras::Quantize (their own box, not clamped); the
 * segments come from obj2Oct with custom_bb (host: min x, min y, max x,
 * max y; NULL: the vertices' box).  Then, up to max_iterations times: the
 * octree of the points (BuildOctree_p), the multi-cells, and FitBoxes'
 * extra points (min_d; the reference passes 1), which join the points of
 * the next build; it stops early when a build yields none.
 * extra_per_iteration (host, max_iterations int64, nullable) receives each
 * build's extra-point count.  With fewer than two points no octree is
 * built and nothing is walked (the reference reads an empty octree there).
 * PGVD_ERR_GEOMETRY where the reference throws (FitBoxes' logic_errors, a
 * walk past 10000 cells).  1 <= bits <= 30.  Synchronizes. */
int pgvd_gvd_refine(pgvd_ctx *ctx, const float *verts, const int64_t *counts, int64_t P, int bits,
                    const float *custom_bb, int max_iterations, float min_d,
                    int64_t *extra_per_iteration, pgvd_gvd_result *out, pgvd_stream s);

/* ---- batched form -------------------------------------------------------
 * Many independent octree builds of uneven size through one segmented
 * pipeline: one launch per stage for the whole batch, replayed as a HIP
 * graph.  The reference has no counterpart: its callers loop over
 * Kernels::BuildOctree_p (opencl/Kernels.cpp:413-429) once per point set.
 * Every instance's results are bit-identical to pgvd_build_octree on that
 * instance alone.  Instances hold 1 .. 2^23 points; all share dim and bits;
 * labels are not carried (use pgvd_build_octree for conflict cells). */
typedef struct pgvd_batch pgvd_batch;

typedef struct pgvd_batch_opts {
  int quantized_input;  /* points are int32 grid coordinates (no bbox)          */
  int clamped;          /* quantize to width instead of width-1 (Karras.cpp:71) */
  int use_graph;        /* replay the pipeline as a captured HIP graph (default) */
  int sort_ipt;         /* keys per thread of the sort tile: 8, 12, 15 or 16    */
                        /* (4, 7 or 8 for 128-bit keys); 0: default           */
  int brt_rows;         /* BRT nodes per thread of a workgroup window (0: def.) */
  int digit_bits;       /* radix digit width of the sort passes: 6, 7 or 8 (8 for
                           128-bit keys); 0: 8                                */
  int sort_waves;       /* waves per SIMD the sort kernel is compiled for: 2, 3
                           or 4; 0: 3                                         */
} pgvd_batch_opts;

typedef struct pgvd_batch_result {
  int64_t count;                 /* instances                                    */
  int dim, bits, mbits, key_bytes, octree_stride;
  int64_t total_octree;          /* nodes over all instances                     */
  /* host arrays owned by the plan, count entries each:                        */
  const int64_t *m;              /* unique keys per instance (incl. padding)    */
  const int64_t *octree_size;    /* octree nodes per instance                   */
  const int64_t *key_offset;     /* first key slot of each instance             */
  const int64_t *octree_offset;  /* first octree node of each instance          */
  /* device arrays owned by the plan, valid until its next run:                */
  const void *keys;              /* instance i: keys[key_offset[i] .. +m[i])     */
  const void *brt;               /* same slots, m[i]-1 nodes                     */
  const int32_t *parent;         /* same slots                                   */
  const uint32_t *prefix_sums;   /* same slots, m[i] entries                     */
  const int32_t *octree;         /* instance i: nodes octree_offset[i] .. with   */
                                 /* child indices local to the instance          */
} pgvd_batch_result;

/* Plans a batch of `count` instances with sizes[i] points each: builds the
 * tile tables, allocates every device buffer once and (with use_graph)
 * captures the pipeline on the first run.  opts may be NULL. */
int pgvd_batch_create(pgvd_ctx *ctx, const int64_t *sizes, int64_t count, int dim, int bits,
                      const pgvd_batch_opts *opts, pgvd_batch **out);
int pgvd_batch_destroy(pgvd_batch *b);
/* Enqueues the pipeline for points pts (device; the instances' points one
 * after another, float or int32 per the plan) on the plan's stream. */
int pgvd_batch_run(pgvd_batch *b, const void *pts);
/* Waits for the last run, grows the octree buffer and refills it if it was
 * too small, and fills *out. */
int pgvd_batch_finish(pgvd_batch *b, pgvd_batch_result *out);
/* Picks the plan's sort tile (sort_ipt), BRT window (brt_rows), radix digit
 * width (digit_bits) and the sort kernel's waves per SIMD (sort_waves) by
 * timing runs on pts: every (sort_ipt, brt_rows) pair, then each digit
 * width, then each wave count, keeping the fastest at each step.  Every
 * setting gives bit-identical results.  Any output may be NULL. */
int pgvd_batch_tune(pgvd_batch *b, const void *pts, int *sort_ipt, int *brt_rows,
                    int *digit_bits, int *sort_waves);
/* Milliseconds of the last pgvd_batch_run, from events around it. */
int pgvd_batch_last_ms(pgvd_batch *b, float *ms);

/* ---- served form ---------------------------------------------------------
 * Octree builds requested online.  The server keeps `slots` contexts (one
 * stream and one pooled set of buffers each, device-resident across
 * requests) and an engine thread.  A submitted request is admitted, in
 * arrival order, as soon as a slot is free and the device memory the server
 * would then hold -- every slot's pool and points buffer as allocated, plus
 * the result buffers, held or cached -- fits the HBM budget (cached result
 * buffers, then idle slots' pools, are freed when that makes the request
 * fit); the builds of admitted
 * requests overlap on the device.  No counterpart in the reference. */
typedef struct pgvd_server pgvd_server;

int pgvd_server_create(int device, int slots, size_t hbm_budget_bytes, pgvd_server **out);
/* Drops requests not yet admitted, waits for the ones on the device, and
 * frees every result. */
int pgvd_server_destroy(pgvd_server *s);

/* Queues a build of n points (host_pts: n*dim floats in host memory, best
 * pinned with pgvd_host_alloc, kept valid until the request is complete)
 * and returns its ticket at once.  PGVD_ERR_CAPACITY when its footprint
 * alone exceeds the budget. */
int pgvd_server_submit(pgvd_server *s, const float *host_pts, int64_t n, int dim, int bits,
                       int64_t *ticket);

typedef struct pgvd_served_result {
  int status;               /* pgvd_status of the build                          */
  int dim, bits;
  int slot;                 /* the slot that built it                            */
  int64_t n, m, octree_size;
  uint64_t octree_digest;   /* pgvd_digest of the octree records                 */
  double arrival_s, admit_s, done_s;   /* seconds since the server was created  */
  const int32_t *octree;    /* device: octree_size records of 2^dim + 1 int32,
                               valid until pgvd_server_release                  */
} pgvd_served_result;

/* The result of a ticket: PGVD_OK with *out filled once the request is
 * complete; PGVD_NOT_READY before that when wait is 0 (wait 1 blocks).
 * The build's own outcome is out->status. */
int pgvd_server_poll(pgvd_server *s, int64_t ticket, int wait, pgvd_served_result *out);
/* Gives a complete request's octree back to the server, which keeps the
 * buffer for a later result (cached buffers still count against the budget
 * and are freed first when a request would not fit).  Call it only once the
 * caller's own reads of the octree (on any stream) have completed. */
int pgvd_server_release(pgvd_server *s, int64_t ticket);

/* A trace replayed in real time through the same engine. */
typedef struct pgvd_request {
  /* in */
  const float *host_pts;    /* n*dim floats in host memory (pinned: pgvd_host_alloc) */
  int64_t n;
  int dim, bits;
  double arrival_s;         /* arrival time, seconds after the trace starts */
  /* out */
  double admit_s, done_s;   /* admission and completion times, same clock */
  int64_t m, octree_size;
  uint64_t octree_digest;   /* pgvd_digest of the octree nodes */
  int status;               /* pgvd_status of the build */
  int slot;
} pgvd_request;

typedef struct pgvd_server_stats {
  double elapsed_s;          /* first arrival to last completion */
  int max_in_flight;         /* most requests on the device at once */
  size_t max_in_flight_bytes;   /* most device bytes the server held: slot pools,
                                   points buffers and unreleased results */
  int64_t completed;
  size_t max_device_bytes;   /* most the device's free memory fell below its
                                level at pgvd_server_create (hipMemGetInfo) */
} pgvd_server_stats;

/* Submits reqs[i] at its arrival time (sorted by arrival_s), waits for all,
 * fills each request's results and releases their octrees; totals into
 * *stats (nullable).  Returns the first failing request's status, if any. */
int pgvd_server_run(pgvd_server *s, pgvd_request *reqs, int64_t count, pgvd_server_stats *stats);
/* Device bytes one whole-path build of n points occupies (its admission size). */
size_t pgvd_build_footprint(int64_t n, int dim, int bits);
/* Pinned host memory for request points. */
int pgvd_host_alloc(size_t bytes, void **out);
int pgvd_host_free(void *p);

#ifdef __cplusplus
}
#endif
#endif /* PGVD_H */
