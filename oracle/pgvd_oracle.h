/*
 * pgvd_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A serial CPU restatement of the reference's octree-construction path
 * (edwardsjohnmartin/pgvd, the `_s` functions of opencl/Kernels.cpp and the
 * per-work-item functions in the C/ directory).  It is the checker for the HIP
 * product path and the CPU baseline that bench.py times.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * Keys of up to 128 bits are held as unsigned __int128 in place of the
 * reference's 11-byte BigUnsigned (C/BigUnsigned.h:8-16); numeric value and
 * ordering are the same for every key the reference can hold (<= 88 bits).
 *
 * Every function names the reference file:line it restates.  Where the
 * reference has undefined or crashing behaviour the restatement documents
 * the definition it uses instead (see the per-function comments).
 */
#ifndef PGVD_ORACLE_H
#define PGVD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* opencl/Kernels.cpp:20 -- max((int)pow(2, ceil(log(n)/log(2))), 8). */
int oracle_next_pow2(int num);

/* BoundingBox.h:166-175 (operator()) over all points, then max_size
 * (BoundingBox.h:80-87).  pts is n*dim floats, point-major.
 * Writes bbmin[dim], bbmax[dim]; returns max_size (dwidth). */
float oracle_bbox(const float *pts, int64_t n, int dim, float *bbmin, float *bbmax);

/* Karras.cpp:42-99 (vector Quantize).  width is Resln::width (power of 2).
 * bbmin/dwidth: pass custom_bb != 0 to use (bbmin, bbmax) given by the
 * caller (Karras.cpp:76-77), else the bbox of the data is computed.
 * out receives n*dim int64 coordinates.  Returns the number of output
 * points (n, or 1 when dwidth == 0, Karras.cpp:84-88), or -1 when a
 * coordinate quantizes below zero (Karras.cpp:56-63 throws there). */
int64_t oracle_quantize(const float *pts, int64_t n, int dim, int64_t width,
                        int custom_bb, const float *bbmin_in,
                        const float *bbmax_in, int clamped, int64_t *out);

/* C/z_order.h:14-47 -- key bit i*dim+k = bit i of coordinate k, i < bits.
 * Returns the key as (lo, hi) 64-bit words. */
void oracle_xyz2z(const int64_t *p, int dim, int bits, uint64_t *lo, uint64_t *hi);

/* xyz2z over n points (no padding). */
void oracle_points_to_keys(const int64_t *p, int64_t n, int dim, int bits,
                           uint64_t *lo, uint64_t *hi);

/* C/BuildBRT.c:54-88 -- the reference's shift-search LCP length. */
int oracle_compute_lcp_length(uint64_t alo, uint64_t ahi, uint64_t blo,
                              uint64_t bhi, int mbits);

/* opencl/Kernels.cpp:125-159 -- inclusive Hillis-Steele scan. */
void oracle_stream_scan(const uint32_t *in, uint32_t *out, int size);

/* ------------------------------------------------------------------ */
/* Whole path: opencl/Kernels.cpp:388-411 (BuildOctree_s).             */
/* ------------------------------------------------------------------ */
typedef struct oracle_tree oracle_tree;

/* From quantized points (n*dim int64).  Runs PointsToMorton_s with
 * zero-key padding, sort, unique, BuildBinaryRadixTree_s and
 * BinaryRadixToOctree_s.  labels (nullable, n entries) feed the
 * point-label conflict pass.  Returns NULL on bad arguments. */
oracle_tree *oracle_build(const int64_t *qpts, int64_t n, int dim, int bits,
                          const uint32_t *labels);

/* Same pipeline from already-sorted unique keys (for stage tests). */
oracle_tree *oracle_build_from_keys(const uint64_t *lo, const uint64_t *hi,
                                    int64_t m, int dim, int mbits);

int64_t oracle_tree_m(const oracle_tree *t);
int64_t oracle_tree_octree_size(const oracle_tree *t);
void oracle_tree_keys(const oracle_tree *t, uint64_t *lo, uint64_t *hi);
/* BrtNode fields (C/BrtNode.h:10-21), m-1 entries each. */
void oracle_tree_brt(const oracle_tree *t, int32_t *left, uint8_t *left_leaf,
                     uint8_t *right_leaf, uint64_t *lcp_lo, uint64_t *lcp_hi,
                     int32_t *lcp_length, int32_t *parent);
/* local splits and their inclusive prefix sums, m entries each. */
void oracle_tree_splits(const oracle_tree *t, uint32_t *ls, uint32_t *ps);
/* OctNode (C/OctNode.h:16-26): children[(1<<dim)] then leaf, per node. */
void oracle_tree_octree(const oracle_tree *t, int32_t *children, int32_t *leaf);
/* Conflict flags per unique key (m) and per-octree-node octant masks (O). */
int oracle_tree_conflicts(const oracle_tree *t, uint8_t *key_flags,
                          uint32_t *node_masks);
void oracle_tree_free(oracle_tree *t);

/* OctreeUtils::FindLeaf (OctreeUtils.cpp:18-71) for each of n grid points
 * p (n*dim int64, coordinates below 2^bits), with the octant masked to dim
 * bits as the commented-out line there (:37-41) intends.  octree: O nodes of
 * 2^dim + 1 int32 (children, leaf).  Per point: parent = the node whose
 * octant is the leaf cell (the OctCell's parent_idx), octant, data =
 * children[octant] of that node, level = the number of digits descended
 * (the cell's width is 2^(bits - level)), status = 0, 1 when no leaf is
 * reached (the reference throws) or 2 when a child index leaves the octree
 * (not an octree the build produced). */
void oracle_find_leaf(const int64_t *p, int64_t n, int dim, int bits, const int32_t *octree,
                      int64_t O, int32_t *parent, int32_t *octant, int32_t *data,
                      int32_t *level, int32_t *status);

/* Text form of an octree: operator<< / operator>> for OctNode and
 * std::vector<OctNode> (C/OctNode.h:58-98, commented out in the reference).
 * format writes "N " then, per node, each child and the leaf mask in
 * decimal, each followed by one space; returns the length (out may be NULL
 * to measure).  parse reads N and then N nodes of 2^dim + 1 integers
 * separated by whitespace (the leaf cast to unsigned char, :76-77); a token
 * that is not an optional sign and digits (leading zeros skipped) with a
 * value in int32 range, a negative count, or too few tokens, returns -1;
 * more than `cap` nodes returns -2.  A token must end at whitespace (where
 * operator>> would read a numeric prefix of the last token and stop). */
/* CellWalk (viewer/Octree2.cpp:265-313) for n 2D segments (segs: ax, ay,
 * bx, by per segment, grid units): the leaf cells it visits, in order,
 * located with oracle_find_leaf.  Per segment counts[i] cells and status[i]
 * (0; 1 after 10000 cells, where the reference throws).  Cells are written
 * as {origin x, origin y, width, parent, octant} int32 records while fewer
 * than cap have been written; returns the total, or -1 when bits is outside
 * 1..30 (the cells are int32). */
int64_t oracle_cell_walk(const float *segs, int64_t n, int bits, const int32_t *octree, int64_t O,
                         int32_t *counts, int32_t *status, int32_t *cells, int64_t cap);

int64_t oracle_octree_format_text(const int32_t *octree, int64_t O, int dim, char *out);
int oracle_octree_parse_text(const char *text, int64_t len, int dim, int32_t *octree,
                             int64_t cap, int64_t *O_out);

/* Timing helper for the CPU baseline: runs the whole serial path from
 * float points (bbox, quantize, BuildOctree_s) and returns the octree size,
 * or -1 on error.  *m_out receives the unique key count. */
int64_t oracle_run_pipeline(const float *pts, int64_t n, int dim, int bits,
                            int64_t *m_out);

/* BuildBinaryRadixTree_s alone on m sorted keys, equal neighbours allowed:
 * m-1 nodes (a child claimed twice keeps its later claimer); -1 on bad
 * arguments. */
int oracle_brt_only(const uint64_t *lo, const uint64_t *hi, int64_t m, int mbits, int32_t *left,
                    uint8_t *left_leaf, uint8_t *right_leaf, uint64_t *lcp_lo, uint64_t *lcp_hi,
                    int32_t *lcp_length, int32_t *parent);

/* ---- GVD refinement (2D) ------------------------------------------- */
/* OctreeUtils::FindIntersections (OctreeUtils.cpp:80-172): rows (t, x, y). */
int oracle_find_intersections(const float *a, const float *b, const int32_t *origin, int width,
                              float *out);
/* Geom::multi_intersection (opencl/Geom.cpp:109-151). */
int oracle_multi_intersection(const float *p1, const float *p2, const float *p3, const float *p4);
/* Geom::FitBoxes (opencl/Geom.cpp:467-716) of segments a, b (ax ay bx by):
 * returns the sample count; samples (2 floats each), box origins (2 each)
 * and widths written while fewer than cap; *boxes_out the box count;
 * *status 0 or the FitBoxes error (1-3 the reference's logic_errors, 4 its
 * loop past 4096 rounds). */
int64_t oracle_fit_boxes(const float *a, const float *b, float min_d, float *samples,
                         float *origins, float *lengths, int64_t cap, int64_t *boxes_out,
                         int *status);
/* 1 when the reference's FitBoxes reads an unset segment on these inputs. */
int oracle_fit_boxes_undefined(const float *a, const float *b, float min_d);
/* FindMultiCells' walks with MCCallback (viewer/Octree2.cpp:366-455):
 * CellIntersections of every node, ci_labels O*8 (-1 empty), ci_segs O*32;
 * nothing is walked when bits is outside 1..30. */
void oracle_multi_cells(const float *segs, const int32_t *labels, int64_t n, int bits,
                        const int32_t *octree, int64_t O, int32_t *ci_labels, float *ci_segs,
                        int32_t *status);
/* FindMultiCells' sampling (viewer/Octree2.cpp:457-505): extra points. */
int64_t oracle_find_multi_cells(const int32_t *ci_labels, const float *ci_segs, int64_t O,
                                float min_d, int32_t *extra, float *samples, float *origins,
                                float *lengths, int64_t cap, int64_t *boxes_out, int *status);
/* Octree2::obj2Oct (viewer/Octree2.cpp:215-225). */
int oracle_obj2oct(const float *v, int64_t n, const float *bbmin, float bbw, int64_t width,
                   float *out);

#ifdef __cplusplus
}
#endif
#endif
