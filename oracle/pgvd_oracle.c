/*
 * pgvd_oracle.c -- TEST INFRASTRUCTURE ONLY (see pgvd_oracle.h).
 *
 * Serial restatement of the reference octree-construction path.  Citations
 * are file:line in edwardsjohnmartin/pgvd.  Compile with
 * -ffp-contract=off: the quantize expression is bit-relevant.
 */
#include "pgvd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static inline u128 mk(uint64_t lo, uint64_t hi) { return ((u128)hi << 64) | lo; }

/* opencl/Kernels.cpp:20 */
int oracle_next_pow2(int num) {
  int p = (int)pow(2, ceil(log(num) / log(2)));
  return p > 8 ? p : 8;
}

/* BoundingBox.h:166-175: first point initialises, then component-wise
 * (a<b?a:b) / (a>b?a:b) (opencl/vec_cpp.h:255-277).  max_size:
 * BoundingBox.h:80-87. */
float oracle_bbox(const float *pts, int64_t n, int dim, float *bbmin, float *bbmax) {
  for (int k = 0; k < dim; ++k) { bbmin[k] = pts[k]; bbmax[k] = pts[k]; }
  for (int64_t i = 1; i < n; ++i) {
    for (int k = 0; k < dim; ++k) {
      const float v = pts[i * dim + k];
      bbmin[k] = (bbmin[k] < v) ? bbmin[k] : v;
      bbmax[k] = (bbmax[k] > v) ? bbmax[k] : v;
    }
  }
  float m = bbmax[0] - bbmin[0];
  for (int k = 1; k < dim; ++k) {
    const float s = bbmax[k] - bbmin[k];
    m = (s > m) ? s : m;
  }
  return m;
}

/* Karras.cpp:42-67 for one point.  The expression is evaluated exactly as
 * the reference writes it: (p-min)/dwidth in float, times the int width
 * converted to float, widened to double, +0.5, truncated.  int64 replaces
 * int so that widths above 2^31 are representable (the reference's int
 * overflows there). */
static int quantize_point(const float *p, const float *bbmin, float dwidth,
                          int64_t eff_width, int dim, int64_t *q) {
  for (int k = 0; k < dim; ++k) {
    const float t = (p[k] - bbmin[k]) / dwidth;
    const float f = (float)eff_width * t;
    const double d = (double)f;
    const int64_t v = (int64_t)(d + 0.5);
    if (v < 0) return -1;
    q[k] = v;
  }
  return 0;
}

/* Karras.cpp:69-99 */
int64_t oracle_quantize(const float *pts, int64_t n, int dim, int64_t width,
                        int custom_bb, const float *bbmin_in,
                        const float *bbmax_in, int clamped, int64_t *out) {
  if (n <= 0) return 0;
  float bbmin[3] = {0, 0, 0}, bbmax[3] = {0, 0, 0}, dwidth;
  if (custom_bb) {
    for (int k = 0; k < dim; ++k) { bbmin[k] = bbmin_in[k]; bbmax[k] = bbmax_in[k]; }
    dwidth = bbmax[0] - bbmin[0];
    for (int k = 1; k < dim; ++k) {
      const float s = bbmax[k] - bbmin[k];
      dwidth = (s > dwidth) ? s : dwidth;
    }
  } else {
    dwidth = oracle_bbox(pts, n, dim, bbmin, bbmax);
  }
  if (dwidth == 0) {              /* Karras.cpp:84-88: a single {0,0} */
    for (int k = 0; k < dim; ++k) out[k] = 0;
    return 1;
  }
  const int64_t eff = clamped ? width : width - 1;   /* Karras.cpp:46-49 */
  for (int64_t i = 0; i < n; ++i)
    if (quantize_point(pts + i * dim, bbmin, dwidth, eff, dim, out + i * dim)) return -1;
  return n;
}

/* C/z_order.h:14-47: bit by bit, key bit i*DIM+k <- bit i of coordinate k.
 * The reference tests `p.x & (1 << i)` (int shift, so it is limited to
 * i <= 30); this restatement tests bit i of a 64-bit coordinate. */
static u128 xyz2z(const int64_t *p, int dim, int bits) {
  u128 r = 0;
  for (int i = 0; i < bits; ++i)
    for (int k = 0; k < dim; ++k)
      if ((uint64_t)p[k] & ((uint64_t)1 << i)) r |= (u128)1 << (i * dim + k);
  return r;
}

void oracle_xyz2z(const int64_t *p, int dim, int bits, uint64_t *lo, uint64_t *hi) {
  const u128 z = xyz2z(p, dim, bits);
  *lo = (uint64_t)z;
  *hi = (uint64_t)(z >> 64);
}

void oracle_points_to_keys(const int64_t *p, int64_t n, int dim, int bits,
                           uint64_t *lo, uint64_t *hi) {
  for (int64_t i = 0; i < n; ++i) oracle_xyz2z(p + i * dim, dim, bits, &lo[i], &hi[i]);
}

/* BigUnsigned right shift (C/BigUnsigned.c:417-471) on a native 128-bit
 * value: every bit shifted past the end is dropped, shifts >= 128 give 0. */
static inline u128 shr(u128 a, int s) { return s >= 128 ? 0 : (a >> s); }

/* C/BuildBRT.c:54-88: binary search over right-shift amounts for the
 * smallest shift that makes a and b equal. */
static int compute_lcp_length(u128 a, u128 b, int mbits) {
  unsigned int v = mbits;
  v--; v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16; v++;
  int offset = v >> 1;
  for (int i = v >> 2; i > 0; i >>= 1) {
    if (shr(a, offset) == shr(b, offset)) offset -= i;
    else offset += i;
  }
  if (shr(a, offset) == shr(b, offset)) {
    if (shr(a, offset - 1) == shr(b, offset - 1)) return mbits - (offset - 1);
    return mbits - offset;
  }
  return mbits - (offset + 1);
}

int oracle_compute_lcp_length(uint64_t alo, uint64_t ahi, uint64_t blo,
                              uint64_t bhi, int mbits) {
  return compute_lcp_length(mk(alo, ahi), mk(blo, bhi), mbits);
}

/* C/BuildBRT.c:34-52: mask the top `length` bits of an mbits-wide value
 * and shift them down. */
static u128 compute_lcp(u128 value, int length, int mbits) {
  u128 mask = 0;
  for (int i = 0; i < length; ++i) mask |= (u128)1 << (mbits - 1 - i);
  return shr(value & mask, mbits - length);
}

/* opencl/Kernels.cpp:125-159 with C/ParallelAlgorithms.c:33-51:
 * zero-padded inclusive Hillis-Steele scan, ping-ponging two buffers. */
void oracle_stream_scan(const uint32_t *in, uint32_t *out, int size) {
  if (size <= 0) return;
  const int p2 = (int)pow(2, ceil(log(size) / log(2)));
  uint32_t *a = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)p2);
  uint32_t *b = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)p2);
  for (int i = 0; i < size; ++i) a[i] = b[i] = in[i];
  for (int i = size; i < p2; ++i) a[i] = b[i] = 0;
  for (unsigned int s = 1; s < (unsigned int)p2; s <<= 1) {
    for (int j = 0; j < p2; ++j)
      b[j] = (j > (int)(s - 1)) ? a[j] + a[j - s] : a[j];
    uint32_t *t = a; a = b; b = t;
  }
  for (int i = 0; i < size; ++i) out[i] = a[i];
  free(a);
  free(b);
}

/* ------------------------------------------------------------------ */

struct oracle_tree {
  int dim, mbits;
  int64_t m, O;
  u128 *keys;                 /* m sorted unique keys                  */
  /* BrtNode, C/BrtNode.h:10-21 */
  int32_t *left, *lcp_length, *parent;
  uint8_t *left_leaf, *right_leaf;
  u128 *lcp;
  uint32_t *ls, *ps;          /* m each                                 */
  int32_t *children, *leaf;   /* O * (1<<dim), O                        */
  uint8_t *key_conflict;      /* m, or NULL without labels             */
};

static int cmp_u128(const void *a, const void *b) {
  const u128 x = *(const u128 *)a, y = *(const u128 *)b;
  return (x > y) - (x < y);
}

typedef struct { u128 key; uint32_t label; uint32_t real; } keyed_label;
static int cmp_keyed(const void *a, const void *b) {
  const keyed_label *x = (const keyed_label *)a, *y = (const keyed_label *)b;
  if (x->key != y->key) return (x->key > y->key) - (x->key < y->key);
  return 0;
}

/* C/BuildBRT.c:90-181 for one gid, including the quirks the reference
 * has:  the l_max loop tests a key loaded one doubling earlier (:118-126),
 * and the split step is the FP32 expression of :151.  The out-of-range
 * read at :119 (gid=1, d=-1 wraps the unsigned index) is not performed;
 * its value is never used because the range test at :120-121 fails
 * first. */
static void build_brt_node(oracle_tree *t, int size, int gid) {
  const u128 *mp = t->keys;
  const int mbits = t->mbits;
  if (!(gid < size - 1)) return;
  const u128 current = mp[gid];
  int d;
  if (gid == 0) d = 1;
  else {
    const int l_pos = compute_lcp_length(current, mp[gid + 1], mbits);
    const int l_neg = compute_lcp_length(current, mp[gid - 1], mbits);
    d = (l_pos - l_neg > 0) - (l_pos - l_neg < 0);
  }
  int l;
  if (gid == 0) {
    l = size - 1;
  } else {
    const int lcp_min = (d == -1) ? compute_lcp_length(current, mp[gid + 1], mbits)
                                  : compute_lcp_length(current, mp[gid - 1], mbits);
    int l_max = 2;
    long long idx = (long long)gid + (long long)l_max * d;
    u128 temp = (idx >= 0 && idx <= size - 1) ? mp[idx] : 0;
    for (;;) {
      idx = (long long)gid + (long long)l_max * d;
      if (!(idx >= 0 && idx <= size - 1)) break;
      if (!(compute_lcp_length(current, temp, mbits) > lcp_min)) break;
      temp = mp[idx];
      l_max = l_max << 1;
    }
    const int l_cutoff = (d == -1) ? gid : size - gid - 1;
    l = 0;
    for (int tt = l_max >> 1; tt >= 1; tt >>= 1) {
      if (l + tt <= l_cutoff) {
        if (compute_lcp_length(current, mp[gid + (l + tt) * d], mbits) > lcp_min) l = l + tt;
      }
    }
  }
  const int j = gid + l * d;
  const int lcp_node = compute_lcp_length(current, mp[j], mbits);
  const int s_cutoff = (d == -1) ? gid - 1 : size - gid - 2;
  int s = 0;
  for (int den = 2; den < 2 * l; den *= 2) {
    const int tt = (int)((l + (float)den - 1) / den);
    if (s + tt <= s_cutoff) {
      if (compute_lcp_length(current, mp[gid + (s + tt) * d], mbits) > lcp_node) s = s + tt;
    }
  }
  const int split = gid + s * d + (d < 0 ? d : 0);
  const int mn = gid < j ? gid : j, mx = gid > j ? gid : j;
  t->left[gid] = split;
  t->left_leaf[gid] = (mn == split);
  t->right_leaf[gid] = (mx == split + 1);
  t->lcp[gid] = compute_lcp(current, lcp_node, mbits);
  t->lcp_length[gid] = lcp_node;
  if (gid == 0) t->parent[gid] = -1;
  if (!t->left_leaf[gid]) t->parent[split] = gid;
  if (!t->right_leaf[gid]) t->parent[split + 1] = gid;
}

/* C/BuildOctree.c:50-64 */
static int quadrant_in_lcp(const oracle_tree *t, int node, int i) {
  const int dim = t->dim;
  const int mask = (dim == 2) ? 3 : 7;
  const int rshift = i * dim + t->lcp_length[node] % dim;
  return (int)(shr(t->lcp[node], rshift) & (u128)mask);
}

static void *zalloc(size_t n) { return calloc(n ? n : 1, 1); }

/* BRT, local splits, scan and octree from t->keys[0..m).  Follows
 * opencl/Kernels.cpp:404-409 and :370-386. */
static void build_tree(oracle_tree *t) {
  const int64_t m = t->m;
  const int dim = t->dim;
  const int nc = 1 << dim;
  const int64_t ni = m > 1 ? m - 1 : 0;
  /* vector<BrtNode> I(numPoints - 1): value-initialised (zero). */
  t->left = zalloc(sizeof(int32_t) * ni);
  t->lcp_length = zalloc(sizeof(int32_t) * ni);
  t->parent = zalloc(sizeof(int32_t) * ni);
  t->left_leaf = zalloc(ni);
  t->right_leaf = zalloc(ni);
  t->lcp = zalloc(sizeof(u128) * ni);
  /* BuildBinaryRadixTree_s, opencl/Kernels.cpp:266-273 */
  for (int i = 0; i < (int)m - 1; ++i) build_brt_node(t, (int)m, i);

  t->ls = zalloc(sizeof(uint32_t) * m);
  t->ps = zalloc(sizeof(uint32_t) * m);
  if (m == 1) {
    /* The reference reads I[0] of an empty vector here
     * (opencl/Kernels.cpp:305-306); defined as lcp_length 0, so one node. */
    t->ls[0] = 1;
    t->ps[0] = 1;
  } else {
    /* ComputeLocalSplits_s, opencl/Kernels.cpp:303-313 with
     * C/BuildOctree.c:17-27 */
    t->ls[0] = 1 + t->lcp_length[0] / dim;
    for (int i = 0; i < (int)m - 1; ++i) {
      const int local = t->lcp_length[i] / dim;
      const int left = t->left[i], right = left + 1;
      if (!t->left_leaf[i]) t->ls[left] = (uint32_t)(t->lcp_length[left] / dim - local);
      if (!t->right_leaf[i]) t->ls[right] = (uint32_t)(t->lcp_length[right] / dim - local);
    }
    oracle_stream_scan(t->ls, t->ps, (int)m);
  }
  /* BinaryRadixToOctree_s, opencl/Kernels.cpp:370-386 */
  const int64_t O = t->ps[m - 1];
  t->O = O;
  t->children = zalloc(sizeof(int32_t) * (size_t)O * nc);
  t->leaf = zalloc(sizeof(int32_t) * (size_t)O);
  /* brt2octree_init, C/BuildOctree.c:134-139.  The reference sets
   * leaf = 15 and uses leaf_masks = {1,2,4,8} (C/OctNode.h:7), which is
   * defined for DIM == 2 only; the 3D definition here is all 2^DIM bits
   * set and mask 1<<octant. */
  const int32_t full = (1 << nc) - 1;
  for (int64_t i = 0; i < O; ++i) {
    t->leaf[i] = full;
    for (int c = 0; c < nc; ++c) t->children[i * nc + c] = -1;
  }
  /* brt2octree, C/BuildOctree.c:73-132, for brt_i = 1 .. m-2 in order */
  for (int brt_i = 1; brt_i < (int)m - 1; ++brt_i) {
    if (t->ls[brt_i] == 0) continue;
    const int num_splits = (int)t->ls[brt_i];
    int current = (int)t->ps[brt_i - 1];
    for (int i = 0; i < num_splits - 1; ++i) {
      const int oct_parent = current + 1;
      const int onode = quadrant_in_lcp(t, brt_i, i);
      t->children[(int64_t)oct_parent * nc + onode] = current;
      t->leaf[oct_parent] &= ~(1 << onode);
      current = oct_parent;
    }
    int brt_parent = t->parent[brt_i];
    while (t->ls[brt_parent] == 0) brt_parent = t->parent[brt_parent];
    const int oct_parent = (brt_parent == 0) ? 0 : (int)t->ps[brt_parent - 1];
    const int q = quadrant_in_lcp(t, brt_i, num_splits - 1);
    t->children[(int64_t)oct_parent * nc + q] = current;
    t->leaf[oct_parent] &= ~(1 << q);
  }
}

static oracle_tree *new_tree(int dim, int mbits) {
  oracle_tree *t = (oracle_tree *)calloc(1, sizeof(oracle_tree));
  t->dim = dim;
  t->mbits = mbits;
  return t;
}

/* opencl/Kernels.cpp:388-411 (BuildOctree_s).  With labels, the key sort
 * carries (label, real) so the point-label conflict flags can be
 * computed per unique key; padding keys carry no label. */
oracle_tree *oracle_build(const int64_t *qpts, int64_t n, int dim, int bits,
                          const uint32_t *labels) {
  if (n <= 0 || (dim != 2 && dim != 3) || bits <= 0 || bits * dim > 128) return NULL;
  oracle_tree *t = new_tree(dim, bits * dim);
  const int64_t rn = oracle_next_pow2((int)n);
  keyed_label *kl = (keyed_label *)malloc(sizeof(keyed_label) * (size_t)rn);
  /* PointsToMorton_s, opencl/Kernels.cpp:62-75 */
  for (int64_t g = 0; g < rn; ++g) {
    if (g < n) {
      kl[g].key = xyz2z(qpts + g * dim, dim, bits);
      kl[g].label = labels ? labels[g] : 0;
      kl[g].real = 1;
    } else {
      kl[g].key = 0;
      kl[g].label = 0;
      kl[g].real = 0;
    }
  }
  /* sort (ascending) and unique, opencl/Kernels.cpp:401-402 */
  qsort(kl, (size_t)rn, sizeof(keyed_label), cmp_keyed);
  t->keys = (u128 *)malloc(sizeof(u128) * (size_t)rn);
  if (labels) t->key_conflict = (uint8_t *)zalloc((size_t)rn);
  int64_t m = 0;
  int64_t run_start = 0;
  for (int64_t i = 0; i <= rn; ++i) {
    if (i == rn || (i > 0 && kl[i].key != kl[i - 1].key)) {
      /* close the run [run_start, i) */
      if (labels) {
        int have = 0;
        uint32_t lo = 0, hi = 0;
        for (int64_t k = run_start; k < i; ++k) {
          if (!kl[k].real) continue;
          if (!have) { lo = hi = kl[k].label; have = 1; }
          else { if (kl[k].label < lo) lo = kl[k].label; if (kl[k].label > hi) hi = kl[k].label; }
        }
        t->key_conflict[m] = (uint8_t)(have && lo != hi);
      }
      t->keys[m++] = kl[run_start].key;
      run_start = i;
    }
  }
  free(kl);
  t->m = m;
  build_tree(t);
  return t;
}

oracle_tree *oracle_build_from_keys(const uint64_t *lo, const uint64_t *hi,
                                    int64_t m, int dim, int mbits) {
  if (m <= 0 || (dim != 2 && dim != 3) || mbits <= 0 || mbits > 128) return NULL;
  oracle_tree *t = new_tree(dim, mbits);
  t->m = m;
  t->keys = (u128 *)malloc(sizeof(u128) * (size_t)m);
  for (int64_t i = 0; i < m; ++i) t->keys[i] = mk(lo[i], hi ? hi[i] : 0);
  build_tree(t);
  return t;
}

int64_t oracle_tree_m(const oracle_tree *t) { return t->m; }
int64_t oracle_tree_octree_size(const oracle_tree *t) { return t->O; }

void oracle_tree_keys(const oracle_tree *t, uint64_t *lo, uint64_t *hi) {
  for (int64_t i = 0; i < t->m; ++i) {
    lo[i] = (uint64_t)t->keys[i];
    if (hi) hi[i] = (uint64_t)(t->keys[i] >> 64);
  }
}

void oracle_tree_brt(const oracle_tree *t, int32_t *left, uint8_t *left_leaf,
                     uint8_t *right_leaf, uint64_t *lcp_lo, uint64_t *lcp_hi,
                     int32_t *lcp_length, int32_t *parent) {
  const int64_t ni = t->m > 1 ? t->m - 1 : 0;
  for (int64_t i = 0; i < ni; ++i) {
    left[i] = t->left[i];
    left_leaf[i] = t->left_leaf[i];
    right_leaf[i] = t->right_leaf[i];
    lcp_lo[i] = (uint64_t)t->lcp[i];
    if (lcp_hi) lcp_hi[i] = (uint64_t)(t->lcp[i] >> 64);
    lcp_length[i] = t->lcp_length[i];
    parent[i] = t->parent[i];
  }
}

void oracle_tree_splits(const oracle_tree *t, uint32_t *ls, uint32_t *ps) {
  memcpy(ls, t->ls, sizeof(uint32_t) * (size_t)t->m);
  memcpy(ps, t->ps, sizeof(uint32_t) * (size_t)t->m);
}

void oracle_tree_octree(const oracle_tree *t, int32_t *children, int32_t *leaf) {
  const int nc = 1 << t->dim;
  memcpy(children, t->children, sizeof(int32_t) * (size_t)t->O * nc);
  memcpy(leaf, t->leaf, sizeof(int32_t) * (size_t)t->O);
}

/* Point-label conflict cells.  The reference's version
 * (CellIntersections.h:9-55, viewer/Octree2.cpp:366-506) walks polyline
 * segments through the octree and keeps up to two labels per
 * (node, octant) leaf slot, `is_multi` meaning a second label is present.
 * This restatement is the point-label form: a unique key (one octree leaf
 * cell) is in conflict when the labels of the points that map to it are
 * not all equal, and the flag is also reported on the (node, octant) slot
 * that holds the key: the octree node of the key's BRT parent (walking up
 * through nodes with no local split, as brt2octree's attach does,
 * C/BuildOctree.c:102-112), octant = the key's digit at that node's
 * level. */
int oracle_tree_conflicts(const oracle_tree *t, uint8_t *key_flags, uint32_t *node_masks) {
  if (!t->key_conflict) return -1;
  memcpy(key_flags, t->key_conflict, (size_t)t->m);
  memset(node_masks, 0, sizeof(uint32_t) * (size_t)t->O);
  const int dim = t->dim, mbits = t->mbits;
  const int mask = (1 << dim) - 1;
  for (int64_t x = 0; x + 1 < t->m; ++x) {
    for (int side = 0; side < 2; ++side) {
      const int is_leaf = side ? t->right_leaf[x] : t->left_leaf[x];
      if (!is_leaf) continue;
      const int64_t c = t->left[x] + side;
      if (!t->key_conflict[c]) continue;
      int64_t q = x;
      while (q != 0 && t->ls[q] == 0) q = t->parent[q];
      const int64_t T = (q == 0) ? 0 : t->ps[q - 1];
      const int lvl = t->lcp_length[q] / dim;
      const int sh = mbits - (lvl + 1) * dim;
      const int o = sh >= 0 ? (int)(shr(t->keys[c], sh) & (u128)mask) : 0;
      node_masks[T] |= 1u << o;
    }
  }
  return 0;
}

void oracle_tree_free(oracle_tree *t) {
  if (!t) return;
  free(t->keys); free(t->left); free(t->lcp_length); free(t->parent);
  free(t->left_leaf); free(t->right_leaf); free(t->lcp);
  free(t->ls); free(t->ps); free(t->children); free(t->leaf);
  free(t->key_conflict);
  free(t);
}

int64_t oracle_run_pipeline(const float *pts, int64_t n, int dim, int bits, int64_t *m_out) {
  int64_t *q = (int64_t *)malloc(sizeof(int64_t) * (size_t)n * dim);
  const int64_t nq = oracle_quantize(pts, n, dim, (int64_t)1 << bits, 0, NULL, NULL, 0, q);
  if (nq < 0) { free(q); return -1; }
  oracle_tree *t = oracle_build(q, nq, dim, bits, NULL);
  free(q);
  if (!t) return -1;
  const int64_t O = t->O;
  if (m_out) *m_out = t->m;
  oracle_tree_free(t);
  return O;
}

/* OctreeUtils::FindLeaf (OctreeUtils.cpp:18-71): the point's key digit by
 * digit from the top (:33-40, the octant masked to dim bits, the fix the
 * commented-out line :37 intends), width halved per digit (:41), and the
 * descent stops at the first octant whose leaf bit is set (:47-51); a node
 * visited without one is an error (:55-69). */
void oracle_find_leaf(const int64_t *p, int64_t n, int dim, int bits, const int32_t *octree,
                      int64_t O, int32_t *parent, int32_t *octant, int32_t *data,
                      int32_t *level, int32_t *status) {
  const int nc = 1 << dim, stride = nc + 1, mbits = bits * dim;
  for (int64_t q = 0; q < n; ++q) {
    uint64_t lo, hi;
    oracle_xyz2z(p + q * dim, dim, bits, &lo, &hi);
    const u128 z = mk(lo, hi);
    int64_t idx = 0;
    int o = 0, lev = 0, st = 1;
    int32_t d = -1;
    for (int i = mbits - dim; i >= 0; i -= dim) {
      o = (int)(shr(z, i) & (u128)(nc - 1));
      ++lev;
      const int32_t *node = octree + idx * stride;
      if (node[nc] & (1 << o)) {
        d = node[o];
        st = 0;
        break;
      }
      if (node[o] < 0 || node[o] >= O) {
        st = 2;
        break;
      }
      idx = node[o];
    }
    parent[q] = (int32_t)idx;
    octant[q] = o;
    data[q] = d;
    level[q] = lev;
    status[q] = st;
  }
}

/* C/OctNode.h:58-87 (operator<< for OctNode and for the vector): the size,
 * a space, then every node followed by a space, a node being its children
 * and its leaf (as int) separated by spaces. */
static int64_t put_int(char *out, int64_t at, int32_t v) {
  char tmp[12];
  int k = 0;
  uint32_t u = v < 0 ? (uint32_t)0 - (uint32_t)v : (uint32_t)v;
  do { tmp[k++] = (char)('0' + u % 10); u /= 10; } while (u);
  if (v < 0) tmp[k++] = '-';
  for (int j = 0; j < k; ++j) { if (out) out[at + j] = tmp[k - 1 - j]; }
  if (out) out[at + k] = ' ';
  return at + k + 1;
}

int64_t oracle_octree_format_text(const int32_t *octree, int64_t O, int dim, char *out) {
  const int stride = (1 << dim) + 1;
  int64_t at = 0;
  char hdr[24];
  int h = 0;
  int64_t n = O;
  do { hdr[h++] = (char)('0' + n % 10); n /= 10; } while (n);
  for (int j = 0; j < h; ++j) { if (out) out[at + j] = hdr[h - 1 - j]; }
  at += h;
  if (out) out[at] = ' ';
  ++at;
  for (int64_t i = 0; i < O * stride; ++i) at = put_int(out, at, octree[i]);
  return at;
}

/* C/OctNode.h:69-98 (operator>>): whitespace-separated integers. */
static int is_ws(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

static int next_int(const char *t, int64_t len, int64_t *at, int64_t *v) {
  int64_t i = *at;
  while (i < len && is_ws(t[i])) ++i;
  if (i >= len) return -1;
  int neg = 0;
  if (t[i] == '+' || t[i] == '-') { neg = t[i] == '-'; ++i; }
  int64_t x = 0;
  int nd = 0, digit = 0;   /* significant digits; leading zeros skipped */
  while (i < len && !is_ws(t[i])) {
    if (t[i] < '0' || t[i] > '9') return -1;
    digit = 1;
    if (nd || t[i] != '0') {
      if (nd >= 10) return -1;
      x = x * 10 + (t[i] - '0');
      ++nd;
    }
    ++i;
  }
  if (!digit) return -1;
  if (neg) x = -x;
  *at = i;
  *v = x;
  return 0;
}

int oracle_octree_parse_text(const char *text, int64_t len, int dim, int32_t *octree,
                             int64_t cap, int64_t *O_out) {
  const int stride = (1 << dim) + 1;
  int64_t at = 0, n;
  if (next_int(text, len, &at, &n) || n < 0) return -1;
  *O_out = n;
  if (n > cap) return -2;
  for (int64_t i = 0; i < n * stride; ++i) {
    int64_t v;
    if (next_int(text, len, &at, &v) || v < INT32_MIN || v > INT32_MAX) return -1;
    octree[i] = (i % stride == stride - 1) ? (int32_t)(uint8_t)v : (int32_t)v;
  }
  return 0;
}

/* ---- CellWalk ---------------------------------------------------------- */
/* OctreeUtils.h:24-30 */
typedef struct { float t; float p[2]; } cell_isect;

/* opencl/vec_cpp.h:526-530, 546-548, 573-575: dot accumulates from 0 in
 * order, length = sqrt(dot(v, v)). */
static float dot2(const float *u, const float *v) {
  float r = 0;
  r += u[0] * v[0];
  r += u[1] * v[1];
  return r;
}

#define CW_EPSILON 1e-6 /* opencl/defs.h:135, a double */

/* OctreeUtils.cpp:80-172: the segment ab against the four edges of the
 * cell (bottom, right, top, left), kept when t is in [-eps, len + eps)
 * (compared in double) and the point is in the cell grown by eps (float);
 * two points closer than eps count once; of three or more, the pair whose
 * direction best matches v (largest |dot|, first such pair) is kept.  When
 * no pair's |dot| exceeds 0 the reference keeps two default-constructed
 * intersections (uninitialised); here the first two are kept. */
static int cw_find_intersections(const float a[2], const float b[2], const int32_t origin[2],
                                 int width, cell_isect ret[4]) {
  float v[2] = {b[0] - a[0], b[1] - a[1]};
  const float len = sqrtf(dot2(v, v));
  v[0] = v[0] / len;
  v[1] = v[1] / len;
  const float bmin[2] = {(float)origin[0], (float)origin[1]};
  const float bmax[2] = {(float)origin[0] + (float)width, (float)origin[1] + (float)width};
  float ts[4], ps[4][2];
  {
    const int y = origin[1];
    ts[0] = ((float)y - a[1]) / v[1];
    ps[0][0] = a[0] + v[0] * ts[0];
    ps[0][1] = (float)y;
  }
  {
    const int x = origin[0] + width;
    ts[1] = ((float)x - a[0]) / v[0];
    ps[1][0] = (float)x;
    ps[1][1] = a[1] + v[1] * ts[1];
  }
  {
    const int y = origin[1] + width;
    ts[2] = ((float)y - a[1]) / v[1];
    ps[2][0] = a[0] + v[0] * ts[2];
    ps[2][1] = (float)y;
  }
  {
    const int x = origin[0];
    ts[3] = ((float)x - a[0]) / v[0];
    ps[3][0] = (float)x;
    ps[3][1] = a[1] + v[1] * ts[3];
  }
  const float eps = (float)CW_EPSILON;
  int n = 0;
  for (int i = 0; i < 4; ++i) {
    const float t = ts[i];
    if ((double)t >= -CW_EPSILON && (double)t < (double)len + CW_EPSILON) {
      int in = 1;
      for (int k = 0; k < 2; ++k)   /* BoundingBox.h:105-111 */
        if (ps[i][k] < bmin[k] - eps || ps[i][k] > bmax[k] + eps) in = 0;
      if (in) {
        ret[n].t = t;
        ret[n].p[0] = ps[i][0];
        ret[n].p[1] = ps[i][1];
        ++n;
      }
    }
  }
  if (n == 2) {
    const float d[2] = {ret[0].p[0] - ret[1].p[0], ret[0].p[1] - ret[1].p[1]};
    if ((double)sqrtf(dot2(d, d)) < CW_EPSILON) n = 1;   /* opencl/Geom.h:16-18 */
  }
  if (n > 2) {
    int bi = 0, bj = 1;
    float best = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        float u[2] = {ret[j].p[0] - ret[i].p[0], ret[j].p[1] - ret[i].p[1]};
        const float l = sqrtf(dot2(u, u));
        u[0] = u[0] / l;
        u[1] = u[1] / l;
        const float cd = fabsf(dot2(u, v));
        if (cd > best) { best = cd; bi = i; bj = j; }
      }
    const cell_isect x = ret[bi], y = ret[bj];
    ret[0] = x;
    ret[1] = y;
    n = 2;
  }
  return n;
}

/* FindLeaf of a float point after convert_int2 (truncation, opencl/vec_cpp.h
 * CONVERT_VEC2); origin = the sum of the octant offsets along the descent
 * (OctreeUtils.cpp:41-47), i.e. the point's low `bits` bits (those its key
 * holds) with the low log2(width) bits cleared: a coordinate of exactly the
 * grid width (obj2Oct's clamp rounds back to it in float from 2^12 up) lands
 * in the cell at 0. */
static void cw_leaf(const float p[2], int bits, const int32_t *octree, int64_t O, int32_t cell[5]) {
  const int64_t q[2] = {(int32_t)p[0], (int32_t)p[1]};
  int32_t parent, octant, data, level, st;
  oracle_find_leaf(q, 1, 2, bits, octree, O, &parent, &octant, &data, &level, &st);
  const int32_t width = (int32_t)(1u << (bits - level));
  const int32_t gmask = (int32_t)((1u << bits) - 1);   /* the key holds the low `bits` bits */
  cell[0] = ((int32_t)q[0] & gmask) & ~(width - 1);
  cell[1] = ((int32_t)q[1] & gmask) & ~(width - 1);
  cell[2] = width;
  cell[3] = parent;
  cell[4] = octant;
}

int64_t oracle_cell_walk(const float *segs, int64_t n, int bits, const int32_t *octree, int64_t O,
                         int32_t *counts, int32_t *status, int32_t *cells, int64_t cap) {
  if (bits < 1 || bits > 30) return -1;   /* int32 cells, as pgvd_cell_walk checks */
  int64_t total = 0;
  for (int64_t s = 0; s < n; ++s) {
    const float *a = segs + 4 * s, *b = segs + 4 * s + 2;
    int dir[2];
    for (int i = 0; i < 2; ++i) dir[i] = b[i] > a[i] ? 1 : (b[i] < a[i] ? -1 : 0);
    cell_isect cur = {0.0f, {a[0], a[1]}};
    int32_t bcell[5], cell[5], ncell[5];
    cw_leaf(b, bits, octree, O, bcell);
    cw_leaf(cur.p, bits, octree, O, cell);
    int count = 0, st = 0, done = 0;
    do {
      if (++count > 10000) { st = 1; --count; break; }
      if (cells && total < cap) memcpy(cells + 5 * total, cell, sizeof cell);
      ++total;
      cell_isect li[4];
      const int nl = cw_find_intersections(a, b, cell, cell[2], li);
      for (int i = 0; i < nl; ++i)
        if (li[i].t > cur.t) cur = li[i];
      for (int i = 0; i < 2; ++i) {
        const int end = cell[i];
        if (cur.p[i] == (float)end && cur.p[i] != 0 && dir[i] == -1) cur.p[i] -= 1;
      }
      cw_leaf(cur.p, bits, octree, O, ncell);
      done = nl == 0 || (cell[0] == bcell[0] && cell[1] == bcell[1]) ||
             (ncell[0] == cell[0] && ncell[1] == cell[1]);
      memcpy(cell, ncell, sizeof cell);
    } while (!done);
    counts[s] = count;
    status[s] = st;
  }
  return total;
}

/* ---- GVD refinement: multi-cells and FitBoxes (2D) ---------------------- */
/* The reference's vector type is OpenCL's cl_float2 with the component-wise
 * operators of opencl/vec_cpp.h (v op scalar and v op v per component;
 * dot accumulates from 0; length = sqrt(dot); unit = v / length), and
 * FloatSegment (opencl/FloatSegment.h) is two of them. */
typedef struct { float x, y; } gv2;
typedef struct { gv2 a, b; } gseg;

static gv2 g_mk(float x, float y) { gv2 r; r.x = x; r.y = y; return r; }
static gv2 g_add(gv2 u, gv2 v) { return g_mk(u.x + v.x, u.y + v.y); }
static gv2 g_sub(gv2 u, gv2 v) { return g_mk(u.x - v.x, u.y - v.y); }
static gv2 g_muls(gv2 u, float s) { return g_mk(u.x * s, u.y * s); }
static gv2 g_divs(gv2 u, float s) { return g_mk(u.x / s, u.y / s); }
static float g_dot(gv2 u, gv2 v) { float r = 0; r += u.x * v.x; r += u.y * v.y; return r; }
static float g_len(gv2 v) { return sqrtf(g_dot(v, v)); }
static gv2 g_unit(gv2 v) { return g_divs(v, g_len(v)); }
static int g_eq(gv2 u, gv2 v) { return u.x == v.x && u.y == v.y; }
static float g_dist(gv2 a, gv2 b) { return g_len(g_sub(a, b)); }          /* opencl/Geom.h:16-18 */
static gseg g_seg(gv2 a, gv2 b) { gseg s; s.a = a; s.b = b; return s; }
static gv2 g_vec(gseg s) { return g_sub(s.b, s.a); }
static gseg g_rev(gseg s) { return g_seg(s.b, s.a); }
static int g_sgn(float f) { return f < 0 ? -1 : (f > 0 ? 1 : 0); }         /* opencl/Geom.h:10-12 */

/* Geom::closest(p, s) (opencl/Geom.cpp:15-29); its EPSILON is a local
 * float 1e-6. */
static gv2 g_closest_pt(gv2 p, gseg s) {
  if (g_eq(s.a, s.b)) return s.a;
  const float eps = (float)1e-6;
  gv2 v = g_vec(s);
  const float len = g_len(v);
  v = g_divs(v, len);
  const gv2 u = g_sub(p, s.a);
  const float t = g_dot(u, v);
  gv2 q = g_add(s.a, g_muls(v, t));
  if (t < 0 || g_dist(q, s.a) < eps) q = s.a;
  else if (t > len || g_dist(q, s.b) < eps) q = s.b;
  return q;
}

/* Geom::x_given_y, y_given_x (opencl/Geom.cpp:31-53). */
static int g_x_given_y(gseg s, float y, float *x) {
  const gv2 p = s.a, v = g_vec(s);
  const float t = (y - p.y) / v.y;
  *x = p.x + t * v.x;
  return t <= 1;
}
static int g_y_given_x(gseg s, float x, float *y) {
  const gv2 p = s.a, v = g_vec(s);
  const float t = (x - p.x) / v.x;
  *y = p.y + t * v.y;
  return t <= 1;
}

/* Geom::clip_to_box (opencl/Geom.cpp:55-96). */
static int g_clip_to_box(gseg s, float d, gseg *out) {
  const gv2 c = s.a;
  const float d2 = d / 2;
  const gv2 sv = g_vec(s);
  int valid;
  float x, y;
  if (fabsf(sv.x) > fabsf(sv.y)) {
    x = c.x + d2;
    if (sv.x < 0) x = c.x - d2;
    valid = g_y_given_x(s, x, &y);
  } else {
    y = c.y + d2;
    if (sv.y < 0) y = c.y - d2;
    valid = g_x_given_y(s, y, &x);
  }
  if (valid) *out = g_seg(g_mk(x, y), s.b);
  return valid;
}

/* Geom::multi_intersection (opencl/Geom.cpp:109-151): parallel segments on
 * one line that overlap in more than a point.  The slope is a float
 * quotient widened to double; the collinearity test compares in double. */
int oracle_multi_intersection(const float *p1, const float *p2, const float *p3, const float *p4) {
  const float x1 = p1[0], x2 = p2[0], x3 = p3[0], x4 = p4[0];
  const float y1 = p1[1], y2 = p2[1], y3 = p3[1], y4 = p4[1];
  const float d = (x1 - x2) * (y3 - y4) - (y1 - y2) * (x3 - x4);
  if (d == 0) {
    if (x2 - x1 != 0) {
      const double m = (y2 - y1) / (x2 - x1);
      if ((double)(y1 - y3) == m * (double)(x1 - x3)) {
        const float len1 = fabsf(x2 - x1), len2 = fabsf(x4 - x3);
        const float len = fmaxf(x1, fmaxf(x2, fmaxf(x3, x4))) - fminf(x1, fminf(x2, fminf(x3, x4)));
        if (len < len1 + len2) return 1;
      }
    } else if (x1 == x3) {
      const float len1 = fabsf(y2 - y1), len2 = fabsf(y4 - y3);
      const float len = fmaxf(y1, fmaxf(y2, fmaxf(y3, y4))) - fminf(y1, fminf(y2, fminf(y3, y4)));
      if (len < len1 + len2) return 1;
    }
  }
  return 0;
}

/* Geom::line_intersection (opencl/Geom.cpp:165-206), in double; the point
 * is rounded to float. */
static gv2 g_line_isect(gv2 p1, gv2 p2, gv2 p3, gv2 p4, int *lines, int *segs) {
  const double x1 = p1.x, x2 = p2.x, x3 = p3.x, x4 = p4.x;
  const double y1 = p1.y, y2 = p2.y, y3 = p3.y, y4 = p4.y;
  const double d = (x1 - x2) * (y3 - y4) - (y1 - y2) * (x3 - x4);
  if (d == 0) {
    *lines = 0;
    *segs = 0;
    return g_mk(0, 0);
  }
  *lines = 1;
  const double pre = (x1 * y2 - y1 * x2), post = (x3 * y4 - y3 * x4);
  const double x = (pre * (x3 - x4) - (x1 - x2) * post) / d;
  const double y = (pre * (y3 - y4) - (y1 - y2) * post) / d;
  *segs = 1;
  if (x < fmin(x1, x2) || x > fmax(x1, x2) || x < fmin(x3, x4) || x > fmax(x3, x4)) *segs = 0;
  if (y < fmin(y1, y2) || y > fmax(y1, y2) || y < fmin(y3, y4) || y > fmax(y3, y4)) *segs = 0;
  return g_mk((float)x, (float)y);
}

/* Geom::closest(a, b, ...) for two segments (opencl/Geom.cpp:259-346). */
static void g_closest_segs(gseg a, gseg b, gv2 *ca, gv2 *cb, int *ca_end, int *cb_end) {
  int li, si;
  const gv2 l = g_line_isect(a.a, a.b, b.a, b.b, &li, &si);
  if (si) {
    *ca = *cb = l;
    *ca_end = *cb_end = 0;
    return;
  }
  gv2 c = g_closest_pt(a.a, b);
  *ca = a.a;
  *cb = c;
  *ca_end = 1;
  *cb_end = g_eq(c, b.a) || g_eq(c, b.b);
  float dist = g_len(g_sub(a.a, c));

  gv2 c_ = g_closest_pt(a.b, b);
  float dist_ = g_len(g_sub(a.b, c_));
  if (dist_ < dist) {
    c = c_;
    *ca = a.b;
    *cb = c;
    *ca_end = 1;
    *cb_end = g_eq(c, b.a) || g_eq(c, b.b);
    dist = dist_;
  }
  c_ = g_closest_pt(b.a, a);
  dist_ = g_len(g_sub(b.a, c_));
  if (dist_ < dist) {
    c = c_;
    *ca = c;
    *cb = b.a;
    *ca_end = g_eq(c, a.a) || g_eq(c, a.b);
    *cb_end = 1;
    dist = dist_;
  }
  c_ = g_closest_pt(b.b, a);
  dist_ = g_len(g_sub(b.b, c_));
  if (dist_ < dist) {
    c = c_;
    *ca = c;
    *cb = b.b;
    *ca_end = g_eq(c, a.a) || g_eq(c, a.b);
    *cb_end = 1;
    dist = dist_;
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const gv2 a0 = i ? a.b : a.a, a1 = i ? a.a : a.b;
      const gv2 b0 = j ? b.b : b.a, b1 = j ? b.a : b.b;
      if (a0.x == b0.x) {
        const int day = g_sgn(a1.y - a0.y), dby = g_sgn(b1.y - b0.y), daby = g_sgn(b0.y - a0.y);
        if (day != dby && day != daby) {
          *ca = a0;
          *cb = b0;
          *ca_end = *cb_end = 1;
        }
      } else if (a0.y == b0.y) {
        const int dax = g_sgn(a1.x - a0.x), dbx = g_sgn(b1.x - b0.x), dabx = g_sgn(b0.x - a0.x);
        if (dax != dbx && dax != dabx) {
          *ca = a0;
          *cb = b0;
          *ca_end = *cb_end = 1;
        }
      }
    }
}

/* Output of FitBoxes: samples, counted past the capacity. */
typedef struct {
  float *samples;
  int64_t cap, n;
} g_out;

static void g_emit(g_out *o, gv2 s0, int two, gv2 s1) {
  const int64_t k = o->n;
  if (o->samples) {
    if (k < o->cap) { o->samples[2 * k] = s0.x; o->samples[2 * k + 1] = s0.y; }
    if (two && k + 1 < o->cap) { o->samples[2 * k + 2] = s1.x; o->samples[2 * k + 3] = s1.y; }
  }
  o->n += two ? 2 : 1;
}

/* Box origins and widths, in the order the reference pushes them. */
typedef struct {
  float *origins, *lengths;
  int64_t cap, n;
} g_boxes;

static void g_box(g_boxes *b, gv2 org, float len) {
  if (b->origins && b->n < b->cap) {
    b->origins[2 * b->n] = org.x;
    b->origins[2 * b->n + 1] = org.y;
    b->lengths[b->n] = len;
  }
  ++b->n;
}

#define FIT_OFFSET_FACTOR 4
#define FIT_MAX_ROUNDS 4096

/* Geom::FitBoxesNoIntersection (opencl/Geom.cpp:471-664).  Returns 0, or
 * the reference's logic_error as 1 ("doesn't support intersections"), 2 ("no
 * end but not intersecting"), 3 ("must include one endpoint"); 4 when the
 * loop runs FIT_MAX_ROUNDS rounds (the reference has no bound). */
static int g_fit_no_isect(gseg A, gseg B, float min_d, g_out *o, g_boxes *bx) {
  int done = 0, rounds = 0;
  while (!done) {
    if (++rounds > FIT_MAX_ROUNDS) return 4;
    gv2 ca, cb, c_cb;
    int ca_end, cb_end, c_cb_end;
    g_closest_segs(A, B, &ca, &cb, &ca_end, &cb_end);
    if (g_eq(ca, cb)) return 1;
    if (!ca_end && !cb_end) return 2;
    if (cb_end) {
      gv2 t = ca; ca = cb; cb = t;
      int te = ca_end; ca_end = cb_end; cb_end = te;
      gseg ts = A; A = B; B = ts;
    }
    if (!ca_end) return 3;
    if (!g_eq(A.a, ca)) A = g_rev(A);
    if (cb_end && !g_eq(B.a, cb)) B = g_rev(B);

    float d;
    gv2 org;
    gv2 off = g_mk(0, 0);
    if (g_dist(ca, cb) < min_d) {
      org = g_sub(g_divs(g_add(ca, cb), 2), g_mk(min_d / 2, min_d / 2));
      d = min_d;
      c_cb = cb;
      c_cb_end = cb_end;
    } else if (cb_end && (A.a.x == B.a.x || A.a.y == B.a.y)) {
      d = g_dist(A.a, B.a);
      org = g_mk(fminf(A.a.x, B.a.x), fminf(A.a.y, B.a.y));
      if (A.a.x == B.a.x) {
        off = g_mk(d / FIT_OFFSET_FACTOR, 0);
        if (g_vec(A).x < 0) {
          org = g_sub(org, g_mk(d, 0));
          off = g_muls(off, -1);
        }
      } else if (A.a.y == B.a.y) {
        off = g_mk(0, d / 16);
        if (g_vec(A).y < 0) {
          org = g_sub(org, g_mk(0, d));
          off = g_muls(off, -1);
        }
      }
      c_cb = cb;
      c_cb_end = 1;
    } else {
      const gv2 v = g_sub(cb, ca);
      d = g_len(v);
      const float cx = (v.x >= 0) ? 2 * d : -2 * d;
      const float cy = (v.y >= 0) ? 2 * d : -2 * d;
      const gseg Cs = g_seg(ca, g_add(ca, g_mk(cx, cy)));
      gv2 c_cc;
      int c_cc_end;
      g_closest_segs(B, Cs, &c_cb, &c_cc, &c_cb_end, &c_cc_end);
      float d_ = g_dist(ca, c_cb) / (float)sqrt(2.0);
      if (!g_eq(c_cb, c_cc)) d_ = fmaxf(fabsf(c_cb.x - ca.x), fabsf(c_cb.y - ca.y));
      if (d_ < d) d = d_;
      const float bxd = (v.x >= 0) ? 0 : -d;
      const float byd = (v.y >= 0) ? 0 : -d;
      org = g_add(ca, g_mk(bxd, byd));
    }
    const gv2 s0 = g_add(g_add(ca, g_divs(g_sub(cb, ca), FIT_OFFSET_FACTOR)), off);
    const gv2 s1 = g_add(g_add(cb, g_divs(g_sub(ca, cb), FIT_OFFSET_FACTOR)), off);
    g_emit(o, s0, 1, s1);
    g_box(bx, org, d);

    if (c_cb_end) {
      if (!g_eq(B.a, c_cb)) B = g_rev(B);
    } else {
      const gseg B0 = g_seg(c_cb, B.a), B1 = g_seg(c_cb, B.b);
      if (g_dot(g_unit(g_vec(A)), g_unit(g_vec(B0))) > 0) B = B0;
      else B = B1;
    }

    if (g_dot(g_unit(g_vec(A)), g_unit(g_vec(B))) <= 0) {
      done = 1;
    } else {
      const gv2 bis = g_unit(g_add(g_unit(g_vec(A)), g_unit(g_vec(B))));
      int horizontal;
      if (A.a.x == B.a.x) horizontal = 1;
      else if (A.a.y == B.a.y) horizontal = 0;
      else horizontal = fabsf(bis.x) > fabsf(bis.y);
      float Ax, Ay, Bx, By;
      int Ai, Bi;
      if (horizontal) {
        Ax = Bx = (bis.x < 0) ? org.x : org.x + d;
        Ai = g_y_given_x(A, Ax, &Ay);
        Bi = g_y_given_x(B, Bx, &By);
      } else {
        Ay = By = (bis.y < 0) ? org.y : org.y + d;
        Ai = g_x_given_y(A, Ay, &Ax);
        Bi = g_x_given_y(B, By, &Bx);
      }
      done = !Ai || !Bi;
      if (g_eq(A.a, g_mk(Ax, Ay)) && g_eq(B.a, g_mk(Bx, By))) done = 1;
      A = g_seg(g_mk(Ax, Ay), A.b);
      B = g_seg(g_mk(Bx, By), B.b);
    }
  }
  return 0;
}

/* Geom::FitBoxes (opencl/Geom.cpp:667-716).  Where the segments meet, the
 * reference clips the four half-segments to the box around the meeting
 * point and reads the unit vectors of the first pair even when clipping
 * failed (default-constructed, uninitialised segments); here a failed clip
 * leaves a zero segment, whose unit vector is NaN, so the pairs are not
 * swapped. */
static int g_fit_boxes_u(gseg A, gseg B, float min_d, g_out *o, g_boxes *bx, int *undef) {
  gv2 ca, cb;
  int ca_end, cb_end;
  g_closest_segs(A, B, &ca, &cb, &ca_end, &cb_end);
  if (undef) *undef = 0;
  if (g_eq(ca, cb)) {
    g_emit(o, ca, 0, ca);
    g_box(bx, g_sub(ca, g_mk(min_d / 2, min_d / 2)), min_d);
    if (A.a.x > A.b.x) A = g_rev(A);
    if (B.a.x > B.b.x) B = g_rev(B);
    const gseg A0o = g_seg(ca, A.a), A1o = g_seg(ca, A.b);
    const gseg B0o = g_seg(cb, B.a), B1o = g_seg(cb, B.b);
    gseg A0, A1, B0, B1;
    memset(&A0, 0, sizeof A0); memset(&A1, 0, sizeof A1);
    memset(&B0, 0, sizeof B0); memset(&B1, 0, sizeof B1);
    int A0v = g_clip_to_box(A0o, min_d, &A0);
    int A1v = g_clip_to_box(A1o, min_d, &A1);
    const int B0v = g_clip_to_box(B0o, min_d, &B0);
    const int B1v = g_clip_to_box(B1o, min_d, &B1);
    if (undef && (!A0v || !B0v)) *undef = 1;   /* the reference reads an unset segment */
    const float dotp = g_dot(g_unit(g_vec(A0)), g_unit(g_vec(B0)));
    if (dotp < 0) {
      gseg t = A0; A0 = A1; A1 = t;
      int tv = A0v; A0v = A1v; A1v = tv;
    }
    int rc = 0;
    if (A0v && B0v) rc = g_fit_no_isect(A0, B0, min_d, o, bx);
    if (!rc && A1v && B1v) rc = g_fit_no_isect(A1, B1, min_d, o, bx);
    return rc;
  }
  return g_fit_no_isect(A, B, min_d, o, bx);
}

static int g_fit_boxes(gseg A, gseg B, float min_d, g_out *o, g_boxes *bx) {
  return g_fit_boxes_u(A, B, min_d, o, bx, NULL);
}

/* 1 when the reference's FitBoxes on these segments reads a segment it
 * never set (the meeting-point branch with a failed clip of A0 or B0,
 * opencl/Geom.cpp:684-697): its result then depends on stack contents. */
int oracle_fit_boxes_undefined(const float *a, const float *b, float min_d) {
  g_out o = {NULL, 0, 0};
  g_boxes bx = {NULL, NULL, 0, 0};
  int undef = 0;
  const gseg A = g_seg(g_mk(a[0], a[1]), g_mk(a[2], a[3]));
  const gseg B = g_seg(g_mk(b[0], b[1]), g_mk(b[2], b[3]));
  (void)g_fit_boxes_u(A, B, min_d, &o, &bx, &undef);
  return undef;
}

int64_t oracle_fit_boxes(const float *a, const float *b, float min_d, float *samples,
                         float *origins, float *lengths, int64_t cap, int64_t *boxes_out,
                         int *status) {
  g_out o = {samples, cap, 0};
  g_boxes bx = {origins, lengths, cap, 0};
  const gseg A = g_seg(g_mk(a[0], a[1]), g_mk(a[2], a[3]));
  const gseg B = g_seg(g_mk(b[0], b[1]), g_mk(b[2], b[3]));
  *status = g_fit_boxes(A, B, min_d, &o, &bx);
  if (boxes_out) *boxes_out = bx.n;
  return o.n;
}

/* CellIntersections::set (CellIntersections.h:20-34): per octant two label
 * slots; a new label takes the first empty slot, a known label keeps its
 * longer segment (strictly longer replaces), a third label is dropped. */
static void ci_set(int32_t *l, float *segs, int octant, int32_t label, gseg s) {
  for (int i = octant * 2; i < (octant + 1) * 2; ++i) {
    if (l[i] == -1) {
      l[i] = label;
      segs[4 * i] = s.a.x; segs[4 * i + 1] = s.a.y;
      segs[4 * i + 2] = s.b.x; segs[4 * i + 3] = s.b.y;
      break;
    }
    if (l[i] == label) {
      const gseg cur = g_seg(g_mk(segs[4 * i], segs[4 * i + 1]), g_mk(segs[4 * i + 2], segs[4 * i + 3]));
      if (g_dot(g_vec(s), g_vec(s)) > g_dot(g_vec(cur), g_vec(cur))) {
        segs[4 * i] = s.a.x; segs[4 * i + 1] = s.a.y;
        segs[4 * i + 2] = s.b.x; segs[4 * i + 3] = s.b.y;
      }
      break;
    }
  }
}

/* MCCallback (viewer/Octree2.cpp:366-428) for one visited cell: the
 * segment's part inside the cell (two edge intersections, or one and the
 * endpoint whose truncated coordinates lie in the cell's half-open box),
 * stored under the cell's parent node and octant.  Returns 1 and the part
 * when it is stored. */
static int mc_part(const float a[2], const float b[2], const int32_t cell[5], gseg *part) {
  cell_isect li[4];
  const int n = cw_find_intersections(a, b, cell, cell[2], li);
  if (n == 2) {
    *part = g_seg(g_mk(li[0].p[0], li[0].p[1]), g_mk(li[1].p[0], li[1].p[1]));
    return 1;
  }
  if (n == 1) {
    const int32_t mn[2] = {cell[0], cell[1]}, mx[2] = {cell[0] + cell[2], cell[1] + cell[2]};
    const int32_t ai[2] = {(int32_t)a[0], (int32_t)a[1]}, bi[2] = {(int32_t)b[0], (int32_t)b[1]};
    gv2 q;
    if (ai[0] >= mn[0] && ai[0] < mx[0] && ai[1] >= mn[1] && ai[1] < mx[1]) q = g_mk(a[0], a[1]);
    else if (bi[0] >= mn[0] && bi[0] < mx[0] && bi[1] >= mn[1] && bi[1] < mx[1]) q = g_mk(b[0], b[1]);
    else return 0;   /* degenerate: nothing stored */
    *part = g_seg(g_mk(li[0].p[0], li[0].p[1]), q);
    return 1;
  }
  return 0;
}

/* FindMultiCells' walk (viewer/Octree2.cpp:430-455): every segment (n, as
 * ax ay bx by in grid units, in the reference's polygon-by-polygon order)
 * walked with CellWalk and the MCCallback visitor into the
 * CellIntersections of every octree node (ci_labels O*8, ci_segs O*32).
 * status[s] = 1 when segment s's walk passed 10000 cells. */
void oracle_multi_cells(const float *segs, const int32_t *labels, int64_t n, int bits,
                        const int32_t *octree, int64_t O, int32_t *ci_labels, float *ci_segs,
                        int32_t *status) {
  for (int64_t i = 0; i < O * 8; ++i) ci_labels[i] = -1;
  memset(ci_segs, 0, sizeof(float) * (size_t)(O * 32));
  if (bits < 1 || bits > 30) return;   /* int32 cells, as pgvd_multi_cells checks */
  for (int64_t s = 0; s < n; ++s) {
    const float *a = segs + 4 * s, *b = segs + 4 * s + 2;
    int dir[2];
    for (int i = 0; i < 2; ++i) dir[i] = b[i] > a[i] ? 1 : (b[i] < a[i] ? -1 : 0);
    cell_isect cur = {0.0f, {a[0], a[1]}};
    int32_t bcell[5], cell[5], ncell[5];
    cw_leaf(b, bits, octree, O, bcell);
    cw_leaf(cur.p, bits, octree, O, cell);
    int count = 0, st = 0, done = 0;
    do {
      if (++count > 10000) { st = 1; break; }
      gseg part;
      if (mc_part(a, b, cell, &part) && cell[3] >= 0 && cell[3] < O)
        ci_set(ci_labels + 8 * (int64_t)cell[3], ci_segs + 32 * (int64_t)cell[3], cell[4], labels[s], part);
      cell_isect li[4];
      const int nl = cw_find_intersections(a, b, cell, cell[2], li);
      for (int i = 0; i < nl; ++i)
        if (li[i].t > cur.t) cur = li[i];
      for (int i = 0; i < 2; ++i) {
        const int end = cell[i];
        if (cur.p[i] == (float)end && cur.p[i] != 0 && dir[i] == -1) cur.p[i] -= 1;
      }
      cw_leaf(cur.p, bits, octree, O, ncell);
      done = nl == 0 || (cell[0] == bcell[0] && cell[1] == bcell[1]) ||
             (ncell[0] == cell[0] && ncell[1] == cell[1]);
      memcpy(cell, ncell, sizeof cell);
    } while (!done);
    if (status) status[s] = st;
  }
}

/* FindMultiCells' sampling (viewer/Octree2.cpp:457-505): for every node and
 * octant holding two labels whose segments do not overlap on one line
 * (Geom::multi_intersection), FitBoxes with min_d; the samples truncated to
 * int (convert_intn) are the extra points, in node, octant, sample order.
 * Returns the number of extra points (written up to cap); boxes get the
 * box origins and widths (nullable).  *status: 0, or the first FitBoxes
 * error (see g_fit_no_isect). */
int64_t oracle_find_multi_cells(const int32_t *ci_labels, const float *ci_segs, int64_t O,
                                float min_d, int32_t *extra, float *samples, float *origins,
                                float *lengths, int64_t cap, int64_t *boxes_out, int *status) {
  g_out o = {samples, cap, 0};
  g_boxes bx = {origins, lengths, cap, 0};
  *status = 0;
  for (int64_t i = 0; i < O; ++i)
    for (int oc = 0; oc < 4; ++oc) {
      const int32_t *l = ci_labels + 8 * i + 2 * oc;
      const int nl = l[0] < 0 ? 0 : (l[1] < 0 ? 1 : 2);
      if (nl < 2) continue;
      const float *s0 = ci_segs + 32 * i + 8 * oc, *s1 = s0 + 4;
      if (oracle_multi_intersection(s0, s0 + 2, s1, s1 + 2)) continue;
      const gseg A = g_seg(g_mk(s0[0], s0[1]), g_mk(s0[2], s0[3]));
      const gseg B = g_seg(g_mk(s1[0], s1[1]), g_mk(s1[2], s1[3]));
      const int64_t before = o.n;
      const int rc = g_fit_boxes(A, B, min_d, &o, &bx);
      if (rc && !*status) *status = rc;
      if (extra && samples)
        for (int64_t k = before; k < o.n && k < cap; ++k) {
          extra[2 * k] = (int32_t)samples[2 * k];
          extra[2 * k + 1] = (int32_t)samples[2 * k + 1];
        }
    }
  if (boxes_out) *boxes_out = bx.n;
  return o.n;
}

/* Octree2::obj2Oct (viewer/Octree2.cpp:215-225): ((v - min) / bbw) * width
 * per component in float; a coordinate at or past the width becomes
 * width - 0.0001 (a double rounded to float).  Returns -1 when a coordinate
 * is below zero (the reference throws). */
int oracle_obj2oct(const float *v, int64_t n, const float *bbmin, float bbw, int64_t width,
                   float *out) {
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 2; ++k) {
      float c = ((v[2 * i + k] - bbmin[k]) / bbw) * (float)width;
      if (c >= (float)width) c = (float)((double)width - 0.0001);
      if (c < 0) return -1;
      out[2 * i + k] = c;
    }
  return 0;
}

/* OctreeUtils::FindIntersections (OctreeUtils.cpp:80-172) of segment ab
 * with the cell at origin of the given width: up to 4 rows (t, x, y). */
int oracle_find_intersections(const float *a, const float *b, const int32_t *origin, int width,
                              float *out) {
  cell_isect r[4];
  const int n = cw_find_intersections(a, b, origin, width, r);
  for (int i = 0; i < n; ++i) {
    out[3 * i] = r[i].t;
    out[3 * i + 1] = r[i].p[0];
    out[3 * i + 2] = r[i].p[1];
  }
  return n;
}

/* BuildBinaryRadixTree_s (opencl/Kernels.cpp:266-273, kernel
 * C/BuildBRT.c:90-181) alone, on m sorted keys that may hold equal
 * neighbours (the reference returns a tree for them): nodes in gid order
 * into a zeroed array, so a child claimed twice keeps its later claimer. */
int oracle_brt_only(const uint64_t *lo, const uint64_t *hi, int64_t m, int mbits, int32_t *left,
                    uint8_t *left_leaf, uint8_t *right_leaf, uint64_t *lcp_lo, uint64_t *lcp_hi,
                    int32_t *lcp_length, int32_t *parent) {
  if (m < 2 || mbits < 1 || mbits > 128) return -1;
  oracle_tree t;
  memset(&t, 0, sizeof t);
  t.mbits = mbits;
  t.m = m;
  t.keys = (u128 *)zalloc(sizeof(u128) * (size_t)m);
  t.left = (int32_t *)zalloc(4 * (size_t)m);
  t.lcp_length = (int32_t *)zalloc(4 * (size_t)m);
  t.parent = (int32_t *)zalloc(4 * (size_t)m);
  t.left_leaf = (uint8_t *)zalloc((size_t)m);
  t.right_leaf = (uint8_t *)zalloc((size_t)m);
  t.lcp = (u128 *)zalloc(sizeof(u128) * (size_t)m);
  for (int64_t i = 0; i < m; ++i) t.keys[i] = mk(lo[i], hi ? hi[i] : 0);
  for (int gid = 0; gid < (int)m - 1; ++gid) build_brt_node(&t, (int)m, gid);
  for (int64_t i = 0; i < m - 1; ++i) {
    left[i] = t.left[i];
    left_leaf[i] = t.left_leaf[i];
    right_leaf[i] = t.right_leaf[i];
    lcp_lo[i] = (uint64_t)t.lcp[i];
    lcp_hi[i] = (uint64_t)(t.lcp[i] >> 64);
    lcp_length[i] = t.lcp_length[i];
    parent[i] = t.parent[i];
  }
  free(t.keys); free(t.left); free(t.lcp_length); free(t.parent);
  free(t.left_leaf); free(t.right_leaf); free(t.lcp);
  return 0;
}
