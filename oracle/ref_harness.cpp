// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Links the reference's own serial C sources (C/BigUnsigned.c,
// C/BuildBRT.c, C/BuildOctree.c, C/ParallelAlgorithms.c, compiled
// unmodified from /root/reference by oracle/Makefile) into
// oracle/_ref/libpgvdref{2,3}.so, and restates the ~25 host lines of
// opencl/Kernels.cpp (PointsToMorton_s :62-75, BuildBinaryRadixTree_s
// :266-273, ComputeLocalSplits_s :303-313, BinaryRadixToOctree_s :370-386,
// BuildOctree_s :388-411) that cannot be compiled here (they include
// CL/cl.hpp and clfw.hpp).  The library is a checker for the oracle
// restatement; nothing in the product loads it.
//
// It also restates the reference's (commented-out) stream input of an
// octree, C/OctNode.h:69-98, on std::istringstream, so the text parser's
// token rules are checked against the standard library's operator>>.
//
// The 2D build (-DOCT2D) also exposes the reference's xyz2z
// (C/z_order.h:14-47).  The 3D build has no Morton entry point: z_order.h
// includes vec_cl.h, which forces OCT2D (C/vec_cl.h:17).
#include <sys/mman.h>

#include <algorithm>
#include <sstream>
#include <string>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {
#include "BigUnsigned.h"
#include "BrtNode.h"
#include "BuildBRT.h"
#include "BuildOctree.h"
#include "OctNode.h"
#include "ParallelAlgorithms.h"
#ifdef OCT2D
#include "z_order.h"
#endif
}

namespace {

// opencl/Kernels.cpp:20
int nextPow2(int num) { return std::max((int)pow(2, ceil(log(num) / log(2))), 8); }

void to_bu(uint64_t lo, uint64_t hi, BigUnsigned *b) {
  initBU(b);
  for (int i = 0; i < BIG_INTEGER_SIZE; ++i) {
    const int bit = 8 * i;
    const uint64_t byte = bit < 64 ? (lo >> bit) & 0xff : (hi >> (bit - 64)) & 0xff;
    b->blk[i] = (Blk)byte;
  }
  b->len = BIG_INTEGER_SIZE;
  zapLeadingZeros(b);
}

void from_bu(const BigUnsigned &b, uint64_t *lo, uint64_t *hi) {
  uint64_t l = 0, h = 0;
  for (int i = 0; i < b.len; ++i) {
    const uint64_t byte = b.blk[i];
    const int bit = 8 * i;
    if (bit < 64) l |= byte << bit;
    else h |= byte << (bit - 64);
  }
  *lo = l;
  *hi = h;
}

// The BRT kernel reads mpoints[gid + 2*d] for gid = 1, d = -1 with an
// unsigned index (C/BuildBRT.c:90,119), i.e. element 2^32-1 (64 GiB past
// the array).  The value is never used.  Keys are therefore placed at
// the start of a large MAP_NORESERVE mapping, so that read lands on an
// untouched (zero) page instead of faulting.
struct KeyArena {
  void *base = nullptr;
  size_t bytes = 0;
  explicit KeyArena(size_t m) {
    bytes = ((size_t)0xFFFFFFFFu + 2 + m) * sizeof(BigUnsigned) + 4096;
    base = mmap(nullptr, bytes, PROT_READ | PROT_WRITE,
                MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (base == MAP_FAILED) base = nullptr;
  }
  ~KeyArena() {
    if (base) munmap(base, bytes);
  }
  BigUnsigned *keys() { return static_cast<BigUnsigned *>(base); }
};

struct RefTree {
  int dim = 0;
  std::vector<BigUnsigned> keys;
  std::vector<BrtNode> I;
  std::vector<unsigned int> ls, ps;
  std::vector<OctNode> octree;
  int status = 0;
};

// BuildBinaryRadixTree_s + BinaryRadixToOctree_s on sorted unique keys.
void build_from_sorted(RefTree *t, int mbits) {
  const int numPoints = (int)t->keys.size();
  t->I.assign(numPoints - 1, BrtNode());
  {
    KeyArena arena(numPoints);
    if (!arena.base) { t->status = -2; return; }
    std::memcpy(arena.keys(), t->keys.data(), sizeof(BigUnsigned) * numPoints);
    for (int i = 0; i < numPoints - 1; ++i)
      BuildBinaryRadixTree(t->I.data(), arena.keys(), mbits, numPoints, i);
  }
  const int size = numPoints;
  t->ls.assign(size, 0);
  if (size > 1) {
    t->ls[0] = 1 + t->I[0].lcp_length / DIM;
    for (int i = 0; i < size - 1; ++i) ComputeLocalSplits(t->ls.data(), t->I.data(), i);
  } else {
    t->ls[0] = 1;   // the reference reads I[0] of an empty vector here
  }
  t->ps.assign(size, 0);
  StreamScan_SerialKernel(t->ls.data(), t->ps.data(), size);
  const int octreeSize = t->ps[size - 1];
  t->octree.resize(octreeSize);
  for (int i = 0; i < octreeSize; ++i) brt2octree_init(i, t->octree.data());
  for (int brt_i = 1; brt_i < size - 1; ++brt_i)
    brt2octree(brt_i, t->I.data(), t->octree.data(), t->ls.data(), t->ps.data(), size, octreeSize);
}

bool bu_greater(BigUnsigned a, BigUnsigned b) { return weakCompareBU(a, b) != 0; }
bool bu_equal(BigUnsigned a, BigUnsigned b) { return weakEqualsBU(a, b); }

}  // namespace

extern "C" {

int ref_dim(void) { return DIM; }
int ref_big_integer_size(void) { return BIG_INTEGER_SIZE; }

int ref_compute_lcp_length(uint64_t alo, uint64_t ahi, uint64_t blo, uint64_t bhi, int mbits) {
  BigUnsigned a, b;
  to_bu(alo, ahi, &a);
  to_bu(blo, bhi, &b);
  return compute_lcp_length(&a, &b, mbits);
}

void ref_compute_lcp(uint64_t lo, uint64_t hi, int length, int mbits, uint64_t *olo, uint64_t *ohi) {
  BigUnsigned v, l;
  to_bu(lo, hi, &v);
  initBU(&l);
  compute_lcp(&l, &v, length, mbits);
  from_bu(l, olo, ohi);
}

void ref_stream_scan(const unsigned int *in, unsigned int *out, int size) {
  StreamScan_SerialKernel(const_cast<unsigned int *>(in), out, size);
}

int ref_next_pow2(int n) { return nextPow2(n); }

#ifdef OCT2D
void ref_xyz2z(int x, int y, int bits, uint64_t *lo, uint64_t *hi) {
  BigUnsigned r;
  intn p;
  p.s[0] = x;
  p.s[1] = y;
  xyz2z(&r, p, bits);
  from_bu(r, lo, hi);
}

// opencl/Kernels.cpp:388-411 (BuildOctree_s) from integer 2D points.
void *ref_build_octree(const int *pts_xy, int n, int bits, int mbits) {
  RefTree *t = new RefTree();
  t->dim = DIM;
  const int roundNumPoints = nextPow2(n);
  std::vector<BigUnsigned> zpoints(roundNumPoints);
  for (int gid = 0; gid < roundNumPoints; ++gid) {   // PointsToMorton_s
    if (gid < n) {
      intn p;
      p.s[0] = pts_xy[2 * gid];
      p.s[1] = pts_xy[2 * gid + 1];
      xyz2z(&zpoints[gid], p, bits);
    } else {
      initBlkBU(&zpoints[gid], 0);
    }
  }
  std::sort(zpoints.rbegin(), zpoints.rend(), bu_greater);
  const int numPoints = std::unique(zpoints.begin(), zpoints.end(), bu_equal) - zpoints.begin();
  zpoints.resize(numPoints);
  t->keys = zpoints;
  build_from_sorted(t, mbits);
  return t;
}
#endif

// Sorted unique keys (<= 88 bits) -> BRT, local splits, scan, octree.
void *ref_build_from_keys(const uint64_t *lo, const uint64_t *hi, int m, int mbits) {
  RefTree *t = new RefTree();
  t->dim = DIM;
  t->keys.resize(m);
  for (int i = 0; i < m; ++i) to_bu(lo[i], hi ? hi[i] : 0, &t->keys[i]);
  build_from_sorted(t, mbits);
  return t;
}

// BuildBinaryRadixTree for every gid of m sorted keys (<= 88 bits; equal
// neighbours allowed) into a zeroed BrtNode array, nothing else.  With three
// equal keys at the end, node m-2 writes the parent field of node m-1, one
// past the reference's m-1 nodes; the array has a spare node for it.
int ref_brt_only(const uint64_t *lo, const uint64_t *hi, int m, int mbits, int32_t *left,
                 uint8_t *left_leaf, uint8_t *right_leaf, uint64_t *lcp_lo, uint64_t *lcp_hi,
                 int32_t *lcp_length, int32_t *parent) {
  std::vector<BrtNode> I((size_t)m);
  std::memset(I.data(), 0, sizeof(BrtNode) * I.size());
  KeyArena arena((size_t)m);
  if (!arena.base) return -2;
  for (int i = 0; i < m; ++i) to_bu(lo[i], hi ? hi[i] : 0, &arena.keys()[i]);
  for (int i = 0; i < m - 1; ++i) BuildBinaryRadixTree(I.data(), arena.keys(), mbits, m, i);
  for (int i = 0; i < m - 1; ++i) {
    left[i] = I[i].left;
    left_leaf[i] = I[i].left_leaf;
    right_leaf[i] = I[i].right_leaf;
    from_bu(I[i].lcp, &lcp_lo[i], &lcp_hi[i]);
    lcp_length[i] = I[i].lcp_length;
    parent[i] = I[i].parent;
  }
  return 0;
}

int ref_tree_status(void *h) { return static_cast<RefTree *>(h)->status; }
int ref_tree_m(void *h) { return (int)static_cast<RefTree *>(h)->keys.size(); }
int ref_tree_octree_size(void *h) { return (int)static_cast<RefTree *>(h)->octree.size(); }

void ref_tree_keys(void *h, uint64_t *lo, uint64_t *hi) {
  RefTree *t = static_cast<RefTree *>(h);
  for (size_t i = 0; i < t->keys.size(); ++i) from_bu(t->keys[i], &lo[i], &hi[i]);
}

void ref_tree_brt(void *h, int32_t *left, uint8_t *left_leaf, uint8_t *right_leaf,
                  uint64_t *lcp_lo, uint64_t *lcp_hi, int32_t *lcp_length, int32_t *parent) {
  RefTree *t = static_cast<RefTree *>(h);
  for (size_t i = 0; i < t->I.size(); ++i) {
    const BrtNode &b = t->I[i];
    left[i] = b.left;
    left_leaf[i] = b.left_leaf;
    right_leaf[i] = b.right_leaf;
    from_bu(b.lcp, &lcp_lo[i], &lcp_hi[i]);
    lcp_length[i] = b.lcp_length;
    parent[i] = b.parent;
  }
}

void ref_tree_splits(void *h, uint32_t *ls, uint32_t *ps) {
  RefTree *t = static_cast<RefTree *>(h);
  std::memcpy(ls, t->ls.data(), sizeof(uint32_t) * t->ls.size());
  std::memcpy(ps, t->ps.data(), sizeof(uint32_t) * t->ps.size());
}

// children[(1<<DIM)] and leaf per node.  In 3D the reference's `leaf`
// reads leaf_masks[4..7] out of bounds (C/OctNode.h:7); compare children
// only there.
void ref_tree_octree(void *h, int32_t *children, int32_t *leaf) {
  RefTree *t = static_cast<RefTree *>(h);
  const int nc = 1 << DIM;
  for (size_t i = 0; i < t->octree.size(); ++i) {
    for (int c = 0; c < nc; ++c) children[i * nc + c] = t->octree[i].children[c];
    leaf[i] = t->octree[i].leaf;
  }
}

void ref_tree_free(void *h) { delete static_cast<RefTree *>(h); }

// The reference's own records, byte for byte: its BigUnsigned keys (16
// bytes each) and BrtNodes, and their layout (sizes and field offsets) as
// its headers define it.
void ref_tree_keys_raw(void *h, void *out) {
  RefTree *t = static_cast<RefTree *>(h);
  std::memcpy(out, t->keys.data(), sizeof(BigUnsigned) * t->keys.size());
}
void ref_tree_brt_raw(void *h, void *out) {
  RefTree *t = static_cast<RefTree *>(h);
  std::memcpy(out, t->I.data(), sizeof(BrtNode) * t->I.size());
}
// sizeof(BigUnsigned), offsetof len/blk, sizeof(BrtNode), offsetof left,
// left_leaf, right_leaf, lcp, lcp_length, parent
void ref_layout(int32_t *out) {
  out[0] = (int32_t)sizeof(BigUnsigned);
  out[1] = (int32_t)offsetof(BigUnsigned, len);
  out[2] = (int32_t)offsetof(BigUnsigned, blk);
  out[3] = (int32_t)sizeof(BrtNode);
  out[4] = (int32_t)offsetof(BrtNode, left);
  out[5] = (int32_t)offsetof(BrtNode, left_leaf);
  out[6] = (int32_t)offsetof(BrtNode, right_leaf);
  out[7] = (int32_t)offsetof(BrtNode, lcp);
  out[8] = (int32_t)offsetof(BrtNode, lcp_length);
  out[9] = (int32_t)offsetof(BrtNode, parent);
}
// A value built the way the reference's tests build random BigUnsigneds
// (tests/KernelTests.cpp:106-112: setBUBit per bit, then len and
// zapLeadingZeros), for nbits bits of (lo, hi).
void ref_bu_from_bits(uint64_t lo, uint64_t hi, int nbits, void *out) {
  BigUnsigned b;
  std::memset(&b, 0, sizeof(b));
  initBU(&b);
  for (int j = 0; j < nbits; ++j) {
    const uint64_t bit = j < 64 ? (lo >> j) & 1 : (hi >> (j - 64)) & 1;
    setBUBit(&b, j, bit != 0);
  }
  b.len = nbits / 8;
  zapLeadingZeros(&b);
  std::memcpy(out, &b, sizeof(b));
}
int ref_compare_bu(const void *x, const void *y) {
  BigUnsigned a, b;
  std::memcpy(&a, x, sizeof(a));
  std::memcpy(&b, y, sizeof(b));
  return compareBU(&a, &b);
}

// operator>> for std::vector<OctNode> (C/OctNode.h:84-95) and OctNode
// (:69-76) on the text: 0 and the nodes as (children, leaf) int32 records;
// -1 when the stream fails (failbit) or the count is negative (the
// vector's resize throws there); -2 when the count exceeds cap.
int ref_istream_octree(const char *text, int64_t len, int32_t *out, int64_t cap, int64_t *O_out) {
  std::istringstream in(std::string(text, (size_t)len));
  int n;
  in >> n;
  if (!in || n < 0) return -1;
  *O_out = n;
  if (n > cap) return -2;
  std::vector<OctNode> octree((size_t)n);
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < (1 << DIM); ++k) in >> octree[i].children[k];
    int leaf;
    in >> leaf;
    octree[i].leaf = static_cast<unsigned char>(leaf);
  }
  if (!in) return -1;
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < (1 << DIM); ++k) out[(size_t)i * ((1 << DIM) + 1) + k] = octree[i].children[k];
    out[(size_t)i * ((1 << DIM) + 1) + (1 << DIM)] = octree[i].leaf;
  }
  return 0;
}

}  // extern "C"
