// query.hip -- batched octree point location: OctreeUtils::FindLeaf
// (OctreeUtils.cpp:18-71), one lane per query point, on an octree as the
// build writes it (nodes of 2^DIM children and a leaf mask).
//
// The key's digits are taken from the top, the octant masked to DIM bits
// (the reference leaves the mask commented out, :37-41, which is wrong below
// the first digit), and the descent stops at the first octant whose leaf bit
// is set.  Each step is two dependent loads from a node (child and leaf
// mask), and below the top levels every lane's load falls on its own line.
//
// Large batches first build a table of the octree's cells at a fixed depth D
// (2^21 cells in 3D, 2^20 in 2D, 8 bytes each: 16 MB, which the Infinity
// Cache holds; 2^18 cells, which L2 holds, left more of the descent's node
// lines to fetch and was slower, and 2^24 cells cost more to build than they
// saved at 2^24 queries): entry e is where the
// descent stands after the D digits e, either the node to continue from or,
// when it ended within them, its parent, octant, level and status.  A query
// then starts with one table read instead of D levels of node reads; its
// outputs are the same as those of the plain descent.
#include "kernels.hpp"

namespace pgvd {

namespace {

// Digit (DIM bits) of the 128-bit key (lo, hi) at bit sh.
template <int DIM>
__device__ inline uint32_t key_digit(uint64_t lo, uint64_t hi, int sh) {
  uint64_t v;
  if (sh >= 64) v = hi >> (sh - 64);
  else if (sh == 0) v = lo;
  else v = (lo >> sh) | (hi << (64 - sh));
  return (uint32_t)v & ((1u << DIM) - 1);
}

// A table entry whose descent ended within the table's depth: meta holds
// level (bits 0-7), octant (8-11), status (12-15) and this flag.
constexpr int32_t kTabDone = 1 << 16;

// Builds the depth-D table: one lane per cell e (its D digits, most
// significant first), the same descent as find_leaf_kernel stopped after D
// levels.  entry.x = node to continue from, or the parent where it ended.
template <int DIM>
__global__ __launch_bounds__(kBlock) void find_leaf_table_kernel(
    const int32_t *__restrict__ octree, int64_t octree_size, int D, int2 *__restrict__ table) {
  constexpr int NC = 1 << DIM, STRIDE = NC + 1;
  const int64_t cells = int64_t(1) << (D * DIM);
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < cells;
       e += (int64_t)gridDim.x * kBlock) {
    int64_t idx = 0;
    int32_t meta = 0;
    for (int l = 1; l <= D; ++l) {
      const uint32_t o = (uint32_t)(e >> ((D - l) * DIM)) & (NC - 1);
      const int32_t *node = octree + idx * STRIDE;
      const int32_t child = node[o];
      if ((node[NC] >> o) & 1) {
        meta = kTabDone | (int32_t)(o << 8) | l;
        break;
      }
      if (child < 0 || child >= octree_size) {
        meta = kTabDone | (2 << 12) | (int32_t)(o << 8) | l;
        break;
      }
      idx = child;
    }
    table[e] = make_int2((int32_t)idx, meta);
  }
}

// The top D digits of the key (lo, hi), D * DIM <= 21, starting at bit sh.
__device__ inline uint32_t key_bits(uint64_t lo, uint64_t hi, int sh, int width) {
  uint64_t v;
  if (sh >= 64) v = hi >> (sh - 64);
  else if (sh == 0) v = lo;
  else v = (lo >> sh) | (hi << (64 - sh));
  return (uint32_t)v & ((1u << width) - 1);
}

template <int DIM, bool TABLE>
__global__ __launch_bounds__(kBlock) void find_leaf_kernel(
    const int32_t *__restrict__ pts, int64_t q, int bits, const int32_t *__restrict__ octree,
    int64_t octree_size, const int2 *__restrict__ table, int D, int32_t *__restrict__ parent,
    int32_t *__restrict__ octant, int32_t *__restrict__ data, int32_t *__restrict__ level,
    uint32_t *__restrict__ status) {
  constexpr int NC = 1 << DIM, STRIDE = NC + 1;
  const int mbits = bits * DIM;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < q;
       i += (int64_t)gridDim.x * kBlock) {
    int64_t c[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) c[k] = pts[i * DIM + k];
    uint64_t lo, hi;
    morton128<DIM>(c, bits, lo, hi);
    int64_t idx = 0;
    uint32_t o = 0, st = 1;
    int32_t d = -1, lev = 0;
    int sh0 = mbits - DIM;
    if (TABLE) {
      const int top = mbits - D * DIM;
      const int2 t = table[key_bits(lo, hi, top, D * DIM)];
      idx = t.x;
      if (t.y & kTabDone) {
        lev = t.y & 0xff;
        o = (uint32_t)(t.y >> 8) & 0xf;
        st = (uint32_t)(t.y >> 12) & 0xf;
        if (st == 0) d = octree[idx * STRIDE + o];
        sh0 = -1;
      } else {
        lev = D;
        o = key_digit<DIM>(lo, hi, top);
        sh0 = top - DIM;
      }
    }
    for (int sh = sh0; sh >= 0; sh -= DIM) {
      o = key_digit<DIM>(lo, hi, sh);
      ++lev;
      const int32_t *node = octree + idx * STRIDE;
      const int32_t child = node[o];
      if ((node[NC] >> o) & 1) {
        d = child;
        st = 0;
        break;
      }
      if (child < 0 || child >= octree_size) {   // not an octree the build wrote
        st = 2;
        break;
      }
      idx = child;
    }
    parent[i] = (int32_t)idx;
    octant[i] = (int32_t)o;
    data[i] = d;
    level[i] = lev;
    if (status) status[i] = st;
  }
}

}  // namespace

int find_leaf_table_depth(int64_t q, int dim, int bits) {
  int D = kFindLeafTableBits / dim;
  if (D > bits) D = bits;
  return q >= (int64_t(2) << (D * dim)) ? D : 0;
}

void launch_find_leaf(const int32_t *pts, int64_t q, int dim, int bits, const int32_t *octree,
                      int64_t octree_size, void *table, int D, int32_t *parent, int32_t *octant,
                      int32_t *data, int32_t *level, uint32_t *status, hipStream_t s) {
  int64_t grid = (q + kBlock - 1) / kBlock;
  if (grid > 8192) grid = 8192;
  if (grid < 1) grid = 1;
  int2 *tab = static_cast<int2 *>(table);
  if (D > 0) {
    const unsigned tgrid = (unsigned)(((int64_t(1) << (D * dim)) + kBlock - 1) / kBlock);
    if (dim == 2)
      find_leaf_table_kernel<2><<<tgrid, kBlock, 0, s>>>(octree, octree_size, D, tab);
    else
      find_leaf_table_kernel<3><<<tgrid, kBlock, 0, s>>>(octree, octree_size, D, tab);
  }
#define PGVD_FIND_LEAF(DIMV, TAB)                                                              \
  find_leaf_kernel<DIMV, TAB><<<(unsigned)grid, kBlock, 0, s>>>(                               \
      pts, q, bits, octree, octree_size, tab, D, parent, octant, data, level, status)
  if (dim == 2) {
    if (D > 0) PGVD_FIND_LEAF(2, true);
    else PGVD_FIND_LEAF(2, false);
  } else {
    if (D > 0) PGVD_FIND_LEAF(3, true);
    else PGVD_FIND_LEAF(3, false);
  }
#undef PGVD_FIND_LEAF
}

}  // namespace pgvd
