// query.hip -- batched octree point location: OctreeUtils::FindLeaf
// (OctreeUtils.cpp:18-71), one lane per query point, on an octree as the
// build writes it (nodes of 2^DIM children and a leaf mask).
//
// The key's digits are taken from the top, the octant masked to DIM bits
// (the reference leaves the mask commented out, :37-41, which is wrong below
// the first digit), and the descent stops at the first octant whose leaf bit
// is set.  Each step is one dependent read of a node; the top levels of the
// tree are shared by every query and stay in cache.
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

template <int DIM>
__global__ __launch_bounds__(kBlock) void find_leaf_kernel(
    const int32_t *__restrict__ pts, int64_t q, int bits, const int32_t *__restrict__ octree,
    int64_t octree_size, int32_t *__restrict__ parent, int32_t *__restrict__ octant,
    int32_t *__restrict__ data, int32_t *__restrict__ level, uint32_t *__restrict__ status) {
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
    for (int sh = mbits - DIM; sh >= 0; sh -= DIM) {
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

void launch_find_leaf(const int32_t *pts, int64_t q, int dim, int bits, const int32_t *octree,
                      int64_t octree_size, int32_t *parent, int32_t *octant, int32_t *data,
                      int32_t *level, uint32_t *status, hipStream_t s) {
  int64_t grid = (q + kBlock - 1) / kBlock;
  if (grid > 8192) grid = 8192;
  if (grid < 1) grid = 1;
  if (dim == 2)
    find_leaf_kernel<2><<<(unsigned)grid, kBlock, 0, s>>>(pts, q, bits, octree, octree_size,
                                                          parent, octant, data, level, status);
  else
    find_leaf_kernel<3><<<(unsigned)grid, kBlock, 0, s>>>(pts, q, bits, octree, octree_size,
                                                          parent, octant, data, level, status);
}

}  // namespace pgvd
