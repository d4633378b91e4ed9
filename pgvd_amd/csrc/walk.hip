// walk.hip -- batched cell walks: CellWalk (viewer/Octree2.cpp:265-313) for
// many 2D segments at once, one lane per segment.  Each step locates the
// current point's leaf (FindLeaf with the octant masked to two bits), visits
// it, intersects the segment with the cell's four edges
// (OctreeUtils::FindIntersections, OctreeUtils.cpp:80-172) and moves to the
// farthest intersection.
//
// The walk's length is not known in advance, so a batch runs it twice: a
// counting pass, a scan of the counts to offsets (decoupled look-back), and
// a pass that writes every visited cell at its segment's offset.  The float
// arithmetic follows the reference operation by operation (no contraction,
// correctly rounded division and square root, epsilon tests in double), so
// both passes and the serial restatement visit the same cells.
#include "kernels.hpp"
#include "tilescan.hpp"
#include "walk_geom.hpp"

namespace pgvd {

namespace {

constexpr int kWalkMax = 10000;        // the reference throws past this many cells

// FindLeaf of the truncated point (convert_int2); cell = origin x, origin
// y, width, parent, octant.
__device__ inline void walk_leaf(float px, float py, int bits, const int32_t *__restrict__ octree,
                                 int64_t size, int32_t cell[5]) {
  const int32_t qx = (int32_t)px, qy = (int32_t)py;
  const int64_t c[2] = {qx, qy};
  uint64_t lo, hi;
  morton128<2>(c, bits, lo, hi);
  int64_t idx = 0;
  uint32_t o = 0;
  int32_t lev = 0;
  for (int sh = 2 * bits - 2; sh >= 0; sh -= 2) {
    o = (uint32_t)(lo >> sh) & 3u;
    ++lev;
    const int32_t *node = octree + idx * 5;
    const int32_t child = node[o];
    if ((node[4] >> o) & 1) break;
    if (child < 0 || child >= size) break;
    idx = child;
  }
  const int32_t w = (int32_t)(1u << (bits - lev));
  const int32_t gmask = (int32_t)((1u << bits) - 1);
  cell[0] = (qx & gmask) & ~(w - 1);
  cell[1] = (qy & gmask) & ~(w - 1);
  cell[2] = w;
  cell[3] = (int32_t)idx;
  cell[4] = (int32_t)o;
}

constexpr int kMaxBits = 30;

// walk_leaf from the deepest node the point shares with the lane's previous
// descent: that descent's key is pkey, its leaf level plev (0: none) and its
// nodes by depth are in path[d * kBlock].  The top digits the two keys share
// select the same nodes and the same octants, none of them a leaf above
// plev, so the result is walk_leaf's.  Updates pkey, plev and the path.
__device__ inline void walk_leaf_path(float px, float py, int bits,
                                      const int32_t *__restrict__ octree, int64_t size,
                                      int32_t *path, uint64_t &pkey, int &plev, int32_t cell[5]) {
  const int32_t qx = (int32_t)px, qy = (int32_t)py;
  const int64_t c[2] = {qx, qy};
  uint64_t lo, hi;
  morton128<2>(c, bits, lo, hi);
  const uint64_t key = lo & ((1ull << (2 * bits)) - 1);
  int d = 0;
  if (plev > 0) {
    const uint64_t x = key ^ pkey;
    const int common = x ? (__clzll((long long)x) - (64 - 2 * bits)) >> 1 : bits;
    d = common < plev - 1 ? common : plev - 1;
  }
  int64_t idx = path[d * kBlock];
  uint32_t o = 0;
  int32_t lev = d;
  for (int sh = 2 * (bits - d) - 2; sh >= 0; sh -= 2) {
    o = (uint32_t)(lo >> sh) & 3u;
    ++lev;
    const int32_t *node = octree + idx * 5;
    const int32_t child = node[o];
    if ((node[4] >> o) & 1) break;
    if (child < 0 || child >= size) break;
    idx = child;
    path[lev * kBlock] = (int32_t)idx;
  }
  pkey = key;
  plev = lev;
  const int32_t w = (int32_t)(1u << (bits - lev));
  const int32_t gmask = (int32_t)((1u << bits) - 1);
  cell[0] = (qx & gmask) & ~(w - 1);
  cell[1] = (qy & gmask) & ~(w - 1);
  cell[2] = w;
  cell[3] = (int32_t)idx;
  cell[4] = (int32_t)o;
}

// One pass over the segments.  MODE kWalkCount: counts and status;
// kWalkWrite: the cells at offsets[i]; kWalkSpec: counts and status, and
// the first K cells of segment i at scratch + 5 * K * i; kWalkRest: for the
// segments with more than K cells, cells K.. at offsets[i] + K.
enum { kWalkCount = 0, kWalkWrite = 1, kWalkSpec = 2, kWalkRest = 3 };

template <int MODE>
__global__ __launch_bounds__(kBlock) void cell_walk_kernel(
    const float *__restrict__ segs, int64_t n, int bits, const int32_t *__restrict__ octree,
    int64_t size, uint32_t *__restrict__ counts, int32_t *__restrict__ status,
    const int64_t *__restrict__ offsets, int32_t *__restrict__ cells, int K) {
  // each lane's last descent, node per depth (depth 0 is the root)
  __shared__ int32_t s_path[(kMaxBits + 1) * kBlock];
  int32_t *path = s_path + threadIdx.x;
  path[0] = 0;
  for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * kBlock) {
    if (MODE == kWalkRest && counts[s] <= (uint32_t)K) continue;
    const float ax = segs[4 * s], ay = segs[4 * s + 1], bx = segs[4 * s + 2], by = segs[4 * s + 3];
    const int dx = bx > ax ? 1 : (bx < ax ? -1 : 0);
    const int dy = by > ay ? 1 : (by < ay ? -1 : 0);
    float ct = 0.0f, cx = ax, cy = ay;
    int32_t bcell[5], cell[5], ncell[5];
    uint64_t pkey = 0;
    int plev = 0;
    walk_leaf(bx, by, bits, octree, size, bcell);
    walk_leaf_path(cx, cy, bits, octree, size, path, pkey, plev, cell);
    int32_t *out = MODE == kWalkWrite || MODE == kWalkRest ? cells + 5 * offsets[s]
                   : MODE == kWalkSpec                     ? cells + 5 * (int64_t)K * s
                                                           : nullptr;
    int count = 0, st = 0;
    bool done = false;
    do {
      if (++count > kWalkMax) {
        st = 1;
        --count;
        break;
      }
      const bool put = MODE == kWalkWrite || (MODE == kWalkSpec && count <= K) ||
                       (MODE == kWalkRest && count > K);
      if (put) {
#pragma unroll
        for (int k = 0; k < 5; ++k) out[5 * (count - 1) + k] = cell[k];
      }
      float ts[4], px[4], py[4];
      const uint32_t keep = walk_isects(ax, ay, bx, by, cell[0], cell[1], cell[2], ts, px, py);
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (((keep >> i) & 1) && ts[i] > ct) {
          ct = ts[i];
          cx = px[i];
          cy = py[i];
        }
      if (cx == (float)cell[0] && cx != 0.0f && dx == -1) cx -= 1.0f;
      if (cy == (float)cell[1] && cy != 0.0f && dy == -1) cy -= 1.0f;
      walk_leaf_path(cx, cy, bits, octree, size, path, pkey, plev, ncell);
      done = keep == 0 || (cell[0] == bcell[0] && cell[1] == bcell[1]) ||
             (ncell[0] == cell[0] && ncell[1] == cell[1]);
#pragma unroll
      for (int k = 0; k < 5; ++k) cell[k] = ncell[k];
    } while (!done);
    if (MODE == kWalkCount || MODE == kWalkSpec) {
      counts[s] = (uint32_t)count;
      if (status) status[s] = st;
    }
  }
}

// Moves the speculative pass's first min(count, K) cells of every segment
// to their offsets; one thread per (segment, slot), K a power of two.
__global__ __launch_bounds__(kBlock) void walk_compact_kernel(
    const int32_t *__restrict__ scratch, const uint32_t *__restrict__ counts,
    const int64_t *__restrict__ offsets, int64_t n, int log2k, int32_t *__restrict__ cells) {
  const int64_t slots = n << log2k;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < slots;
       i += (int64_t)gridDim.x * kBlock) {
    const int64_t s = i >> log2k;
    const uint32_t k = (uint32_t)(i & ((int64_t(1) << log2k) - 1));
    if (k < counts[s]) {
      const int32_t *src = scratch + 5 * i;
      int32_t *dst = cells + 5 * (offsets[s] + k);
#pragma unroll
      for (int j = 0; j < 5; ++j) dst[j] = src[j];
    }
  }
}

// offsets[0] = 0, offsets[i + 1] = counts[0] + ... + counts[i]; 1024 counts
// per tile; ws: word 0 tile counter, status array from word 4.
__global__ __launch_bounds__(kBlock) void walk_scan_kernel(const uint32_t *__restrict__ counts,
                                                           int64_t n,
                                                           int64_t *__restrict__ offsets,
                                                           uint64_t *__restrict__ ws) {
  __shared__ uint32_t s_w[kBlock / 64];
  __shared__ int s_tile;
  const int t = threadIdx.x;
  if (t == 0) s_tile = (int)atomicAdd(reinterpret_cast<unsigned int *>(ws), 1u);
  __syncthreads();
  const int tile = s_tile;
  const int64_t i0 = ((int64_t)tile * kBlock + t) * 4;
  uint32_t c[4], sum = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    c[k] = i0 + k < n ? counts[i0 + k] : 0u;
    sum += c[k];
  }
  uint32_t total;
  const uint32_t ex = block_excl_sum<uint32_t>(sum, s_w, &total);
  uint64_t run = tile_prefix<kBlock>(ws + 4, tile, total) + ex;
  if (tile == 0 && t == 0) offsets[0] = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    run += c[k];
    if (i0 + k < n) offsets[i0 + k + 1] = (int64_t)run;
  }
}

__global__ void walk_ws_zero_kernel(uint64_t *ws, int64_t words) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < words;
       i += (int64_t)gridDim.x * kBlock)
    ws[i] = 0;
}

}  // namespace

int64_t walk_scan_ws_words(int64_t n) { return 4 + (n + 4 * kBlock - 1) / (4 * kBlock); }

void launch_cell_walk_count(const float *segs, int64_t n, int bits, const int32_t *octree,
                            int64_t size, uint32_t *counts, int32_t *status, int64_t *offsets,
                            uint64_t *ws, int32_t *scratch, int log2k, hipStream_t s) {
  int64_t grid = (n + kBlock - 1) / kBlock;
  if (grid > 8192) grid = 8192;
  if (scratch)
    cell_walk_kernel<kWalkSpec><<<(unsigned)grid, kBlock, 0, s>>>(
        segs, n, bits, octree, size, counts, status, nullptr, scratch, 1 << log2k);
  else
    cell_walk_kernel<kWalkCount><<<(unsigned)grid, kBlock, 0, s>>>(
        segs, n, bits, octree, size, counts, status, nullptr, nullptr, 0);
  const int64_t words = walk_scan_ws_words(n);
  walk_ws_zero_kernel<<<(unsigned)std::min<int64_t>((words + kBlock - 1) / kBlock, 1024), kBlock,
                        0, s>>>(ws, words);
  const int64_t tiles = (n + 4 * kBlock - 1) / (4 * kBlock);
  walk_scan_kernel<<<(unsigned)tiles, kBlock, 0, s>>>(counts, n, offsets, ws);
}

void launch_offsets_scan(const uint32_t *counts, int64_t n, int64_t *offsets, uint64_t *ws,
                         hipStream_t s) {
  const int64_t words = walk_scan_ws_words(n);
  walk_ws_zero_kernel<<<(unsigned)std::min<int64_t>((words + kBlock - 1) / kBlock, 1024), kBlock,
                        0, s>>>(ws, words);
  const int64_t tiles = (n + 4 * kBlock - 1) / (4 * kBlock);
  if (tiles > 0) walk_scan_kernel<<<(unsigned)tiles, kBlock, 0, s>>>(counts, n, offsets, ws);
}

void launch_cell_walk_write(const float *segs, int64_t n, int bits, const int32_t *octree,
                            int64_t size, const int64_t *offsets, int32_t *cells,
                            const int32_t *scratch, uint32_t *counts, int log2k, hipStream_t s) {
  int64_t grid = (n + kBlock - 1) / kBlock;
  if (grid > 8192) grid = 8192;
  if (!scratch) {
    cell_walk_kernel<kWalkWrite><<<(unsigned)grid, kBlock, 0, s>>>(
        segs, n, bits, octree, size, nullptr, nullptr, offsets, cells, 0);
    return;
  }
  int64_t cgrid = ((n << log2k) + kBlock - 1) / kBlock;
  if (cgrid > 16384) cgrid = 16384;
  walk_compact_kernel<<<(unsigned)cgrid, kBlock, 0, s>>>(scratch, counts, offsets, n, log2k,
                                                          cells);
  cell_walk_kernel<kWalkRest><<<(unsigned)grid, kBlock, 0, s>>>(
      segs, n, bits, octree, size, counts, nullptr, offsets, cells, 1 << log2k);
}

}  // namespace pgvd
