// brt.hip -- Karras binary radix tree over sorted unique keys.
//
// Per internal node i (one lane each) the kernel follows
// BuildBinaryRadixTree (C/BuildBRT.c:90-181):
//   * direction d from delta(i,i+1) vs delta(i,i-1); node 0 spans [0, m-1];
//   * range end by exponential + binary search (the reference's l_max loop
//     tests a key one doubling stale, :118-126, which changes only how far
//     the doubling goes, never l);
//   * split by binary search whose step is the reference's FP32 expression
//     t = (int)((l + (float)den - 1) / den) (:151) -- for ranges longer than
//     2^23 keys that expression can exceed the exact ceil and select a
//     different split, which is reproduced;
//   * lcp = key >> (mbits - lcp_length) (compute_lcp, :34-52);
//   * parent links: the reference's serial loop lets the last (highest)
//     claiming node win (:169-179); here every claim is an atomicMax of
//     (claimer + 1) into a zeroed array, which gives the same winner.
// delta is mbits - bitlen(a ^ b), the value compute_lcp_length's shift
// search (:54-88) returns for every pair of distinct keys.
#include "tilescan.hpp"

namespace pgvd {

namespace {

enum { CTR_BRT = kMaxPasses + 3 };

// One BRT node.  Writes brt[i]; returns lcp_min (the smaller of the two
// neighbour deltas, which is the parent's lcp_length in a well-formed
// radix tree) through *lcp_min_out and the node's lcp_length as result.
// Claims are made by `claim(child, i)`.
template <typename K, typename Claim>
__device__ inline int brt_node(const K *__restrict__ keys, int mbits, int m, int i,
                               BrtNodeT<K> *__restrict__ brt, RunState *st, int *lcp_min_out,
                               Claim claim) {
  const K cur = keys[i];
  const K nxt = keys[i + 1];
  // The reference requires distinct keys (unique runs first); equal
  // neighbours are reported, and every search below stays range-bounded.
  if (keq(cur, nxt)) atomicOr(&st->status, (uint32_t)ST_DUP_KEY);
  int d, l;
  *lcp_min_out = -1;
  if (i == 0) {
    d = 1;
    l = m - 1;
  } else {
    const int l_pos = delta(cur, nxt, mbits);
    const int l_neg = delta(cur, keys[i - 1], mbits);
    d = l_pos >= l_neg ? 1 : -1;
    const int lcp_min = (d == -1) ? l_pos : l_neg;
    *lcp_min_out = lcp_min;
    int l_max = 2;
    for (;;) {
      const long long idx = (long long)i + (long long)l_max * d;
      if (idx < 0 || idx > m - 1) break;
      if (delta(cur, keys[idx], mbits) <= lcp_min) break;
      l_max <<= 1;
    }
    const int l_cutoff = (d == -1) ? i : m - i - 1;
    l = 0;
    for (int t = l_max >> 1; t >= 1; t >>= 1) {
      if (l + t <= l_cutoff && delta(cur, keys[i + (l + t) * d], mbits) > lcp_min) l += t;
    }
  }
  const int j = i + l * d;
  const int lcp_node = delta(cur, keys[j], mbits);
  const int s_cutoff = (d == -1) ? i - 1 : m - i - 2;
  int s = 0;
  for (int den = 2; den < 2 * l; den *= 2) {
    // (l + (float)den - 1) / den, every operation rounded as in C.
    const float num = __fadd_rn(__fadd_rn((float)l, (float)den), -1.0f);
    const int t = (int)__fdiv_rn(num, (float)den);
    if (s + t <= s_cutoff && delta(cur, keys[i + (s + t) * d], mbits) > lcp_node) s += t;
  }
  if (l > (1 << 23)) {
    // For l <= 2^23 every intermediate of the FP32 step is an integer
    // <= 2^24 (den < 2l is a power of two), so it equals the exact ceil.
    // Above that it can differ; if it changed the split, the tree is not a
    // well-formed radix tree and the octree stage must take the
    // order-exact scatter path.
    int se = 0;
    for (int den = 2; den < 2 * l; den *= 2) {
      const int t = (int)(((long long)l + den - 1) / den);
      if (se + t <= s_cutoff && delta(cur, keys[i + (se + t) * d], mbits) > lcp_node) se += t;
    }
    if (se != s) st->malformed = 1u;
  }
  const int split = i + s * d + (d < 0 ? d : 0);
  BrtNodeT<K> nd;
  nd.left = split;
  nd.left_leaf = (min(i, j) == split);
  nd.right_leaf = (max(i, j) == split + 1);
  nd.lcp = kshr(cur, mbits - lcp_node);
  nd.lcp_length = (uint16_t)lcp_node;
  brt[i] = nd;
  if (!nd.left_leaf) claim(split, i);
  if (!nd.right_leaf) claim(split + 1, i);
  return lcp_node;
}

// Stage form (BuildBinaryRadixTree_p): claims as atomicMax(claimer + 1) into
// a zeroed array, which gives the serial loop's last (highest) claimer.
template <typename K>
__global__ __launch_bounds__(kBlock) void brt_kernel(const K *__restrict__ keys, int mbits,
                                                     BrtNodeT<K> *__restrict__ brt,
                                                     int32_t *__restrict__ parent, RunState *st,
                                                     int64_t m_host) {
  const int m = m_host >= 0 ? (int)m_host : (int)st->m;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m - 1) return;
  int lcp_min;
  brt_node(keys, mbits, m, i, brt, st, &lcp_min,
           [&](int c, int p) { atomicMax(&parent[c], p + 1); });
}

// Pipeline form: BRT node, its local split and the inclusive scan of the
// local splits in one pass over striped tiles of kRows*256 nodes.
//  * In a well-formed radix tree the parent's lcp_length is the node's
//    lcp_min, so ls[c] = lcp_length(c)/DIM - lcp_min(c)/DIM needs no parent
//    lookup (C/BuildOctree.c:17-27 computes the same from the parent).
//  * Every non-root node has exactly one claimer, so parent links are
//    plain stores in the reference form (root -1).
// If some node took an FP32-quirk split (st->malformed), parent links, local
// splits and prefix sums are recomputed by the order-exact claim path.
template <typename K, int DIM>
__global__ __launch_bounds__(kBlock) void brt_scan_kernel(const K *__restrict__ keys, int mbits,
                                                          BrtNodeT<K> *__restrict__ brt,
                                                          int32_t *__restrict__ parent,
                                                          uint32_t *__restrict__ ps,
                                                          uint64_t *lookback, RunState *st) {
  constexpr int R = 8;
  constexpr int TILE = R * kBlock;
  __shared__ uint32_t s_w[R * 4];
  __shared__ uint64_t s_excl;
  __shared__ uint32_t s_tile;
  const int t = threadIdx.x;
  const int m = (int)st->m;
  const int ntiles = (m + TILE - 1) / TILE;
  if (t == 0) s_tile = atomicAdd(&st->counters[CTR_BRT], 1u);
  __syncthreads();
  const int tile = (int)s_tile;
  if (tile >= ntiles) return;
  const int base = tile * TILE;
  uint32_t v[R], ex[R];
#pragma unroll 1
  for (int r = 0; r < R; ++r) {
    const int c = base + r * kBlock + t;
    uint32_t ls = 0;
    if (c < m - 1) {
      int lcp_min;
      const int len = brt_node(keys, mbits, m, c, brt, st, &lcp_min,
                               [&](int child, int p) { parent[child] = p; });
      ls = (c == 0) ? 1u + (uint32_t)(len / DIM) : (uint32_t)(len / DIM - lcp_min / DIM);
      if (c == 0) parent[0] = -1;
    } else if (c == m - 1 && m == 1) {
      ls = 1;     // single key: one octree node (the reference reads I[0] of
                  // an empty vector here, opencl/Kernels.cpp:305-306)
    }
    v[r] = ls;
  }
  const uint32_t total = striped_excl_scan<R, uint32_t>(v, ex, s_w);
  const uint64_t excl = tile_prefix(lookback, tile, total, &s_excl);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = base + r * kBlock + t;
    if (c < m) ps[c] = (uint32_t)(excl + ex[r] + v[r]);
  }
  if (tile == ntiles - 1 && t == 0) st->octree_size = (uint32_t)(excl + total);
}

// Pipeline form, one node per lane: BRT node + its local split, parent links
// as plain stores (see brt_scan_kernel for why both are exact in a
// well-formed tree).  ls[m-1] = 0; m == 1 gives ls[0] = 1.
template <typename K, int DIM>
__global__ __launch_bounds__(kBlock) void brt_ls_kernel(const K *__restrict__ keys, int mbits,
                                                        BrtNodeT<K> *__restrict__ brt,
                                                        int32_t *__restrict__ parent,
                                                        uint32_t *__restrict__ ls, RunState *st) {
  const int m = (int)st->m;
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= m) return;
  if (c == m - 1) {
    ls[c] = (m == 1) ? 1u : 0u;
    return;
  }
  int lcp_min;
  const int len = brt_node(keys, mbits, m, c, brt, st, &lcp_min,
                           [&](int child, int p) { parent[child] = p; });
  if (c == 0) parent[0] = -1;
  ls[c] = (c == 0) ? 1u + (uint32_t)(len / DIM) : (uint32_t)(len / DIM - lcp_min / DIM);
}

// Order-exact recomputation of the parent claims when some split was an
// FP32-quirk split: zero the claim array, then atomicMax(claimer + 1).
__global__ __launch_bounds__(kBlock) void claims_zero_kernel(int32_t *parent, RunState *st) {
  if (!st->malformed) return;
  const int m = (int)st->m;
  for (int c = blockIdx.x * kBlock + threadIdx.x; c < m - 1; c += gridDim.x * kBlock) parent[c] = 0;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void claims_kernel(const BrtNodeT<K> *__restrict__ brt,
                                                        int32_t *parent, RunState *st) {
  if (!st->malformed) return;
  const int m = (int)st->m;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < m - 1; i += gridDim.x * kBlock) {
    const BrtNodeT<K> b = brt[i];
    if (!b.left_leaf) atomicMax(&parent[b.left], i + 1);
    if (!b.right_leaf) atomicMax(&parent[b.left + 1], i + 1);
  }
}

__global__ __launch_bounds__(kBlock) void finalize_parents_kernel(int32_t *parent, RunState *st,
                                                                  int64_t m_host) {
  const int m = m_host >= 0 ? (int)m_host : (int)st->m;
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= m - 1) return;
  const int32_t pc = parent[c];
  parent[c] = (c == 0) ? -1 : (pc ? pc - 1 : 0);
}

}  // namespace

template <typename K>
void launch_brt(const K *keys, int64_t max_m, int64_t m_host, int mbits, BrtNodeT<K> *brt,
                int32_t *parent, RunState *st, hipStream_t s) {
  const int64_t nodes = max_m > 1 ? max_m - 1 : 1;
  const unsigned grid = (unsigned)((nodes + kBlock - 1) / kBlock);
  brt_kernel<K><<<grid, kBlock, 0, s>>>(keys, mbits, brt, parent, st, m_host);
}

void launch_finalize_parents(int32_t *parent, int64_t max_m, int64_t m_host, RunState *st,
                             hipStream_t s) {
  const int64_t nodes = max_m > 1 ? max_m - 1 : 1;
  const unsigned grid = (unsigned)((nodes + kBlock - 1) / kBlock);
  finalize_parents_kernel<<<grid, kBlock, 0, s>>>(parent, st, m_host);
}

int64_t brt_scan_tiles(int64_t max_m) { return (max_m + 8 * kBlock - 1) / (8 * kBlock); }

template <typename K>
void launch_brt_scan(const K *keys, int64_t max_m, int mbits, int dim, BrtNodeT<K> *brt,
                     int32_t *parent, uint32_t *ps, uint64_t *lookback, RunState *st,
                     hipStream_t s) {
  const unsigned grid = (unsigned)brt_scan_tiles(max_m > 0 ? max_m : 1);
  if (dim == 2) brt_scan_kernel<K, 2><<<grid, kBlock, 0, s>>>(keys, mbits, brt, parent, ps, lookback, st);
  else brt_scan_kernel<K, 3><<<grid, kBlock, 0, s>>>(keys, mbits, brt, parent, ps, lookback, st);
}

template <typename K>
void launch_brt_ls(const K *keys, int64_t max_m, int mbits, int dim, BrtNodeT<K> *brt,
                   int32_t *parent, uint32_t *ls, RunState *st, hipStream_t s) {
  const unsigned grid = (unsigned)((max_m + kBlock - 1) / kBlock);
  if (dim == 2) brt_ls_kernel<K, 2><<<grid, kBlock, 0, s>>>(keys, mbits, brt, parent, ls, st);
  else brt_ls_kernel<K, 3><<<grid, kBlock, 0, s>>>(keys, mbits, brt, parent, ls, st);
}

template <typename K>
void launch_claims_if_malformed(const BrtNodeT<K> *brt, int32_t *parent, int64_t max_m,
                                RunState *st, hipStream_t s) {
  int64_t grid = (max_m + kBlock - 1) / kBlock;
  if (grid > 1024) grid = 1024;
  if (grid < 1) grid = 1;
  claims_zero_kernel<<<(unsigned)grid, kBlock, 0, s>>>(parent, st);
  claims_kernel<K><<<(unsigned)grid, kBlock, 0, s>>>(brt, parent, st);
}

#define PGVD_INST(K)                                                                          \
  template void launch_brt<K>(const K *, int64_t, int64_t, int, BrtNodeT<K> *, int32_t *,      \
                              RunState *, hipStream_t);                                        \
  template void launch_brt_scan<K>(const K *, int64_t, int, int, BrtNodeT<K> *, int32_t *,     \
                                   uint32_t *, uint64_t *, RunState *, hipStream_t);           \
  template void launch_claims_if_malformed<K>(const BrtNodeT<K> *, int32_t *, int64_t,         \
                                              RunState *, hipStream_t);                        \
  template void launch_brt_ls<K>(const K *, int64_t, int, int, BrtNodeT<K> *, int32_t *,       \
                                 uint32_t *, RunState *, hipStream_t);
PGVD_INST(uint32_t)
PGVD_INST(uint64_t)
PGVD_INST(Key128)
#undef PGVD_INST

}  // namespace pgvd
