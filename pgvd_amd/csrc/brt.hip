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
#include "kernels.hpp"

namespace pgvd {

namespace {

template <typename K>
__global__ __launch_bounds__(kBlock) void brt_kernel(const K *__restrict__ keys, int mbits,
                                                     BrtNodeT<K> *__restrict__ brt,
                                                     int32_t *__restrict__ parent, RunState *st,
                                                     int64_t m_host) {
  const int m = m_host >= 0 ? (int)m_host : (int)st->m;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m - 1) return;
  const K cur = keys[i];
  const K nxt = keys[i + 1];
  // The reference requires distinct keys (unique runs first); equal
  // neighbours are reported, and every search below stays range-bounded.
  if (keq(cur, nxt)) atomicOr(&st->status, (uint32_t)ST_DUP_KEY);
  int d, l;
  if (i == 0) {
    d = 1;
    l = m - 1;
  } else {
    const int l_pos = delta(cur, nxt, mbits);
    const int l_neg = delta(cur, keys[i - 1], mbits);
    d = l_pos >= l_neg ? 1 : -1;
    const int lcp_min = (d == -1) ? l_pos : l_neg;
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
  if (!nd.left_leaf) atomicMax(&parent[split], i + 1);
  if (!nd.right_leaf) atomicMax(&parent[split + 1], i + 1);
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

template void launch_brt<uint32_t>(const uint32_t *, int64_t, int64_t, int, BrtNodeT<uint32_t> *,
                                   int32_t *, RunState *, hipStream_t);
template void launch_brt<uint64_t>(const uint64_t *, int64_t, int64_t, int, BrtNodeT<uint64_t> *,
                                   int32_t *, RunState *, hipStream_t);
template void launch_brt<Key128>(const Key128 *, int64_t, int64_t, int, BrtNodeT<Key128> *,
                                 int32_t *, RunState *, hipStream_t);

}  // namespace pgvd
