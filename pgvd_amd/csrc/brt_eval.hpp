// brt_eval.hpp -- one node of the Karras binary radix tree (BuildBinaryRadixTree,
// C/BuildBRT.c:90-181), shared by every BRT kernel.  Host-callable as well,
// so that the search logic itself is checked against the oracle on the CPU.
#pragma once

#include "kernels.hpp"

namespace pgvd {

// The reference's split step t = (int)((l + (float)den - 1) / den)
// (C/BuildBRT.c:151), every operation rounded as in C.
__host__ __device__ inline int fp32_split_step(int l, int den) {
#if defined(__HIP_DEVICE_COMPILE__)
  const float num = __fadd_rn(__fadd_rn((float)l, (float)den), -1.0f);
  return (int)__fdiv_rn(num, (float)den);
#else
  volatile float a = (float)l + (float)den;   // no contraction or reassociation
  volatile float num = a - 1.0f;
  return (int)(num / (float)den);
#endif
}

// Result of one BRT node.
template <typename K>
struct BrtOut {
  BrtNodeT<K> nd;
  int lcp_min;    // smaller neighbour delta (-1 for node 0)
  bool dup;       // keys[i] == keys[i+1]
  bool quirk;     // the FP32 split step chose a split the exact ceil does not
};

// One BRT node from a key oracle: key(x) = keys[x] for 0 <= x < m.  An
// oracle that cannot supply a key may answer kflip(keys[i], mbits - 1), whose
// delta to keys[i] is 0, so every search stays range-bounded; the caller then
// discards the result.  The searches test "delta(keys[i], k) > T", i.e.
// bitlen(keys[i] ^ k) < mbits - T, as (keys[i] ^ k) >> (mbits - T - 1) == 0
// (never true for T >= mbits), and step by +-offset as
// i + ((off ^ dn) - dn) with dn = 0 (d = +1) or -1 (d = -1).
template <typename K, typename Key>
__host__ __device__ inline BrtOut<K> brt_eval(K cur, Key key, int mbits, int m, int i) {
  BrtOut<K> o;
  o.quirk = false;
  o.lcp_min = -1;
  const int l_pos = delta(cur, key(i + 1), mbits);
  // The reference requires distinct keys (unique runs first); equal
  // neighbours (delta == mbits) are reported, and every search below stays
  // range-bounded.
  o.dup = l_pos == mbits;
  int dn = 0, l, limit;     // limit: largest offset that stays inside [0, m-1]
  if (i == 0) {
    l = m - 1;
    limit = m - 1;
  } else {
    const int l_neg = delta(cur, key(i - 1), mbits);
    dn = l_pos >= l_neg ? 0 : -1;
    const int lcp_min = dn ? l_pos : l_neg;
    o.lcp_min = lcp_min;
    limit = dn ? i : m - 1 - i;
    const int sh = mbits - lcp_min - 1;
    auto above = [&](int off) {
      return sh >= 0 && kzero(kshr(kxor(cur, key(i + ((off ^ dn) - dn))), sh));
    };
    int l_max = 2;
    while (l_max <= limit && above(l_max)) l_max <<= 1;
    l = 0;
    for (int t = l_max >> 1; t >= 1; t >>= 1)
      if (l + t <= limit && above(l + t)) l += t;
  }
  const int j = i + ((l ^ dn) - dn);
  const int lcp_node = delta(cur, key(j), mbits);
  const int s_cutoff = limit - 1;
  const int shn = mbits - lcp_node - 1;
  auto above_node = [&](int off) {
    return shn >= 0 && kzero(kshr(kxor(cur, key(i + ((off ^ dn) - dn))), shn));
  };
  int s = 0;
  if (l <= (1 << 23)) {
    // The reference's step t = (int)((l + (float)den - 1) / den): for
    // l <= 2^23 every intermediate is an integer <= 2^24 and den a power of
    // two, so t is exactly ceil(l / den).
    int k = 1;
    for (int den = 2; den < 2 * l; den <<= 1, ++k) {
      const int t = (l + den - 1) >> k;
      if (s + t <= s_cutoff && above_node(s + t)) s += t;
    }
  } else {
    // Above 2^23 the FP32 step can exceed the exact ceil; if that changes
    // the split, the tree is not a well-formed radix tree and the octree
    // stage must take the order-exact scatter path.
    for (long long den = 2; den < 2ll * l; den *= 2) {
      const int t = fp32_split_step(l, (int)den);
      if (s + t <= s_cutoff && above_node(s + t)) s += t;
    }
    int se = 0;
    for (long long den = 2; den < 2ll * l; den *= 2) {
      const int t = (int)(((long long)l + den - 1) / den);
      if (se + t <= s_cutoff && above_node(se + t)) se += t;
    }
    o.quirk = se != s;
  }
  const int split = i + ((s ^ dn) - dn) + dn;
  o.nd.left = split;
  o.nd.left_leaf = ((i < j ? i : j) == split);
  o.nd.right_leaf = ((i > j ? i : j) == split + 1);
  o.nd.lcp = kshr(cur, mbits - lcp_node);
  o.nd.lcp_length = (uint16_t)lcp_node;
  return o;
}

}  // namespace pgvd
