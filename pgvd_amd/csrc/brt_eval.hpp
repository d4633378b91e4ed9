// brt_eval.hpp -- one node of the Karras binary radix tree (BuildBinaryRadixTree,
// C/BuildBRT.c:90-181), shared by every BRT kernel.  Host-callable as well,
// so that the search logic itself is checked against the oracle on the CPU.
#pragma once

#include <cstdlib>

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
    // stage must take the order-exact scatter path.  The two searches step
    // side by side, so that their probes do not wait on each other.
    int se = 0;
    for (long long den = 2; den < 2ll * l; den *= 2) {
      const int t = fp32_split_step(l, (int)den);
      const int te = (int)(((long long)l + den - 1) / den);
      const bool up = s + t <= s_cutoff && above_node(s + t);
      const bool upe = se + te <= s_cutoff && above_node(se + te);
      if (up) s += t;
      if (upe) se += te;
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

// Largest x in [lo, hi] with pred(x), for a predicate that holds on a prefix
// of (lo, hi] (pred(lo) is taken as true): each round probes 7 evenly spaced
// offsets of the candidates at once (their loads are independent), then keeps
// the part after the last probe that holds.  ~log8(hi - lo) rounds.
template <typename Pred>
__host__ __device__ inline int kary_last_true(int lo, int hi, Pred pred) {
  while (hi > lo) {
    const int step = (hi - lo + 8) / 8;
    // every probe loads (a probe past hi reads hi and is dropped), so that
    // the seven loads are issued together rather than each behind a branch
    bool p[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      const int x = lo + (j + 1) * step;
      p[j] = pred(x < hi ? x : hi) & (x <= hi);
    }
    int best = lo;
#pragma unroll
    for (int j = 0; j < 7; ++j)
      if (p[j]) best = lo + (j + 1) * step;
    const int top = best + step - 1;
    hi = top < hi ? top : hi;
    lo = best;
  }
  return lo;
}

// kary_last_true(0, hi, pred) for a prefix that is short against hi (a
// node's range against the whole key array): one round at 8, 64, ..., 8^7
// brackets the answer by its power of 8, and the 8-ary search runs inside
// the bracket.
template <typename Pred>
__host__ __device__ inline int gallop_last_true(int hi, Pred pred) {
  bool p[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    const int x = 8 << (3 * j);
    p[j] = pred(x < hi ? x : hi) & (x <= hi);
  }
  int k = -1;
#pragma unroll
  for (int j = 0; j < 7; ++j)
    if (p[j]) k = j;
  const int lo = k >= 0 ? 8 << (3 * k) : 0;
  const int top = k < 6 ? (8 << (3 * (k + 1))) - 1 : hi;   // below the first probe that fails
  return kary_last_true(lo, top < hi ? top : hi, pred);
}

// The reference's split search (Karras' ceil-halving steps; the FP32 step
// of fp32_split_step above 2^23) replayed on a predicate that holds exactly
// on offsets [0, p]: no key is read.  *quirk: the FP32 search ends elsewhere
// than the exact one.  brt_eval's searches, on a monotone predicate whose
// last true offset is p, take the same steps.
__host__ __device__ inline int split_replay(int l, int s_cutoff, int p, bool *quirk) {
  int s = 0;
  *quirk = false;
  if (l <= (1 << 23)) {
    int k = 1;
    for (int den = 2; den < 2 * l; den <<= 1, ++k) {
      const int t = (l + den - 1) >> k;
      if (s + t <= s_cutoff && s + t <= p) s += t;
    }
  } else {
    int se = 0;
    for (long long den = 2; den < 2ll * l; den *= 2) {
      const int t = fp32_split_step(l, (int)den);
      const int te = (int)(((long long)l + den - 1) / den);
      if (s + t <= s_cutoff && s + t <= p) s += t;
      if (se + te <= s_cutoff && se + te <= p) se += te;
    }
    *quirk = se != s;
  }
  return s;
}

// brt_eval for the nodes whose ranges reach beyond the window kernels'
// windows (the overflow pass over global memory, where every probe of
// brt_eval's binary searches is a dependent load).  Both of its searches
// ask a monotone predicate (a key shares more than lcp bits with the
// node's key: true on a run of offsets from 0): the range end by a
// gallop and 8-ary search (seven independent loads a round), and the
// split by finding the predicate's last true offset the same way and
// replaying the reference's step sequence on it (split_replay), the FP32
// steps above 2^23 included.
template <typename K, typename Key>
__host__ __device__ inline BrtOut<K> brt_eval_wide(K cur, Key key, int mbits, int m, int i) {
  BrtOut<K> o;
  o.lcp_min = -1;
  const int l_pos = delta(cur, key(i + 1), mbits);
  o.dup = l_pos == mbits;
  int dn = 0, l, limit;
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
    l = sh < 0 ? 0 : gallop_last_true(limit, [&](int off) {
      return kzero(kshr(kxor(cur, key(i + ((off ^ dn) - dn))), sh));
    });
  }
  const int j = i + ((l ^ dn) - dn);
  const int lcp_node = delta(cur, key(j), mbits);
  const int shn = mbits - lcp_node - 1;
  // last offset in [0, l - 1] sharing more than lcp_node bits (0 when none
  // but the node's own key: the replay then never steps)
  const int p = (shn < 0 || l < 2) ? 0 : kary_last_true(0, l - 1, [&](int off) {
    return kzero(kshr(kxor(cur, key(i + ((off ^ dn) - dn))), shn));
  });
  const int s = split_replay(l, limit - 1, p, &o.quirk);
  const int split = i + ((s ^ dn) - dn) + dn;
  o.nd.left = split;
  o.nd.left_leaf = ((i < j ? i : j) == split);
  o.nd.right_leaf = ((i > j ? i : j) == split + 1);
  o.nd.lcp = kshr(cur, mbits - lcp_node);
  o.nd.lcp_length = (uint16_t)lcp_node;
  return o;
}

// ---- the window form: a sparse table of adjacent-key deltas ----------------
// For sorted distinct keys, delta(keys[i], keys[x]) is the minimum of the
// adjacent deltas D_k = delta(keys[k], keys[k+1]) for k between i and x, so
// a node's range end is the nearest adjacent delta <= its lcp_min and its
// split is the (unique) position of the range's minimum adjacent delta.  A
// window over global positions [w0, w0 + WN) holds, per level j < LV,
// mn(j, k) = 1 + min D over window positions [k, k + 2^j), and 0 (a
// sentinel below every delta) where that block reaches outside the window
// or outside [0, m-2].  Level j stores k = -2^j .. WN, so every probe of a
// descent is in range without a test:
//   byte offset of (j, k) = base(j) + k,  base(j) = 2^(j+1) - 1 + j (WN + 1),
//   table size = 2^LV - 1 + LV (WN + 1) bytes.
template <int WN, int LV>
struct DeltaTab {
  static_assert((1 << (LV - 1)) <= WN && WN < (1 << LV), "levels cover the window");
  static constexpr int kBytes = (1 << LV) - 1 + LV * (WN + 1);
  __host__ __device__ static constexpr int base(int j) { return (2 << j) - 1 + j * (WN + 1); }
  const uint8_t *t;
  __host__ __device__ int mn(int j, int k) const { return t[base(j) + k]; }
  // Smallest k >= pos with mn(0, k) <= v (fwd), or largest k <= pos with
  // mn(0, k) <= v (backward); WN resp. -1 if the window holds none.  Needs
  // -1 <= pos <= WN.  Greedy descent: a block of 2^j entries all above v is
  // skipped; blocks that reach outside the window read 0 and are not, so
  // pos stays in [-1, WN] and every probe in [-2^j, WN].
  __host__ __device__ int find(int pos, int v, bool fwd) const {
#pragma unroll
    for (int j = LV - 1; j >= 0; --j) {
      const int step = 1 << j;
      const int at = fwd ? pos : pos - step + 1;
#if !defined(__HIP_DEVICE_COMPILE__)
      if (at < -step || at > WN) std::abort();   // host builds check the probe bound
#endif
      const bool skip = t[base(j) + at] > v;
      pos += skip ? (fwd ? step : -step) : 0;
    }
    return pos;
  }
};

// BRT node i (0 <= i < m-1) from the window (w0 <= i-1, i < w0 + WN).
// Returns false when the node's range or split is not inside the window; the
// node then goes to brt_eval over the whole key array.  Well-formed trees
// only: ranges here are far below 2^23 keys.  Branch-free: every probe is
// made (clamped into the window) whatever the outcome, so that the searches
// of several nodes of one lane can overlap; o holds the node only when the
// result is true.
template <typename K, int WN, int LV>
__host__ __device__ inline bool brt_node_window(K cur, const DeltaTab<WN, LV> &tab, int w0,
                                                int mbits, int m, int i, BrtOut<K> &o) {
  const int wi = i - w0;
  const int dpos = tab.mn(0, wi) - 1;
  const int dneg = tab.mn(0, wi - 1) - 1;
  const bool root = i == 0;
  const bool fwd = dpos >= dneg;
  const int dmin = fwd ? dneg : dpos;
  const int k = tab.find(fwd ? wi : wi - 1, dmin + 1, fwd);
  bool ok = root || (k >= 0 && k < WN);
  const int lo = root ? 0 : (fwd ? i : w0 + k + 1);
  const int hi = root ? m - 1 : (fwd ? w0 + k : i);
  const int a = lo - w0, b = hi - 1 - w0;
  ok = ok && a >= 0 && b < WN && b >= a;
  const int ac = ok ? a : 0, bc = ok ? b : 0;
  const int jl = 31 - __builtin_clz((unsigned)(bc - ac + 1));
  const int lc = tab.mn(jl, ac), rc = tab.mn(jl, bc - (1 << jl) + 1);
  const int lcp_node = (lc < rc ? lc : rc) - 1;
  const int split = w0 + tab.find(ac, lcp_node + 1, true);
  // Holds for distinct keys; equal keys (reported through o.dup) may break
  // it, and then the bounded search over the whole array takes the node.
  ok = ok && hi > lo && split >= lo && split < hi;
  o.dup = dpos == mbits;
  o.quirk = false;
  o.lcp_min = root ? -1 : dmin;
  o.nd.left = split;
  o.nd.left_leaf = (lo == split);
  o.nd.right_leaf = (hi == split + 1);
  o.nd.lcp = kshr(cur, mbits - lcp_node);
  o.nd.lcp_length = (uint16_t)lcp_node;
  return ok;
}

}  // namespace pgvd
