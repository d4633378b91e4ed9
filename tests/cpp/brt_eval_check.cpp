// CPU check of the node search every BRT kernel runs (brt_eval.hpp),
// compiled for the host: for sorted distinct keys of 32, 63 and 126 bits,
// uniform and clustered, every node computed from the whole key array equals
// the oracle's BuildBinaryRadixTree node, and every node that a window of
// the kernels' shape resolves -- from the keys themselves, or from the
// sparse table of adjacent deltas the window kernels use -- is that same node.
//
//     brt_eval_check        prints one line per case, exit status 0 if all match
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "brt_eval.hpp"

extern "C" {
#include "pgvd_oracle.h"
}

using namespace pgvd;

namespace {

template <typename K> K from_words(uint64_t lo, uint64_t hi);
template <> uint32_t from_words<uint32_t>(uint64_t lo, uint64_t) { return (uint32_t)lo; }
template <> uint64_t from_words<uint64_t>(uint64_t lo, uint64_t) { return lo; }
template <> Key128 from_words<Key128>(uint64_t lo, uint64_t hi) { return {lo, hi}; }

uint64_t lo_of(uint32_t k) { return k; }
uint64_t lo_of(uint64_t k) { return k; }
uint64_t lo_of(const Key128s &k) { return k.lo; }
uint64_t hi_of(uint32_t) { return 0; }
uint64_t hi_of(uint64_t) { return 0; }
uint64_t hi_of(const Key128s &k) { return k.hi; }

// m distinct keys below 2^mbits, sorted.  Clustered keys share long
// prefixes: a few random centres, each key a centre plus small offsets.
void make_keys(int mbits, int64_t m, uint64_t seed, bool clustered, std::vector<uint64_t> &lo,
               std::vector<uint64_t> &hi) {
  std::mt19937_64 rng(seed);
  auto mask = [&](uint64_t &l, uint64_t &h) {
    if (mbits < 64) { l &= (1ull << mbits) - 1; h = 0; }
    else if (mbits == 64) h = 0;
    else if (mbits < 128) h &= (1ull << (mbits - 64)) - 1;
  };
  std::vector<std::pair<uint64_t, uint64_t>> v;
  std::vector<std::pair<uint64_t, uint64_t>> centres;
  for (int c = 0; c < 6; ++c) centres.push_back({rng(), rng()});
  while ((int64_t)v.size() < m * 2) {
    uint64_t l = rng(), h = rng();
    if (clustered) {
      const auto &c = centres[rng() % centres.size()];
      const int spread = 8 + (int)(rng() % 20);
      h = c.second;
      l = (c.first & ~((1ull << spread) - 1)) | (rng() & ((1ull << spread) - 1));
      if (mbits <= 64) { l = (l >> (64 - mbits)) << 0; l = (c.first >> 1) ^ (rng() & ((1ull << spread) - 1)); }
    }
    mask(l, h);
    v.push_back({h, l});
  }
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  if ((int64_t)v.size() > m) v.resize(m);
  lo.resize(v.size());
  hi.resize(v.size());
  for (size_t i = 0; i < v.size(); ++i) { lo[i] = v[i].second; hi[i] = v[i].first; }
  // keep sorted after masking of the clustered low words
  std::vector<size_t> idx(v.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
    return hi[a] != hi[b] ? hi[a] < hi[b] : lo[a] < lo[b];
  });
  std::vector<uint64_t> l2, h2;
  for (size_t i : idx) {
    if (!l2.empty() && l2.back() == lo[i] && h2.back() == hi[i]) continue;
    l2.push_back(lo[i]);
    h2.push_back(hi[i]);
  }
  lo.swap(l2);
  hi.swap(h2);
}

// the window shape of the BRT kernels: 1024 nodes and 256 positions of
// halo either side
constexpr int WN = 1536, LV = 11;

template <typename K>
int check(const char *name, int mbits, int64_t want_m, uint64_t seed, bool clustered, int seg,
          int halo) {
  std::vector<uint64_t> lo, hi;
  make_keys(mbits, want_m, seed, clustered, lo, hi);
  const int64_t m = (int64_t)lo.size();
  std::vector<K> keys(m);
  for (int64_t i = 0; i < m; ++i) keys[i] = from_words<K>(lo[i], hi[i]);
  oracle_tree *t = oracle_build_from_keys(lo.data(), hi.data(), m, 3, mbits);
  if (!t) { std::printf("%s: oracle failed\n", name); return 1; }
  const int64_t n = m - 1;
  std::vector<int32_t> left(n), lcp_len(n), parent(n);
  std::vector<uint8_t> ll(n), rl(n);
  std::vector<uint64_t> clo(n), chi(n);
  oracle_tree_brt(t, left.data(), ll.data(), rl.data(), clo.data(), chi.data(), lcp_len.data(),
                  parent.data());
  oracle_tree_free(t);
  int bad = 0;
  std::vector<BrtOut<K>> full(n);
  for (int64_t i = 0; i < n; ++i) {
    const K cur = keys[i];
    full[i] = brt_eval(cur, [&](int x) { return keys[x]; }, mbits, (int)m, (int)i);
    const BrtOut<K> &o = full[i];
    const bool ok = o.nd.left == left[i] && o.nd.left_leaf == ll[i] && o.nd.right_leaf == rl[i] &&
                    o.nd.lcp_length == lcp_len[i] && lo_of(o.nd.lcp) == clo[i] &&
                    hi_of(o.nd.lcp) == chi[i] && !o.dup && !o.quirk;
    if (!ok && bad++ < 5)
      std::printf("%s: node %lld differs: left %d/%d leaves %d%d/%d%d lcp_length %d/%d\n", name,
                  (long long)i, o.nd.left, left[i], o.nd.left_leaf, o.nd.right_leaf, ll[i], rl[i],
                  o.nd.lcp_length, lcp_len[i]);
  }
  // the kernels' window: keys [b0 - halo, b0 + seg + halo) for nodes [b0, b0 + seg)
  int64_t misses = 0;
  for (int64_t b0 = 0; b0 < n; b0 += seg) {
    const int64_t w0 = b0 - halo, w1 = b0 + seg + halo;
    for (int64_t i = b0; i < std::min<int64_t>(b0 + seg, n); ++i) {
      const K cur = keys[i];
      const K far = kflip(cur, mbits - 1);
      bool miss = false;
      const BrtOut<K> o = brt_eval(
          cur,
          [&](int x) -> K {
            if (x < w0 || x >= w1) { miss = true; return far; }
            return keys[x];
          },
          mbits, (int)m, (int)i);
      if (miss) { ++misses; continue; }
      const BrtOut<K> &f = full[i];
      const bool ok = o.nd.left == f.nd.left && o.nd.left_leaf == f.nd.left_leaf &&
                      o.nd.right_leaf == f.nd.right_leaf && o.nd.lcp_length == f.nd.lcp_length &&
                      lo_of(o.nd.lcp) == lo_of(f.nd.lcp) && hi_of(o.nd.lcp) == hi_of(f.nd.lcp) &&
                      o.lcp_min == f.lcp_min;
      if (!ok && bad++ < 5) std::printf("%s: window node %lld differs\n", name, (long long)i);
    }
  }
  // the kernels' delta-table window: nodes [b0, b0 + seg), adjacent deltas of
  // positions [b0 - halo, b0 + seg + halo)
  int64_t rmq_misses = 0;
  {
    // the kernels' layout (DeltaTab), filled here from its definition: entry
    // (j, k) for k = -2^j .. WN is 1 + min D over [k, k + 2^j) when that
    // block lies in [0, WN) and in [0, m-2], else 0
    constexpr int wn = WN, levels = LV;
    using Tab = DeltaTab<WN, LV>;
    std::vector<uint8_t> tab(Tab::kBytes, 0xEE);
    for (int64_t b0 = 0; b0 < n; b0 += seg) {
      const int64_t w0 = b0 - halo;
      auto d1 = [&](int k) -> int {     // level-0 value of window position k
        const int64_t g = w0 + k;
        return (k >= 0 && k < wn && g >= 0 && g < m - 1) ? 1 + delta(keys[g], keys[g + 1], mbits)
                                                         : 0;
      };
      // min over [k, k + 2^j) as the min of its two halves
      std::vector<int> cur(wn + 1), nxt(wn + 1);
      for (int k = 0; k <= wn; ++k) cur[k] = d1(k);
      for (int j = 0; j < levels; ++j) {
        for (int k = -(1 << j); k <= wn; ++k) {
          const bool inside = k >= 0 && k + (1 << j) <= wn;
          tab[(size_t)(Tab::base(j) + k)] = (uint8_t)(inside ? cur[k] : 0);
        }
        for (int k = 0; k <= wn; ++k)
          nxt[k] = k + (2 << j) <= wn ? std::min(cur[k], cur[k + (1 << j)]) : 0;
        cur.swap(nxt);
      }
      const Tab view{tab.data()};
      for (int64_t i = b0; i < std::min<int64_t>(b0 + seg, n); ++i) {
        BrtOut<K> o;
        if (!brt_node_window(keys[i], view, (int)w0, mbits, (int)m, (int)i, o)) {
          ++rmq_misses;
          continue;
        }
        const BrtOut<K> &f = full[i];
        const bool ok = o.nd.left == f.nd.left && o.nd.left_leaf == f.nd.left_leaf &&
                        o.nd.right_leaf == f.nd.right_leaf &&
                        o.nd.lcp_length == f.nd.lcp_length && lo_of(o.nd.lcp) == lo_of(f.nd.lcp) &&
                        hi_of(o.nd.lcp) == hi_of(f.nd.lcp) && o.lcp_min == f.lcp_min && !o.dup;
        if (!ok && bad++ < 5)
          std::printf("%s: table node %lld differs: left %d/%d lcp_length %d/%d lcp_min %d/%d\n",
                      name, (long long)i, o.nd.left, f.nd.left, o.nd.lcp_length,
                      f.nd.lcp_length, o.lcp_min, f.lcp_min);
      }
    }
  }
  std::printf("%s: m=%lld nodes=%lld window misses=%lld (%.2f%%) table misses=%lld (%.2f%%) %s\n",
              name, (long long)m, (long long)n, (long long)misses, 100.0 * misses / (n ? n : 1),
              (long long)rmq_misses, 100.0 * rmq_misses / (n ? n : 1), bad ? "FAIL" : "ok");
  return bad;
}

}  // namespace

int main() {
  int bad = 0;
  for (bool cl : {false, true}) {
    const char *tag = cl ? "clustered" : "uniform";
    char name[64];
    std::snprintf(name, sizeof name, "32-bit %s", tag);
    bad += check<uint32_t>(name, 32, 200000, 7, cl, 1024, 256);
    std::snprintf(name, sizeof name, "63-bit %s", tag);
    bad += check<uint64_t>(name, 63, 200000, 8, cl, 1024, 256);
    std::snprintf(name, sizeof name, "126-bit %s", tag);
    bad += check<Key128>(name, 126, 100000, 9, cl, 1024, 256);
  }
  for (int64_t m : {2, 3, 4, 5, 9, 17, 100}) {
    char name[64];
    std::snprintf(name, sizeof name, "63-bit m=%lld", (long long)m);
    bad += check<uint64_t>(name, 63, m, 10 + m, false, 1024, 256);
  }
  std::printf(bad ? "MISMATCHES\n" : "ALL MATCH\n");
  return bad ? 1 : 0;
}
