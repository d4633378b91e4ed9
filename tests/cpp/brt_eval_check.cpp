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
    const BrtOut<K> w = brt_eval_wide(cur, [&](int x) { return keys[x]; }, mbits, (int)m, (int)i);
    if ((w.nd.left != o.nd.left || w.nd.left_leaf != o.nd.left_leaf ||
         w.nd.right_leaf != o.nd.right_leaf || w.nd.lcp_length != o.nd.lcp_length ||
         lo_of(w.nd.lcp) != lo_of(o.nd.lcp) || hi_of(w.nd.lcp) != hi_of(o.nd.lcp) ||
         w.lcp_min != o.lcp_min || w.dup != o.dup || w.quirk != o.quirk) &&
        bad++ < 5)
      std::printf("%s: k-ary node %lld differs: left %d/%d\n", name, (long long)i, w.nd.left,
                  o.nd.left);
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

// The searches both node forms rely on, over every prefix predicate of
// small ranges: kary_last_true's last true offset, and the Karras split
// search's exact ceil steps (brt_eval for l <= 2^23), which must find the
// same offset for a predicate that holds on [0, p] and fails at l.
int check_searches() {
  int bad = 0;
  for (int hi = 0; hi <= 1200; ++hi)
    for (int p = 0; p <= hi; ++p) {
      const int got = kary_last_true(0, hi, [&](int x) { return x <= p; });
      if (got != p && bad++ < 5) std::printf("kary_last_true(0, %d) = %d, want %d\n", hi, got, p);
    }
  {
    std::vector<int> his;
    for (int hi = 0; hi <= 1200; ++hi) his.push_back(hi);
    for (int e = 11; e <= 25; ++e) {
      his.push_back((1 << e) - 1);
      his.push_back(1 << e);
      his.push_back((1 << e) + 12345);
    }
    uint64_t x = 12345;
    for (int hi : his) {
      std::vector<int> ps;
      for (int p = 0; p <= hi && p <= 600; ++p) ps.push_back(p);
      for (int e = 0; e < 31 && (1 << e) <= hi; ++e) {
        ps.push_back((1 << e) - 1);
        ps.push_back(1 << e);
      }
      for (int r = 0; r < 64 && hi > 0; ++r) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        ps.push_back((int)(x % (uint64_t)(hi + 1)));
      }
      ps.push_back(hi);
      for (int p : ps) {
        if (p > hi) continue;
        const int got = gallop_last_true(hi, [&](int y) { return y <= p; });
        if (got != p && bad++ < 5)
          std::printf("gallop_last_true(%d) = %d, want %d\n", hi, got, p);
      }
    }
  }
  for (int l = 1; l <= 1200; ++l)
    for (int p = 0; p < l; ++p) {
      int s = 0, k = 1;
      for (int den = 2; den < 2 * l; den <<= 1, ++k) {
        const int t = (l + den - 1) >> k;
        if (s + t <= p) s += t;
      }
      if (s != p && bad++ < 5) std::printf("ceil split search l=%d gives %d, want %d\n", l, s, p);
    }
  // equal neighbouring keys: the k-ary form against brt_eval itself
  {
    std::vector<uint64_t> keys;
    uint64_t x = 0x9e3779b97f4a7c15ull, v = 0;
    for (int i = 0; i < 20000; ++i) {
      x ^= x << 13; x ^= x >> 7; x ^= x << 17;
      v += (x % 5 == 0) ? 0 : (x >> 40);     // runs of equal keys
      keys.push_back(v & ((1ull << 63) - 1));
    }
    std::sort(keys.begin(), keys.end());
    const int m = (int)keys.size();
    for (int i = 0; i + 1 < m; ++i) {
      auto key = [&](int y) { return keys[y]; };
      const BrtOut<uint64_t> a = brt_eval(keys[i], key, 63, m, i);
      const BrtOut<uint64_t> b = brt_eval_wide(keys[i], key, 63, m, i);
      if ((a.nd.left != b.nd.left || a.nd.lcp_length != b.nd.lcp_length ||
           a.nd.left_leaf != b.nd.left_leaf || a.nd.right_leaf != b.nd.right_leaf ||
           a.dup != b.dup || a.lcp_min != b.lcp_min) && bad++ < 5)
        std::printf("equal keys: node %d differs: left %d/%d\n", i, b.nd.left, a.nd.left);
    }
  }
  // a root range above 2^23: the FP32 split steps, replayed by the k-ary
  // form, against brt_eval's own search (random keys; the root and a
  // sample of the other nodes)
  {
    const int m = (1 << 23) + 4097;
    std::vector<uint64_t> keys(m);
    uint64_t x = 0x2545f4914f6cdd1dull;
    for (int i = 0; i < m; ++i) {
      x ^= x << 13; x ^= x >> 7; x ^= x << 17;
      keys[i] = x >> 1;
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const int mm = (int)keys.size();
    auto key = [&](int y) { return keys[y]; };
    int quirks = 0;
    for (int i = 0; i + 1 < mm; i = i < 4096 ? i + 1 : i + 257) {
      const BrtOut<uint64_t> a = brt_eval(keys[i], key, 63, mm, i);
      const BrtOut<uint64_t> b = brt_eval_wide(keys[i], key, 63, mm, i);
      quirks += a.quirk;
      if ((a.nd.left != b.nd.left || a.nd.lcp_length != b.nd.lcp_length ||
           a.nd.left_leaf != b.nd.left_leaf || a.nd.right_leaf != b.nd.right_leaf ||
           a.quirk != b.quirk || a.lcp_min != b.lcp_min) && bad++ < 5)
        std::printf("m=%d: node %d differs: left %d/%d quirk %d/%d\n", mm, i, b.nd.left,
                    a.nd.left, b.quirk, a.quirk);
    }
    std::printf("m=%d: %d quirk splits sampled\n", mm, quirks);
  }
  // a root whose FP32 split differs from the exact one (l = 2^24 - 1,
  // where l + den - 1 rounds in FP32): the first p + 1 keys below 2^62,
  // the rest above, for a p where the steps diverge
  {
    const int mm = 1 << 24, l = mm - 1;
    int pq = -1;
    for (int p = l / 2; p < l && pq < 0; ++p) {
      bool q;
      (void)split_replay(l, l - 1, p, &q);
      if (q) pq = p;
    }
    if (pq < 0) {
      std::printf("no diverging split for l=%d\n", l);
      ++bad;
    } else {
      std::vector<uint64_t> keys(mm);
      for (int i = 0; i < mm; ++i) keys[i] = (i <= pq ? 0ull : 1ull << 62) | (uint64_t)i;
      auto key = [&](int y) { return keys[y]; };
      const BrtOut<uint64_t> a = brt_eval(keys[0], key, 63, mm, 0);
      const BrtOut<uint64_t> b = brt_eval_wide(keys[0], key, 63, mm, 0);
      if (!a.quirk || a.nd.left != b.nd.left || a.quirk != b.quirk ||
          a.nd.lcp_length != b.nd.lcp_length) {
        std::printf("quirk root (p=%d): left %d/%d quirk %d/%d\n", pq, b.nd.left, a.nd.left,
                    (int)b.quirk, (int)a.quirk);
        ++bad;
      }
    }
  }
  // the replayed split search against the probing one, over every last-true
  // offset p of ranges l near and above 2^23 (the FP32 steps)
  {
    uint64_t x = 99;
    for (int r = 0; r < 4000; ++r) {
      x ^= x << 13; x ^= x >> 7; x ^= x << 17;
      const int l = (1 << 23) - 64 + (int)(x % (uint64_t)(3 << 23));
      const int p = (int)((x >> 32) % (uint64_t)l);
      const int cutoff = l + (int)((x >> 16) & 7) - 1;
      int s = 0, se = 0;
      if (l <= (1 << 23)) {
        int k = 1;
        for (int den = 2; den < 2 * l; den <<= 1, ++k) {
          const int t = (l + den - 1) >> k;
          if (s + t <= cutoff && s + t <= p) s += t;
        }
        se = s;
      } else {
        for (long long den = 2; den < 2ll * l; den *= 2) {
          const int t = fp32_split_step(l, (int)den);
          const int te = (int)(((long long)l + den - 1) / den);
          if (s + t <= cutoff && s + t <= p) s += t;
          if (se + te <= cutoff && se + te <= p) se += te;
        }
      }
      bool q;
      const int got = split_replay(l, cutoff, p, &q);
      if ((got != s || q != (se != s) || se != p) && bad++ < 5)
        std::printf("split_replay(l=%d, p=%d) = %d quirk %d; probing search %d exact %d\n", l,
                    p, got, (int)q, s, se);
    }
  }
  std::printf("searches: %s\n", bad ? "FAIL" : "ok");
  return bad;
}

int main() {
  int bad = check_searches();
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
