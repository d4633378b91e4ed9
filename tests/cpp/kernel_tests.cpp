// kernel_tests.cpp -- the reference's stage tests (tests/KernelTests.cpp)
// restated against the C++ host layer (include/pgvd_kernels.hpp).  Where
// the reference compares each `_p` stage with its serial `_s` twin, this
// compares with the oracle (oracle/pgvd_oracle.h).  Inputs are seeded
// (the reference used srand(time(NULL))).  Prints one line per scenario and
// exits non-zero on the first failure.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "pgvd_kernels.hpp"
#include "pgvd_oracle.h"

using namespace pgvd;

static int failures = 0;
#define CHECK(cond, what)                                                   \
  do {                                                                      \
    if (!(cond)) {                                                          \
      std::printf("FAIL %s (%s:%d)\n", what, __FILE__, __LINE__);           \
      ++failures;                                                           \
      return;                                                               \
    }                                                                       \
  } while (0)

static const int kBits = 30;          // tests/KernelTests.cpp:8
static const int kMbits = kBits * 2;  // :9 (DIM 2)

// "Points can be mapped to a Z-Order curve" (:47-92): (i, i), i < 10^6.
static void morton_scenario() {
  const int n = 1000000;
  std::vector<int2> pts(n);
  for (int i = 0; i < n; ++i) pts[i] = {i, i};
  Buffer p, z;
  CHECK(Kernels::UploadPoints(pts, p) == 0, "upload");
  CHECK(Kernels::PointsToMorton_p(p, z, n, kBits) == 0, "PointsToMorton_p");
  const int g = Kernels::nextPow2(n);
  std::vector<uint64_t> lo, hi;
  CHECK(Kernels::ReadKeys(z, g, lo, hi) == 0, "read keys");
  for (int i = 0; i < g; ++i) {
    uint64_t elo = 0, ehi = 0;
    if (i < n) {
      const int64_t c[2] = {i, i};
      oracle_xyz2z(c, 2, kBits, &elo, &ehi);
    }
    CHECK(lo[i] == elo && hi[i] == ehi, "Morton keys equal the serial xyz2z (padding zero)");
  }
  std::printf("PASS Points can be mapped to a Z-Order curve\n");
}

static std::vector<uint64_t> random_keys(int n, int mbits, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::vector<uint64_t> k(n);
  for (auto &x : k) x = rng() & ((mbits >= 64) ? ~0ull : ((1ull << mbits) - 1));
  return k;
}

static Buffer keys_buffer(const std::vector<uint64_t> &keys, int global) {
  Context &c = Context::Default();
  std::vector<uint64_t> host(keys);
  host.resize(global, 0);                 // zero-key padding
  Buffer b = c.alloc(sizeof(uint64_t) * global);
  c.upload(b, host.data(), sizeof(uint64_t) * global);
  b.key_bytes = 8;
  b.dim = 2;
  return b;
}

// "Big Unsigneds can be sorted using a parallel radix sort" (:96-146)
static void sort_scenario() {
  const int n = 1000000, g = Kernels::nextPow2(n);
  auto keys = random_keys(n, kMbits, 1);
  Buffer b = keys_buffer(keys, g);
  CHECK(Kernels::RadixSortBigUnsigned(b, n, 88) == 0, "RadixSortBigUnsigned");
  std::vector<uint64_t> lo, hi;
  Kernels::ReadKeys(b, g, lo, hi);
  keys.resize(g, 0);
  std::sort(keys.begin(), keys.end());
  CHECK(lo == keys, "sorted keys equal std::sort");
  std::printf("PASS Big Unsigneds can be sorted using a parallel radix sort\n");
}

// The reference's own bit-serial sort (opencl/Kernels.cpp:219-247) driven
// through the elementary stages: BitPredicate, StreamScan_p, DoubleCompact.
static void bit_serial_scenario() {
  const int n = 50000, g = Kernels::nextPow2(n), mbits = 20;
  auto keys = random_keys(n, mbits, 2);
  Buffer in = keys_buffer(keys, g), pred, addr, tmp;
  for (unsigned index = 0; index < (unsigned)mbits; ++index) {
    CHECK(Kernels::BitPredicate(in, pred, index, 0, g) == 0, "BitPredicate");
    CHECK(Kernels::StreamScan_p(pred, addr, g) == 0, "StreamScan_p");
    CHECK(Kernels::DoubleCompact(in, tmp, pred, addr, g) == 0, "DoubleCompact");
    std::swap(in, tmp);
  }
  std::vector<uint64_t> lo, hi;
  Kernels::ReadKeys(in, g, lo, hi);
  keys.resize(g, 0);
  std::sort(keys.begin(), keys.end());
  CHECK(lo == keys, "bit-serial split sort equals std::sort");
  // UniquePredicate + StreamScan_p + SingleCompact = UniqueSorted's steps
  Buffer up, ua, uo;
  CHECK(Kernels::UniquePredicate(in, up, g) == 0, "UniquePredicate");
  CHECK(Kernels::StreamScan_p(up, ua, g) == 0, "StreamScan_p");
  CHECK(Kernels::SingleCompact(in, uo, up, ua, g) == 0, "SingleCompact");
  const int m = (int)(std::unique(keys.begin(), keys.end()) - keys.begin());
  Kernels::ReadKeys(uo, m, lo, hi);
  CHECK(std::equal(lo.begin(), lo.end(), keys.begin()), "single compaction equals std::unique");
  std::printf("PASS the bit-serial split sort and compaction stages\n");
}

// "Sorted BigUnsigneds can be unique'd in parallel" (:148-207)
static void unique_scenario() {
  const int n = 1000000, g = Kernels::nextPow2(n);
  auto keys = random_keys(n, 24, 3);        // many duplicates
  keys.resize(g, 0);
  std::sort(keys.begin(), keys.end());
  Buffer b = keys_buffer(keys, g);
  int size = n;
  CHECK(Kernels::UniqueSorted(b, size) == 0, "UniqueSorted");
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  CHECK(size == (int)keys.size(), "unique count");
  std::vector<uint64_t> lo, hi;
  Kernels::ReadKeys(b, size, lo, hi);
  CHECK(lo == keys, "unique keys equal std::unique");
  std::printf("PASS Sorted BigUnsigneds can be unique'd in parallel\n");
}

// "binary radix tree" (:209-277), "local splits" (:279-315), "octree" (:317-350)
static void tree_scenarios() {
  auto keys = random_keys(1000000, kMbits, 4);
  keys.push_back(0);
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  const int m = (int)keys.size();
  Buffer z = keys_buffer(keys, m), I, ls;
  CHECK(Kernels::BuildBinaryRadixTree_p(z, I, m, kMbits) == 0, "BuildBinaryRadixTree_p");
  oracle_tree *t = oracle_build_from_keys(keys.data(), nullptr, m, 2, kMbits);
  std::vector<int32_t> left(m - 1), len(m - 1), par(m - 1);
  std::vector<uint8_t> ll(m - 1), rl(m - 1);
  std::vector<uint64_t> lcp_lo(m - 1), lcp_hi(m - 1);
  oracle_tree_brt(t, left.data(), ll.data(), rl.data(), lcp_lo.data(), lcp_hi.data(), len.data(), par.data());
  std::vector<BrtNode> nodes;
  Kernels::ReadBrt(I, m, nodes);
  bool ok = true;
  for (int i = 0; i < m - 1 && ok; ++i)
    ok = nodes[i].left == left[i] && nodes[i].left_leaf == (bool)ll[i] &&
         nodes[i].right_leaf == (bool)rl[i] && nodes[i].lcp_lo == lcp_lo[i] &&
         nodes[i].lcp_hi == lcp_hi[i] && nodes[i].lcp_length == len[i] && nodes[i].parent == par[i];
  CHECK(ok, "BRT nodes equal the serial BRT");
  std::printf("PASS Sorted Z-Order numbers can be used to construct a binary radix tree\n");

  CHECK(Kernels::ComputeLocalSplits_p(I, ls, m) == 0, "ComputeLocalSplits_p");
  std::vector<uint32_t> gls(m), ols(m), ops(m);
  Context::Default().download(gls.data(), ls, sizeof(uint32_t) * m);
  oracle_tree_splits(t, ols.data(), ops.data());
  CHECK(gls == ols, "local splits equal the serial local splits");
  std::printf("PASS A binary radix tree can be used to calculate a vector of local splits\n");

  std::vector<OctNode2> oct;
  CHECK(Kernels::BinaryRadixToOctree_p(I, oct, m) == 0, "BinaryRadixToOctree_p");
  const int64_t O = oracle_tree_octree_size(t);
  CHECK((int64_t)oct.size() == O, "octree size");
  std::vector<int32_t> ch(O * 4), lf(O);
  oracle_tree_octree(t, ch.data(), lf.data());
  for (int64_t i = 0; i < O && ok; ++i) {
    for (int c = 0; c < 4; ++c) ok &= oct[i].children[c] == ch[i * 4 + c];
    ok &= oct[i].leaf == lf[i];
  }
  CHECK(ok, "octree equals the serial octree");
  oracle_tree_free(t);
  std::printf("PASS A binary radix tree can be used to construct a Octree/Quadtree\n");
}

// The same scenarios read back the way the reference's tests read their
// buffers: as BigUnsigned records (tests/KernelTests.cpp:74,130,190) and
// BrtNodes (:209-277), compared with compareBU (C/BigUnsigned.c:132-153)
// and compareBrtNode's fields (C/BrtNode.h:23-31).  The sort takes
// BIG_INTEGER_SIZE * 8 = 88-bit values built bit by bit (:106-112).
static BigUnsigned bu_of(uint64_t lo, uint64_t hi) {
  BigUnsigned b;
  uint64_t w[2] = {lo, hi};
  pgvd_keys_to_big_unsigned(w, 16, 1, &b);
  return b;
}

static void big_unsigned_scenarios() {
  {   // Morton keys of (i, i), 30 bits per axis
    const int n = 1000000, g = Kernels::nextPow2(n);
    std::vector<int2> pts(n);
    for (int i = 0; i < n; ++i) pts[i] = {i, i};
    Buffer p, z;
    CHECK(Kernels::UploadPoints(pts, p) == 0, "upload");
    CHECK(Kernels::PointsToMorton_p(p, z, n, kBits) == 0, "PointsToMorton_p");
    std::vector<BigUnsigned> gpu;
    CHECK(Kernels::ReadBigUnsigned(z, g, gpu) == 0, "ReadBigUnsigned");
    int cmp = 0;
    for (int i = 0; i < g && cmp == 0; ++i) {
      uint64_t lo = 0, hi = 0;
      if (i < n) {
        const int64_t c[2] = {i, i};
        oracle_xyz2z(c, 2, kBits, &lo, &hi);
      }
      const BigUnsigned want = bu_of(lo, hi);
      cmp = CompareBU(want, gpu[i]);
    }
    CHECK(cmp == 0, "Morton keys read as BigUnsigned compare equal (compareBU)");
  }
  {   // radix sort of 88-bit BigUnsigneds, then unique
    const int n = 1 << 20;
    std::mt19937_64 rng(9);
    std::vector<BigUnsigned> host(n);
    for (auto &b : host) {
      const uint64_t lo = rng(), hi = rng() & ((1ull << 24) - 1);
      b = bu_of(lo, (rng() & 7) ? hi : 0);   // some shorter values
    }
    for (int i = 0; i < n; i += 97) host[i] = host[(i * 31) % n];   // equal values
    Buffer b;
    CHECK(Kernels::WriteBigUnsigned(host, 16, b) == 0, "WriteBigUnsigned");
    b.dim = 2;
    CHECK(Kernels::RadixSortBigUnsigned(b, n, 88) == 0, "RadixSortBigUnsigned (88 bits)");
    std::vector<BigUnsigned> gpu;
    CHECK(Kernels::ReadBigUnsigned(b, n, gpu) == 0, "ReadBigUnsigned");
    std::sort(host.begin(), host.end(),
              [](const BigUnsigned &x, const BigUnsigned &y) { return CompareBU(x, y) < 0; });
    int cmp = 0;
    for (int i = 0; i < n && cmp == 0; ++i) cmp = CompareBU(host[i], gpu[i]);
    CHECK(cmp == 0, "sorted BigUnsigneds compare equal to std::sort's (compareBU)");
    int size = n;
    CHECK(Kernels::UniqueSorted(b, size) == 0, "UniqueSorted");
    host.erase(std::unique(host.begin(), host.end(),
                           [](const BigUnsigned &x, const BigUnsigned &y) { return CompareBU(x, y) == 0; }),
               host.end());
    CHECK(size == (int)host.size(), "unique count");
    CHECK(Kernels::ReadBigUnsigned(b, size, gpu) == 0, "ReadBigUnsigned");
    for (int i = 0; i < size && cmp == 0; ++i) cmp = CompareBU(host[i], gpu[i]);
    CHECK(cmp == 0, "unique BigUnsigneds compare equal to std::unique's");
  }
  {   // BRT over 87-bit keys, read as BrtNodes
    std::mt19937_64 rng(10);
    std::vector<std::pair<uint64_t, uint64_t>> kv(200000);
    for (auto &k : kv) k = {rng() & ((1ull << 23) - 1), rng()};
    kv.push_back({0, 0});
    std::sort(kv.begin(), kv.end());
    kv.erase(std::unique(kv.begin(), kv.end()), kv.end());
    const int m = (int)kv.size();
    std::vector<BigUnsigned> keys(m);
    std::vector<uint64_t> lo(m), hi(m);
    for (int i = 0; i < m; ++i) {
      hi[i] = kv[i].first;
      lo[i] = kv[i].second;
      keys[i] = bu_of(lo[i], hi[i]);
    }
    Buffer z, I;
    CHECK(Kernels::WriteBigUnsigned(keys, 16, z) == 0, "WriteBigUnsigned");
    z.dim = 3;
    CHECK(Kernels::BuildBinaryRadixTree_p(z, I, m, 87) == 0, "BuildBinaryRadixTree_p");
    std::vector<RefBrtNode> nodes;
    CHECK(Kernels::ReadBrtNodes(I, m, nodes) == 0, "ReadBrtNodes");
    oracle_tree *t = oracle_build_from_keys(lo.data(), hi.data(), m, 3, 87);
    std::vector<int32_t> left(m - 1), len(m - 1), par(m - 1);
    std::vector<uint8_t> ll(m - 1), rl(m - 1);
    std::vector<uint64_t> lcp_lo(m - 1), lcp_hi(m - 1);
    oracle_tree_brt(t, left.data(), ll.data(), rl.data(), lcp_lo.data(), lcp_hi.data(), len.data(),
                    par.data());
    oracle_tree_free(t);
    bool ok = (int)nodes.size() == m - 1;
    for (int i = 0; i < m - 1 && ok; ++i) {
      const BigUnsigned want = bu_of(lcp_lo[i], lcp_hi[i]);
      ok = CompareBU(nodes[i].lcp, want) == 0 && nodes[i].lcp_length == len[i] &&
           nodes[i].left == left[i] && nodes[i].left_leaf == ll[i] && nodes[i].parent == par[i] &&
           nodes[i].right_leaf == rl[i];
    }
    CHECK(ok, "BrtNodes equal the serial BRT (compareBrtNode fields)");
  }
  std::printf("PASS results read back as BigUnsigned and BrtNode records\n");
}

static bool same_octree(const std::vector<OctNode2> &a, oracle_tree *t) {
  const int64_t O = oracle_tree_octree_size(t);
  if ((int64_t)a.size() != O) return false;
  std::vector<int32_t> ch(O * 4), lf(O);
  oracle_tree_octree(t, ch.data(), lf.data());
  for (int64_t i = 0; i < O; ++i) {
    for (int c = 0; c < 4; ++c)
      if (a[i].children[c] != ch[i * 4 + c]) return false;
    if (a[i].leaf != lf[i]) return false;
  }
  return true;
}

// "An octree can be build using a bunch of points" (:352-446) and the
// stress test (:448-490), which in the reference compared an empty octree.
static void build_scenarios() {
  std::mt19937 rng(5);
  for (int round = 0; round < 60; ++round) {
    const int n = round == 0 ? 1000000 : 1000 - round * 7;
    std::vector<int2> pts(n);
    std::vector<int64_t> q(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      pts[i] = {(int32_t)(rng() & 0x7fffffff), (int32_t)(rng() & 0x7fffffff)};
      q[2 * i] = pts[i].x;
      q[2 * i + 1] = pts[i].y;
    }
    std::vector<OctNode2> oct;
    CHECK(Kernels::BuildOctree_p(pts, oct, kBits, kMbits) == 0, "BuildOctree_p");
    oracle_tree *t = oracle_build(q.data(), n, 2, kBits, nullptr);
    const bool ok = same_octree(oct, t);
    oracle_tree_free(t);
    CHECK(ok, "octree equals BuildOctree_s");
  }
  std::printf("PASS An octree can be build using a bunch of points / stress test\n");
  bool threw = false;
  try {
    std::vector<OctNode2> oct;
    Kernels::BuildOctree_p(std::vector<int2>(), oct, kBits, kMbits);
  } catch (const std::logic_error &) {
    threw = true;
  }
  CHECK(threw, "zero points throw logic_error");
  std::printf("PASS Zero points are rejected like the reference\n");
}

// Karras facade: float points -> Quantize -> BuildOctreeInParallel.
static void karras_scenario() {
  std::mt19937 rng(6);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  const int n = 100000;
  std::vector<float3> pts(n);
  std::vector<float> flat(3 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    pts[i] = {u(rng), u(rng), u(rng)};
    flat[3 * i] = pts[i].x; flat[3 * i + 1] = pts[i].y; flat[3 * i + 2] = pts[i].z;
  }
  const Resln r = make_resln(1 << 21, 3);
  auto q = Karras::Quantize(pts, r);
  std::vector<int64_t> oq(3 * (size_t)n);
  CHECK(oracle_quantize(flat.data(), n, 3, r.width, 0, nullptr, nullptr, 0, oq.data()) == n, "oracle quantize");
  bool ok = (int)q.size() == n;
  for (int i = 0; i < n && ok; ++i) ok = q[i].x == oq[3 * i] && q[i].y == oq[3 * i + 1] && q[i].z == oq[3 * i + 2];
  CHECK(ok, "Quantize equals the serial quantize");
  auto oct = Karras::BuildOctreeInParallel(q, r);
  auto oct2 = Karras::BuildOctreeFromPoints(pts, r);
  oracle_tree *t = oracle_build(oq.data(), n, 3, r.bits, nullptr);
  const int64_t O = oracle_tree_octree_size(t);
  std::vector<int32_t> ch(O * 8), lf(O);
  oracle_tree_octree(t, ch.data(), lf.data());
  ok = (int64_t)oct.size() == O && (int64_t)oct2.size() == O;
  for (int64_t i = 0; i < O && ok; ++i) {
    for (int c = 0; c < 8; ++c) ok &= oct[i].children[c] == ch[i * 8 + c] && oct2[i].children[c] == ch[i * 8 + c];
    ok &= oct[i].leaf == lf[i] && oct2[i].leaf == lf[i];
  }
  oracle_tree_free(t);
  CHECK(ok, "3D octree equals the serial octree");
  std::printf("PASS Karras facade (Quantize, BuildOctreeInParallel) in 3D\n");
}

// Octree consumers: FindLeaf (OctreeUtils.cpp:18-71), CellWalk
// (viewer/Octree2.cpp:265-313) with a visitor, and the octree's stream
// operators (C/OctNode.h:58-98), each against the oracle.
struct WalkLog {
  std::vector<OctCell2> cells;
};
static void log_cell(const OctCell2 &cell, const float2 &, const float2 &,
                     const std::vector<OctNode2> &, const Resln &, void *data) {
  static_cast<WalkLog *>(data)->cells.push_back(cell);
}

static void query_scenarios() {
  std::mt19937 rng(9);
  // FindLeaves, 3D
  {
    const Resln r = make_resln(1 << 21, 3);
    std::uniform_int_distribution<int> u(0, (1 << 21) - 1);
    std::vector<int3> pts(20000);
    for (auto &p : pts) p = {u(rng), u(rng), u(rng)};
    auto oct = Karras::BuildOctreeInParallel(pts, r);
    std::vector<int3> q(5000);
    for (auto &p : q) p = {u(rng), u(rng), u(rng)};
    auto cells = OctreeUtils::FindLeaves(q, oct, r);
    std::vector<int64_t> oq(3 * q.size());
    for (size_t i = 0; i < q.size(); ++i) { oq[3 * i] = q[i].x; oq[3 * i + 1] = q[i].y; oq[3 * i + 2] = q[i].z; }
    std::vector<int32_t> par(q.size()), octn(q.size()), dat(q.size()), lev(q.size()), st(q.size());
    oracle_find_leaf(oq.data(), (int64_t)q.size(), 3, r.bits, &oct[0].children[0], (int64_t)oct.size(),
                     par.data(), octn.data(), dat.data(), lev.data(), st.data());
    bool ok = cells.size() == q.size();
    for (size_t i = 0; i < q.size() && ok; ++i) {
      const int32_t w = 1 << (r.bits - lev[i]);
      ok = st[i] == 0 && cells[i].parent_idx == par[i] && cells[i].octant == octn[i] &&
           cells[i].data == dat[i] && cells[i].width == w && cells[i].origin[0] == (q[i].x & ~(w - 1)) &&
           cells[i].origin[2] == (q[i].z & ~(w - 1));
    }
    CHECK(ok, "FindLeaves equals the serial FindLeaf");
    const OctCell3 one = OctreeUtils::FindLeaf(q[7], oct, r);
    CHECK(one.parent_idx == par[7] && one.octant == octn[7], "FindLeaf of one point");
  }
  // CellWalk with a visitor, 2D
  {
    const Resln r = make_resln(1 << 12, 2);
    std::uniform_int_distribution<int> u(0, (1 << 12) - 1);
    std::uniform_real_distribution<float> f(0.f, 4095.9f);
    std::vector<int2> pts(3000);
    for (auto &p : pts) p = {u(rng), u(rng)};
    auto oct = Karras::BuildOctreeInParallel(pts, r);
    bool ok = true;
    for (int s = 0; s < 50 && ok; ++s) {
      const float2 a = {f(rng), f(rng)}, b = {f(rng), f(rng)};
      WalkLog log;
      OctreeUtils::CellWalk(a, b, oct, r, log_cell, &log);
      const float seg[4] = {a.x, a.y, b.x, b.y};
      int32_t count = 0, st = 0;
      const int64_t tot = oracle_cell_walk(seg, 1, r.bits, &oct[0].children[0], (int64_t)oct.size(),
                                           &count, &st, nullptr, 0);
      std::vector<int32_t> want(5 * (size_t)tot);
      oracle_cell_walk(seg, 1, r.bits, &oct[0].children[0], (int64_t)oct.size(), &count, &st,
                       want.data(), tot);
      ok = st == 0 && (int64_t)log.cells.size() == tot;
      for (int64_t k = 0; k < tot && ok; ++k) {
        const OctCell2 &c = log.cells[k];
        ok = c.origin[0] == want[5 * k] && c.origin[1] == want[5 * k + 1] && c.width == want[5 * k + 2] &&
             c.parent_idx == want[5 * k + 3] && c.octant == want[5 * k + 4];
      }
    }
    CHECK(ok, "CellWalk visits the serial walk's cells");
    // stream operators
    std::ostringstream os;
    os << oct;
    const std::string text = os.str();
    const int64_t len = oracle_octree_format_text(&oct[0].children[0], (int64_t)oct.size(), 2, nullptr);
    std::string want((size_t)len, ' ');
    oracle_octree_format_text(&oct[0].children[0], (int64_t)oct.size(), 2, &want[0]);
    CHECK(text == want, "operator<< prints the serial text");
    std::istringstream is(text);
    std::vector<OctNode2> back;
    is >> back;
    ok = !is.fail() && back.size() == oct.size();
    for (size_t i = 0; i < oct.size() && ok; ++i)
      for (int k = 0; k < 4; ++k) ok &= back[i].children[k] == oct[i].children[k] && back[i].leaf == oct[i].leaf;
    CHECK(ok, "operator>> reads the octree back");
    std::istringstream bad("3 1 2 x");
    is.clear();
    bad >> back;
    CHECK(bad.fail(), "operator>> sets failbit on malformed text");
  }
  std::printf("PASS octree queries (FindLeaf, CellWalk) and stream operators\n");
}

// The GVD refinement: Geom::FitBoxes against the oracle's, its logic_error
// on meeting segments that the reference rejects, and Octree2's polyline
// build with one FindMultiCells (viewer/Octree2.cpp:127-190) against the
// oracle's octree, CellIntersections and extra points.
static void gvd_scenarios() {
  std::mt19937 rng(8);
  std::uniform_real_distribution<float> u(0.f, 1000.f);
  for (int k = 0; k < 200; ++k) {
    const FloatSegment A{{u(rng), u(rng)}, {u(rng), u(rng)}};
    FloatSegment B{{A.a.x + 3.f, A.a.y + 7.f}, {A.b.x + 5.f, A.b.y + 2.f}};
    if (k % 2) B = FloatSegment{{u(rng), u(rng)}, {u(rng), u(rng)}};
    const float a[4] = {A.a.x, A.a.y, A.b.x, A.b.y}, b[4] = {B.a.x, B.a.y, B.b.x, B.b.y};
    int64_t nb = 0;
    int st = 0;
    const int64_t ns = oracle_fit_boxes(a, b, 1.f, nullptr, nullptr, nullptr, 0, &nb, &st);
    std::vector<float> ws(2 * (size_t)std::max<int64_t>(ns, 1)), wo(2 * (size_t)std::max<int64_t>(nb, 1)),
        wl((size_t)std::max<int64_t>(nb, 1));
    oracle_fit_boxes(a, b, 1.f, ws.data(), wo.data(), wl.data(), std::max(ns, nb), &nb, &st);
    std::vector<float2> smp, org;
    std::vector<float> len;
    bool threw = false;
    try {
      Geom::FitBoxes(A, B, 1.f, &smp, &org, &len);
    } catch (const std::logic_error &) {
      threw = true;
    }
    CHECK(threw == (st != 0), "FitBoxes throws where the oracle reports an error");
    bool ok = (int64_t)smp.size() == ns && (int64_t)org.size() == nb;
    for (int64_t i = 0; i < ns && ok; ++i) ok = smp[i].x == ws[2 * i] && smp[i].y == ws[2 * i + 1];
    for (int64_t i = 0; i < nb && ok; ++i)
      ok = org[i].x == wo[2 * i] && org[i].y == wo[2 * i + 1] && len[i] == wl[i];
    CHECK(ok, "FitBoxes equals the oracle's");
  }
  // two polylines facing each other across a narrow gap
  std::vector<std::vector<float2>> polys = {
      {{0.f, 0.f}, {0.595f, 0.01f}, {0.575f, 1.f}, {0.f, 1.f}, {0.f, 0.f}},
      {{0.605f, 0.f}, {1.f, 0.f}, {1.f, 1.f}, {0.615f, 0.99f}, {0.605f, 0.f}}};
  const int bits = 12;
  const Resln r = make_resln(1 << bits, 2);
  Octree2 o2(r, 1);
  o2.build(polys);
  std::vector<float> flat;
  for (const auto &p : polys)
    for (const auto &v : p) { flat.push_back(v.x); flat.push_back(v.y); }
  const int64_t nv = (int64_t)flat.size() / 2;
  std::vector<int64_t> q(2 * (size_t)nv);
  CHECK(oracle_quantize(flat.data(), nv, 2, r.width, 0, nullptr, nullptr, 0, q.data()) == nv, "quantize");
  oracle_tree *t = oracle_build(q.data(), nv, 2, bits, nullptr);
  const bool same = same_octree(o2.octree(), t);
  const int64_t O = oracle_tree_octree_size(t);
  std::vector<int32_t> ch(O * 4), lf(O), oct(O * 5);
  oracle_tree_octree(t, ch.data(), lf.data());
  oracle_tree_free(t);
  CHECK(same, "Octree2 octree equals the serial octree");
  for (int64_t i = 0; i < O; ++i) {
    for (int c = 0; c < 4; ++c) oct[5 * i + c] = ch[4 * i + c];
    oct[5 * i + 4] = lf[i];
  }
  float bbmin[2] = {flat[0], flat[1]}, bbmax[2] = {flat[0], flat[1]};
  for (int64_t i = 0; i < nv; ++i)
    for (int k = 0; k < 2; ++k) {
      bbmin[k] = std::min(bbmin[k], flat[2 * i + k]);
      bbmax[k] = std::max(bbmax[k], flat[2 * i + k]);
    }
  const float bbw = std::max(bbmax[0] - bbmin[0], bbmax[1] - bbmin[1]);
  std::vector<float> octv(flat.size());
  CHECK(oracle_obj2oct(flat.data(), nv, bbmin, bbw, r.width, octv.data()) == 0, "obj2oct");
  std::vector<float> segs;
  std::vector<int32_t> labels;
  int64_t at = 0;
  for (size_t j = 0; j < polys.size(); ++j) {
    for (size_t i = 0; i + 1 < polys[j].size(); ++i) {
      for (int k = 0; k < 2; ++k) segs.push_back(octv[2 * (at + i) + k]);
      for (int k = 0; k < 2; ++k) segs.push_back(octv[2 * (at + i + 1) + k]);
      labels.push_back((int32_t)j);
    }
    at += (int64_t)polys[j].size();
  }
  std::vector<int32_t> cil(O * 8), wst(labels.size());
  std::vector<float> cis(O * 32);
  oracle_multi_cells(segs.data(), labels.data(), (int64_t)labels.size(), bits, oct.data(), O,
                     cil.data(), cis.data(), wst.data());
  bool ok = (int64_t)o2.cell_intersections().size() == O;
  int multi = 0;
  for (int64_t i = 0; i < O && ok; ++i) {
    const CellIntersections &c = o2.cell_intersections()[i];
    for (int j = 0; j < 8; ++j) {
      ok &= c.l[j] == cil[8 * i + j];
      ok &= c.segs[j].a.x == cis[32 * i + 4 * j] && c.segs[j].a.y == cis[32 * i + 4 * j + 1] &&
            c.segs[j].b.x == cis[32 * i + 4 * j + 2] && c.segs[j].b.y == cis[32 * i + 4 * j + 3];
    }
    for (int oc = 0; oc < 4; ++oc) multi += c.is_multi(oc);
  }
  CHECK(ok, "Octree2 cell_intersections equal the oracle's");
  CHECK(multi > 0, "the scene has multi-cells");
  int64_t nb = 0;
  int st = 0;
  const int64_t ne = oracle_find_multi_cells(cil.data(), cis.data(), O, 1.f, nullptr, nullptr,
                                             nullptr, nullptr, 0, &nb, &st);
  std::vector<int32_t> ex(2 * (size_t)std::max<int64_t>(ne, 1));
  std::vector<float> smp(2 * (size_t)std::max<int64_t>(ne, 1));
  oracle_find_multi_cells(cil.data(), cis.data(), O, 1.f, ex.data(), smp.data(), nullptr, nullptr,
                          ne, &nb, &st);
  ok = st == 0 && (int64_t)o2.extra_qpoints().size() == ne && ne > 0;
  for (int64_t i = 0; i < ne && ok; ++i)
    ok = o2.extra_qpoints()[i].x == ex[2 * i] && o2.extra_qpoints()[i].y == ex[2 * i + 1];
  CHECK(ok, "Octree2 extra_qpoints equal the oracle's");
  std::printf("PASS GVD refinement (FitBoxes, Octree2 polyline build with FindMultiCells)\n");
}

int main() {
  morton_scenario();
  sort_scenario();
  bit_serial_scenario();
  unique_scenario();
  tree_scenarios();
  big_unsigned_scenarios();
  build_scenarios();
  karras_scenario();
  query_scenarios();
  gvd_scenarios();
  std::printf(failures ? "FAILED %d\n" : "ALL PASSED\n", failures);
  return failures ? 1 : 0;
}
