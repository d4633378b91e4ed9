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
#include "brt_eval.hpp"
#include "tilescan.hpp"

namespace pgvd {

namespace {

// Stores a BRT record as whole words (one 16-byte store for 64-bit keys)
// instead of the field-by-field byte and short stores of a struct copy.
__device__ inline void store_brt(BrtNodeT<uint32_t> *p, const BrtNodeT<uint32_t> &n) {
  const uint32_t tail = (uint32_t)n.lcp_length | ((uint32_t)n.left_leaf << 16) |
                        ((uint32_t)n.right_leaf << 24);
  uint32_t *q = reinterpret_cast<uint32_t *>(p);   // 12-byte records: 4-byte aligned
  q[0] = n.lcp;
  q[1] = (uint32_t)n.left;
  q[2] = tail;
}
__device__ inline void store_brt(BrtNodeT<uint64_t> *p, const BrtNodeT<uint64_t> &n) {
  const uint32_t tail = (uint32_t)n.lcp_length | ((uint32_t)n.left_leaf << 16) |
                        ((uint32_t)n.right_leaf << 24);
  *reinterpret_cast<uint4 *>(p) =   // 16-byte records from a 256-byte aligned base
      make_uint4((uint32_t)n.lcp, (uint32_t)(n.lcp >> 32), (uint32_t)n.left, tail);
}
__device__ inline void store_brt(BrtNodeT<Key128> *p, const BrtNodeT<Key128> &n) {
  const uint32_t tail = (uint32_t)n.lcp_length | ((uint32_t)n.left_leaf << 16) |
                        ((uint32_t)n.right_leaf << 24);
  uint2 *q = reinterpret_cast<uint2 *>(p);         // 24-byte records: 8-byte aligned
  q[0] = make_uint2((uint32_t)n.lcp.lo, (uint32_t)(n.lcp.lo >> 32));
  q[1] = make_uint2((uint32_t)n.lcp.hi, (uint32_t)(n.lcp.hi >> 32));
  q[2] = make_uint2((uint32_t)n.left, tail);
}

template <typename K>
__device__ inline void brt_flags(const BrtOut<K> &o, RunState *st) {
  if (o.dup) atomicOr(&st->status, (uint32_t)ST_DUP_KEY);
  if (o.quirk) st->malformed = 1u;
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
  const K cur = keys[i];
  const BrtOut<K> o = brt_eval(cur, [&](int x) { return keys[x]; }, mbits, m, i);
  brt_flags(o, st);
  brt[i] = o.nd;
  // (with three equal keys at the end, node m-2 claims node m-1, one past
  // the m-1 nodes; the reference writes past its array there)
  if (!o.nd.left_leaf) atomicMax(&parent[o.nd.left], i + 1);
  if (!o.nd.right_leaf && o.nd.left + 1 < m - 1) atomicMax(&parent[o.nd.left + 1], i + 1);
}

// Pipeline store of node c: the record, its children's parent links as
// plain stores (every non-root node of a well-formed radix tree has exactly
// one claimer; root -1), and lsx[c] = ls | octant << 8 where
//  * ls = lcp_length(c)/DIM - lcp_min(c)/DIM: in a well-formed radix tree
//    the parent's lcp_length is the node's lcp_min, so this is the local
//    split C/BuildOctree.c:17-27 computes from the parent (root: 1 + len/DIM);
//  * octant = digit of keys[c] at the parent's octree level: key c lies in
//    node c's range, so this is quadrantInLcp(c, ls - 1)
//    (C/BuildOctree.c:47-55), the octant c's chain attaches in.
// If some node took an FP32-quirk split (st->malformed), parent links, local
// splits and prefix sums are recomputed by the order-exact claim path.
template <typename K, int DIM>
__device__ inline uint32_t brt_store(int c, K cur, const BrtOut<K> &o, int mbits,
                                     BrtNodeT<K> *__restrict__ brt, int32_t *__restrict__ parent,
                                     uint32_t *__restrict__ lsx) {
  store_brt(brt + c, o.nd);
  if (!o.nd.left_leaf) parent[o.nd.left] = c;
  if (!o.nd.right_leaf) parent[o.nd.left + 1] = c;
  const int len = o.nd.lcp_length;
  if (c == 0) {
    parent[0] = -1;
    const uint32_t ls = 1u + (uint32_t)(len / DIM);
    lsx[0] = ls;
    return ls;
  }
  const int plvl = o.lcp_min / DIM;
  const uint32_t ls = (uint32_t)(len / DIM - plvl);
  lsx[c] = ls | ((uint32_t)digit_at<K, DIM>(cur, plvl, mbits) << 8);
  return ls;
}

// Adjacent-key deltas staged per workgroup: its SEG = R*256 nodes plus H
// positions either side.
template <typename K> struct BrtWin { static constexpr int R = 4, H = 256; };

// Pipeline, first pass.  A workgroup builds, in LDS, the sparse table of
// adjacent-key deltas over positions [b0-H, b0+SEG+H) of keys[0..m) (one
// byte per entry and level) and resolves nodes [b0, b0+SEG) from it
// (brt_node_window): each node is a few table descents instead of a chain of
// key probes.  A node whose range or split leaves the window is handed to
// defer(true, c) and left to the global-memory pass.  lsx[m-1] = 0;
// m == 1 gives lsx[0] = 1.  Every lane calls defer once per row, converged.
// With chunk_tot, each row of 256 nodes (a chunk) also stores the sum of
// its local splits, deferred nodes left out (their pass adds them).
template <typename K, int DIM, int R, typename Defer>
__device__ inline void brt_window(const K *__restrict__ keys, int m, int b0, int mbits,
                                  BrtNodeT<K> *__restrict__ brt, int32_t *__restrict__ parent,
                                  uint32_t *__restrict__ lsx, uint32_t *__restrict__ chunk_tot,
                                  RunState *st, Defer defer) {
  __shared__ uint32_t s_rowsum[R * (kBlock / 64)];
  constexpr int H = BrtWin<K>::H, SEG = R * kBlock, WN = SEG + 2 * H;
  constexpr int LV = (WN >= 2048) ? 12 : (WN >= 1024 ? 11 : (WN >= 512 ? 10 : 9));
  constexpr int PT = WN / kBlock;     // window positions per thread
  static_assert(WN % kBlock == 0 && H == kBlock, "thread t holds positions t + 256 i");
  using Tab = DeltaTab<WN, LV>;
  constexpr int kWords = (Tab::kBytes + 3) / 4;
  __shared__ __attribute__((aligned(16))) uint32_t s_tab[kWords];
  uint8_t *s_mn = reinterpret_cast<uint8_t *>(s_tab);
  const int w0 = b0 - H;
  const int t = threadIdx.x;
  // Window keys: position t + 256 i needs keys[g] and keys[g+1], g = w0 + t +
  // 256 i; all loads are issued before the first is used.  Node rows are
  // positions 1 .. R, so the nodes' own keys come from the same loads.
  K ka[PT], kb[PT];
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const int g = w0 + t + i * kBlock;
    const bool v = g >= 0 && g < m - 1;
    ka[i] = keys[v ? g : 0];
    kb[i] = keys[v ? g + 1 : 0];
  }
  // sentinels: padding and the block ends of every level
  for (int r = t; r < kWords; r += kBlock) s_tab[r] = 0;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PT; ++i) {
    const int g = w0 + t + i * kBlock;
    s_mn[Tab::base(0) + t + i * kBlock] =
        (g >= 0 && g < m - 1) ? (uint8_t)(1 + delta(ka[i], kb[i], mbits)) : 0;
  }
  __syncthreads();
  // Two levels per pass from the level below them: entry (j, r) is the min
  // of two entries of level j-1, entry (j+1, r) of four; probes past WN read
  // the level's last entry (0) and only feed entries left at 0.
#pragma unroll
  for (int j = 1; j < LV; j += 2) {
    const int q = 1 << (j - 1);
    const uint8_t *prev = s_mn + Tab::base(j - 1);
    uint8_t x[PT][4];
#pragma unroll
    for (int i = 0; i < PT; ++i) {
      const int r = t + i * kBlock;
#pragma unroll
      for (int u = 0; u < 4; ++u) x[i][u] = prev[min(r + u * q, WN)];
    }
#pragma unroll
    for (int i = 0; i < PT; ++i) {
      const int r = t + i * kBlock;
      const uint8_t a = x[i][0] < x[i][1] ? x[i][0] : x[i][1];
      const uint8_t b = x[i][2] < x[i][3] ? x[i][2] : x[i][3];
      if (r + 2 * q <= WN) s_mn[Tab::base(j) + r] = a;
      if (j + 1 < LV && r + 4 * q <= WN) s_mn[Tab::base(j + 1) + r] = a < b ? a : b;
    }
    __syncthreads();
  }
  const Tab tab{s_mn};
  // every row's searches first (independent, so they overlap), then stores
  BrtOut<K> o[R];
  bool ok[R];
#pragma unroll
  for (int r = 0; r < R; ++r)
    ok[r] = brt_node_window(ka[1 + r], tab, w0, mbits, m, b0 + r * kBlock + t, o[r]);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = b0 + r * kBlock + t;
    bool miss = false;
    uint32_t ls = 0;
    if (c < m - 1) {
      miss = !ok[r];
      if (!miss) {
        brt_flags(o[r], st);
        ls = brt_store<K, DIM>(c, ka[1 + r], o[r], mbits, brt, parent, lsx);
      }
    } else if (c == m - 1) {
      ls = (m == 1) ? 1u : 0u;
      lsx[c] = ls;
    }
    defer(miss, c);
    if (chunk_tot) {
      const uint32_t sum = wave_reduce_add(ls);
      if ((t & 63) == 0) s_rowsum[r * (kBlock / 64) + (t >> 6)] = sum;
    }
  }
  if (chunk_tot) {
    __syncthreads();
    if (threadIdx.x < R && b0 + (int)threadIdx.x * kBlock < m) {
      uint32_t tot = 0;
      for (int w = 0; w < kBlock / 64; ++w) tot += s_rowsum[threadIdx.x * (kBlock / 64) + w];
      chunk_tot[(b0 >> 8) + threadIdx.x] = tot;
    }
  }
}

// Waves per SIMD the BRT window kernel is compiled for: 6 caps it at 80
// VGPRs (it needs 81 uncapped, which leaves 5), without spills.
#ifndef PGVD_BRT_WPE
#define PGVD_BRT_WPE 6
#endif
template <typename K, int DIM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(PGVD_BRT_WPE))) void
brt_ls_kernel(const K *__restrict__ keys, int mbits,
                                                        BrtNodeT<K> *__restrict__ brt,
                                                        int32_t *__restrict__ parent,
                                                        uint32_t *__restrict__ lsx,
                                                        uint32_t *__restrict__ chunk_tot,
                                                        uint64_t *__restrict__ miss_words,
                                                        RunState *st) {
  constexpr int R = BrtWin<K>::R, SEG = R * kBlock;
  const int m = (int)st->m;
  const int b0 = (int)xcd_segment(blockIdx.x, gridDim.x) * SEG;
  if (b0 >= m) return;                      // uniform over the workgroup
  brt_window<K, DIM, R>(keys, m, b0, mbits, brt, parent, lsx, chunk_tot, st,
                        [&](bool miss, int c) { wave_mask(miss, miss_words, (uint32_t)c >> 6); });
}

// Pipeline, second pass: the nodes the window could not hold (set bits of
// the first pass's miss words, one word per 64 nodes), against keys in
// global memory, with the 8-ary searches (7 independent loads a round in
// place of one dependent load a step).
template <typename K, int DIM>
__global__ __launch_bounds__(kBlock) void brt_ls_defer_kernel(const K *__restrict__ keys,
                                                              int mbits,
                                                              BrtNodeT<K> *__restrict__ brt,
                                                              int32_t *__restrict__ parent,
                                                              uint32_t *__restrict__ lsx,
                                                              uint32_t *__restrict__ chunk_tot,
                                                              const uint64_t *__restrict__ miss_words,
                                                              RunState *st) {
  const int m = (int)st->m;
  const uint32_t nwords = m > 1 ? (uint32_t)(m - 1 + 63) / 64u : 0u;
  for_each_masked(miss_words, nwords, [&](uint32_t cu) {
    const int c = (int)cu;
    const K cur = keys[c];
    const BrtOut<K> o = brt_eval_wide(cur, [&](int x) { return keys[x]; }, mbits, m, c);
    brt_flags(o, st);
    const uint32_t ls = brt_store<K, DIM>(c, cur, o, mbits, brt, parent, lsx);
    if (ls) atomicAdd(&chunk_tot[c >> 8], ls);
  });
}

// Batched form: BRT tiles of R*256 key slots per instance (n+1 slots each);
// instance s's arrays start at key slot kofs, m = ss[s].m.  Deferred nodes
// are recorded by key slot.
template <typename K, int DIM, int R>
__global__ __launch_bounds__(kBlock) void brt_seg_kernel(const K *__restrict__ keys, int mbits,
                                                         const SegInfo *__restrict__ seg,
                                                         TileTable tt,
                                                         const SegState *__restrict__ ss,
                                                         BrtNodeT<K> *__restrict__ brt,
                                                         int32_t *__restrict__ parent,
                                                         uint32_t *__restrict__ lsx,
                                                         uint32_t *__restrict__ defer,
                                                         RunState *st) {
  constexpr int SEG = R * kBlock;
  const uint32_t tile = xcd_segment(blockIdx.x, gridDim.x);
  const uint32_t s = tt.seg[tile];
  const int m = (int)ss[s].m;
  const int b0 = (int)(tile - tt.first[s]) * SEG;
  if (b0 >= m) return;
  const uint32_t kofs = seg[s].kofs;
  brt_window<K, DIM, R>(keys + kofs, m, b0, mbits, brt + kofs, parent + kofs, lsx + kofs, nullptr, st,
                        [&](bool miss, int c) {
                          wave_append(miss, kofs + (uint32_t)c, &st->counters[CTR_BRT_DEFER],
                                      defer);
                        });
}

template <typename K, int DIM>
__global__ __launch_bounds__(kBlock) void brt_seg_defer_kernel(const K *__restrict__ keys,
                                                               int mbits,
                                                               const uint32_t *__restrict__ kofs_all,
                                                               uint32_t B,
                                                               const SegState *__restrict__ ss,
                                                               BrtNodeT<K> *__restrict__ brt,
                                                               int32_t *__restrict__ parent,
                                                               uint32_t *__restrict__ lsx,
                                                               const uint32_t *__restrict__ defer,
                                                               RunState *st) {
  const uint32_t cnt = st->counters[CTR_BRT_DEFER];
  for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < cnt; k += gridDim.x * kBlock) {
    const uint32_t slot = defer[k];
    const uint32_t s = find_seg(kofs_all, B, slot);
    const uint32_t kofs = kofs_all[s];
    const int c = (int)(slot - kofs);
    const int m = (int)ss[s].m;
    const K *kk = keys + kofs;
    const K cur = kk[c];
    const BrtOut<K> o = brt_eval_wide(cur, [&](int x) { return kk[x]; }, mbits, m, c);
    brt_flags(o, st);
    (void)brt_store<K, DIM>(c, cur, o, mbits, brt + kofs, parent + kofs, lsx + kofs);
  }
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
  // the last claimer in node order; node 0 sets its own -1 before any
  // claim (C/BuildBRT.c:169-179), others keep the zero they start with
  const int32_t pc = parent[c];
  parent[c] = pc ? pc - 1 : (c == 0 ? -1 : 0);
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

template <typename K, int DIM>
void launch_brt_ls_dim(const K *keys, int64_t max_m, int mbits, BrtNodeT<K> *brt,
                       int32_t *parent, uint32_t *lsx, uint32_t *chunk_tot, uint32_t *defer,
                       RunState *st, hipStream_t s) {
  constexpr int SEG = BrtWin<K>::R * kBlock;
  const unsigned grid = (unsigned)((max_m + SEG - 1) / SEG);
  uint64_t *words = reinterpret_cast<uint64_t *>(defer);
  brt_ls_kernel<K, DIM><<<grid, kBlock, 0, s>>>(keys, mbits, brt, parent, lsx, chunk_tot, words,
                                                st);
  // The overflow pass: one thread per miss word.
  const int64_t g2 = ((max_m + 63) / 64 + kBlock - 1) / kBlock;
  const unsigned grid2 = (unsigned)(g2 < 4096 ? g2 : 4096);
  brt_ls_defer_kernel<K, DIM><<<grid2, kBlock, 0, s>>>(keys, mbits, brt, parent, lsx, chunk_tot,
                                                       words, st);
}

template <typename K>
void launch_brt_ls(const K *keys, int64_t max_m, int mbits, int dim, BrtNodeT<K> *brt,
                   int32_t *parent, uint32_t *lsx, uint32_t *chunk_tot, uint32_t *defer,
                   RunState *st, hipStream_t s) {
  if (max_m < 1) max_m = 1;
  if (dim == 2) launch_brt_ls_dim<K, 2>(keys, max_m, mbits, brt, parent, lsx, chunk_tot, defer, st, s);
  else launch_brt_ls_dim<K, 3>(keys, max_m, mbits, brt, parent, lsx, chunk_tot, defer, st, s);
}

bool brt_seg_rows_ok(int key_bytes, int rows) {
  (void)key_bytes;
  return rows == 2 || rows == 4;
}

template <typename K, int DIM, int R>
void brt_seg_dim(const K *keys, int mbits, const BatchDev &b, BrtNodeT<K> *brt, int32_t *parent,
                 uint32_t *lsx, uint32_t *defer, RunState *st, hipStream_t s) {
  brt_seg_kernel<K, DIM, R><<<(unsigned)b.n_brt, kBlock, 0, s>>>(keys, mbits, b.seg, b.brt, b.ss,
                                                                 brt, parent, lsx, defer, st);
  brt_seg_defer_kernel<K, DIM><<<2048, kBlock, 0, s>>>(keys, mbits, b.kofs, (uint32_t)b.B, b.ss,
                                                       brt, parent, lsx, defer, st);
}

template <typename K>
void launch_brt_seg(const K *keys, int mbits, int dim, const BatchDev &b, BrtNodeT<K> *brt,
                    int32_t *parent, uint32_t *lsx, uint32_t *defer, RunState *st,
                    hipStream_t s) {
  if (b.n_brt <= 0) return;
  const bool small = b.brt_rows == 2;
  if (dim == 2) {
    if (small) brt_seg_dim<K, 2, 2>(keys, mbits, b, brt, parent, lsx, defer, st, s);
    else brt_seg_dim<K, 2, 4>(keys, mbits, b, brt, parent, lsx, defer, st, s);
  } else {
    if (small) brt_seg_dim<K, 3, 2>(keys, mbits, b, brt, parent, lsx, defer, st, s);
    else brt_seg_dim<K, 3, 4>(keys, mbits, b, brt, parent, lsx, defer, st, s);
  }
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
  template void launch_claims_if_malformed<K>(const BrtNodeT<K> *, int32_t *, int64_t,         \
                                              RunState *, hipStream_t);                        \
  template void launch_brt_ls<K>(const K *, int64_t, int, int, BrtNodeT<K> *, int32_t *,       \
                                 uint32_t *, uint32_t *, uint32_t *, RunState *, hipStream_t); \
  template void launch_brt_seg<K>(const K *, int, int, const BatchDev &, BrtNodeT<K> *,        \
                                  int32_t *, uint32_t *, uint32_t *, RunState *, hipStream_t);
PGVD_INST(uint32_t)
PGVD_INST(uint64_t)
PGVD_INST(Key128)
#undef PGVD_INST

}  // namespace pgvd
