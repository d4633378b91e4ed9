// octree.hip -- BRT to octree (brt2octree_init / brt2octree,
// C/BuildOctree.c:73-139, driven as in opencl/Kernels.cpp:370-386), plus the
// per-node conflict octant masks.
//
// Layout: BRT node b with local split ls_b > 0 owns the octree nodes
// [ps[b-1], ps[b]) (ps inclusive, ps[-1] = 0).  Its lowest node ps[b-1] is the
// attach target of b's octree children; nodes ps[b-1]+k (k >= 1) form b's
// chain, node ps[b-1]+k holding child ps[b-1]+k-1 in octant
// quadrantInLcp(b, k-1) (C/BuildOctree.c:88-101).  A child chain c attaches
// its top node ps[c]-1 in octant quadrantInLcp(c, ls_c-1) of the lowest node
// of the first ancestor with a local split (:102-130).  The root's chain is
// never linked (the reference starts at brt_i = 1), and `leaf` has bit o set
// exactly when children[o] is -1 (init 2^(2^DIM)-1, the reference's 15 in
// 2D).
//
// Two exact strategies:
//   gather  (tree well formed): one lane per BRT node writes every node it
//           owns once, finding the lowest node's children by descending
//           through the same-level (ls == 0) BRT nodes below it -- at most
//           2^DIM - 1 of them.  No atomics, no init pass.
//   scatter (some node took an FP32-quirk split): init pass, then the
//           reference's walk-up per BRT node, with colliding attaches
//           resolved by atomicMax -- the top-of-chain index ps[c]-1 grows
//           with c, so the largest value is the serial loop's last writer.
#include "kernels.hpp"
#include "tilescan.hpp"

namespace pgvd {

namespace {

template <typename K, int DIM>
__device__ inline int quadrant(const BrtNodeT<K> &b, int i) {
  const int rshift = i * DIM + b.lcp_length % DIM;
  return (int)(klow32(kshr((K)b.lcp, rshift)) & ((1u << DIM) - 1));
}

__device__ inline bool test_bit(const uint32_t *bits, uint32_t i) {
  return (bits[i >> 5] >> (i & 31)) & 1u;
}

__device__ inline uint32_t ls_of(const uint32_t *ps, uint32_t c) {
  return c == 0 ? ps[0] : ps[c] - ps[c - 1];
}

// A BRT record's child links packed as left | left_leaf << 30 | right_leaf
// << 31 (left < 2^30), read as one word pair where the record allows.
template <typename K>
__device__ inline uint32_t pack_links(uint32_t left, uint32_t tail) {
  return left | (((tail >> 16) & 1u) << 30) | (((tail >> 24) & 1u) << 31);
}
__device__ inline uint32_t load_links(const BrtNodeT<uint32_t> *p) {
  const uint32_t *q = reinterpret_cast<const uint32_t *>(p);
  return pack_links<uint32_t>(q[1], q[2]);
}
__device__ inline uint32_t load_links(const BrtNodeT<uint64_t> *p) {
  const uint2 v = reinterpret_cast<const uint2 *>(p)[1];
  return pack_links<uint64_t>(v.x, v.y);
}
__device__ inline uint32_t load_links(const BrtNodeT<Key128> *p) {
  const uint2 v = reinterpret_cast<const uint2 *>(p)[2];
  return pack_links<Key128>(v.x, v.y);
}

template <int DIM>
__device__ inline void write_node(int32_t *octree, uint32_t idx, const int32_t *ch) {
  constexpr int NC = 1 << DIM;
  int32_t *o = octree + (size_t)idx * (NC + 1);
  int32_t leaf = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    o[c] = ch[c];
    leaf |= (ch[c] == -1) << c;
  }
  o[NC] = leaf;
}

// Octree nodes owned by BRT node q: the lowest, excl = ps[q-1], takes the
// children found by descending through the same-level (ls == 0) BRT nodes
// below q; chain nodes excl+k (1 <= k < ls_q) each take excl+k-1.
//   links(x, left, ll, rl) -> false if x's record is unavailable
//   child(c, ps_c, lsx_c)  -> false if c's entries are unavailable
// Returns false, writing nothing, when a read was unavailable.
template <typename K, int DIM, typename Links, typename Child, typename Put>
__device__ __forceinline__ bool gather_one(uint32_t q, const BrtNodeT<K> &nq, uint32_t excl, uint32_t lsq,
                                  Links links, Child child, const K *__restrict__ keys,
                                  const uint32_t *__restrict__ conflict_bits, int mbits, Put put,
                                  uint32_t *__restrict__ node_conflict, RunState *st) {
  constexpr int NC = 1 << DIM;
  int32_t ch[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) ch[c] = -1;
  const int lvl = nq.lcp_length / DIM;
  uint32_t cmask = 0;
  uint32_t stack[8];
  int sp = 0;
  uint32_t xl = (uint32_t)nq.left;
  bool xll = nq.left_leaf, xrl = nq.right_leaf;
  // A well-formed tree visits at most 2^DIM - 1 nodes here; the cap only
  // guards against a malformed tree the BRT stage did not flag.
  int visits = 0;
  for (;;) {
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const uint32_t c = xl + side;
      if (side ? xrl : xll) {
        if (conflict_bits && test_bit(conflict_bits, c))
          cmask |= 1u << digit_at<K, DIM>(keys[c], lvl, mbits);
        continue;
      }
      uint32_t psc, lxc;
      if (!child(c, psc, lxc)) return false;
      if ((lxc & 0xffu) > 0) {
        ch[lxc >> 8] = (int32_t)(psc - 1);     // c's chain top, in its octant
      } else if (sp < 8) {
        stack[sp++] = c;
      } else {
        atomicOr(&st->status, (uint32_t)ST_STACK);
      }
    }
    if (sp == 0) break;
    if (++visits > 64) {
      atomicOr(&st->status, (uint32_t)ST_STACK);
      break;
    }
    if (!links(stack[--sp], xl, xll, xrl)) return false;
  }
  put(excl, ch);
  if (node_conflict) node_conflict[excl] = cmask;
  for (uint32_t k = 1; k < lsq; ++k) {
#pragma unroll
    for (int c = 0; c < NC; ++c) ch[c] = -1;
    if (q != 0) ch[quadrant<K, DIM>(nq, (int)k - 1)] = (int32_t)(excl + k - 1);
    put(excl + k, ch);
    if (node_conflict) node_conflict[excl + k] = 0;
  }
  return true;
}

// A window position's prefix sum (relative to the window's base) and its
// packed local split (local split | attach octant << 8 < 2^11) in one word.
constexpr int kPlShift = 11;
constexpr uint32_t kPlLsx = (1u << kPlShift) - 1;

// A slot of gather_one_bfs's rounds in one word: BRT index (30 bits), then
// its leaf flag (bit 30) and live flag (bit 31) -- so that a round's state
// stays in vector registers rather than in lane masks.
constexpr uint32_t kSlotIdx = 0x3fffffffu, kSlotLeaf = 1u << 30, kSlotLive = 1u << 31;

// One breadth-first round of gather_one_bfs over W slots: a live internal
// slot inside the window is either the top of a child chain (local split >
// 0: recorded in node0 at its octant) or a same-level node, whose two
// children become slots of the next round.  All reads of a round are
// independent.  fetch(x, internal, need_links, ps, lx, lnk) reads BRT node
// x's prefix sum, packed local split (local split | attach octant << 8)
// and, when need_links, its packed child links; it returns false when x is
// unavailable (outside a window), and reads nothing that matters unless
// internal.  CONF: collect the octants of leaf children whose keys carry a
// conflict flag.
template <typename K, int DIM, int W, bool LAST, bool CONF, typename Fetch>
__device__ __forceinline__ void bfs_round(const uint32_t (&sl)[W], uint32_t (&nsl)[2 * W],
                                          Fetch fetch, const K *__restrict__ keys,
                                          const uint32_t *__restrict__ conflict_bits, int lvl,
                                          int mbits, int32_t *node0, uint32_t &cmask,
                                          bool &miss, RunState *st) {
#pragma unroll
  for (int s = 0; s < W; ++s) {
    const uint32_t x = sl[s] & kSlotIdx;
    const bool internal = (sl[s] & (kSlotLive | kSlotLeaf)) == kSlotLive;
    uint32_t psx, lx, v = 0;
    const bool inwin = fetch(x, internal, !LAST, psx, lx, v);
    miss |= internal && !inwin;
    const bool use = internal && inwin;
    const bool term = use && (lx & 0xffu) != 0;
    const bool expand = use && (lx & 0xffu) == 0;
    // the child chain's top, in its octant; otherwise a dummy store to the
    // leaf word, which is written last
    const uint32_t o = term ? lx >> 8 : (uint32_t)(1 << DIM);
    node0[o] = (int32_t)(psx - 1);
    if (CONF && (sl[s] & (kSlotLive | kSlotLeaf)) == (kSlotLive | kSlotLeaf) &&
        test_bit(conflict_bits, x))
      cmask |= 1u << digit_at<K, DIM>(keys[x], lvl, mbits);
    if (LAST) {
      // a well-formed tree has no same-level node this deep; a malformed one
      // the BRT stage did not flag is reported
      if (expand) atomicOr(&st->status, (uint32_t)ST_STACK);
    } else {
      // v: left | left leaf << 30 | right leaf << 31
      const uint32_t a = (v & (kSlotIdx | kSlotLeaf)) | kSlotLive;
      const uint32_t b = ((v & kSlotIdx) + 1) | ((v >> 1) & kSlotLeaf) | kSlotLive;
      nsl[2 * s] = expand ? a : 0u;
      nsl[2 * s + 1] = expand ? b : 0u;
    }
  }
}

// gather_one in breadth-first form: q's nodes node0[0 .. lsq STRIDE) are
// written whole here (children -1 unless found, leaf masks).  The
// same-level BRT nodes below q split on later bits of the digit q splits,
// so they form a binary trie of fewer than DIM levels: DIM rounds over
// fixed slots (2, 4, 8) find every child chain without a loop or stack,
// and every read of a round is independent of the others.  node0 is q's
// lowest node (staged in LDS by the window pass, in the octree for the
// global-memory pass); its leaf mask is read back from the children it
// holds.  lk is q's packed child links; fetch reads the entries of the
// trie's nodes (bfs_round).  q's record is read from brt only for a chain
// (lsq > 1) or the conflict octants (CONF: conflict_bits, keys and
// node_conflict are given).  Returns false when fetch could not read an
// internal node of the trie; the global-memory pass then rewrites q's
// nodes whole.
template <typename K, int DIM, bool CONF, typename Fetch>
__device__ __forceinline__ bool gather_one_bfs(uint32_t q, uint32_t lk,
                                      const BrtNodeT<K> *__restrict__ brt, uint32_t excl,
                                      uint32_t lsq, Fetch fetch,
                                      const K *__restrict__ keys,
                                      const uint32_t *__restrict__ conflict_bits, int mbits,
                                      int32_t *node0, uint32_t *__restrict__ node_conflict,
                                      RunState *st) {
  constexpr int NC = 1 << DIM;
  constexpr int STRIDE = NC + 1;
  constexpr int32_t kFull = (1 << NC) - 1;
  const int lvl = CONF ? brt[q].lcp_length / DIM : 0;
  uint32_t cmask = 0;
  bool miss = false;
#pragma unroll
  for (int c = 0; c < NC; ++c) node0[c] = -1;
  const uint32_t s1[2] = {(lk & (kSlotIdx | kSlotLeaf)) | kSlotLive,
                          ((lk & kSlotIdx) + 1) | ((lk >> 1) & kSlotLeaf) | kSlotLive};
  uint32_t s2[4];
  bfs_round<K, DIM, 2, DIM == 1, CONF>(s1, s2, fetch, keys, conflict_bits, lvl, mbits, node0,
                                       cmask, miss, st);
  uint32_t s3[8];
  bfs_round<K, DIM, 4, DIM == 2, CONF>(s2, s3, fetch, keys, conflict_bits, lvl, mbits, node0,
                                       cmask, miss, st);
  if (DIM == 3) {
    uint32_t s4[16];
    bfs_round<K, DIM, 8, true, CONF>(s3, s4, fetch, keys, conflict_bits, lvl, mbits, node0,
                                     cmask, miss, st);
  }
  if (miss) return false;
  int32_t leaf = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) leaf |= (node0[c] == -1) << c;
  node0[NC] = leaf;
  if (CONF) node_conflict[excl] = cmask;
  if (lsq > 1) {
    const BrtNodeT<K> nq = brt[q];
    for (uint32_t k = 1; k < lsq; ++k) {
      int32_t *nd = node0 + k * STRIDE;
#pragma unroll
      for (int c = 0; c < NC; ++c) nd[c] = -1;
      nd[NC] = kFull;
      if (q != 0) {
        const int o = quadrant<K, DIM>(nq, (int)k - 1);
        nd[o] = (int32_t)(excl + k - 1);
        nd[NC] = kFull & ~(1 << o);
      }
      if (CONF) node_conflict[excl + k] = 0;
    }
  }
  return true;
}

// Appends val to a wave's queue from each lane where want holds (lanes in
// order); entries the queue cannot hold (more than cap - (tail - head)
// queued) are dropped and flag their lane's miss.  Returns the new tail.
// Call with the whole wave converged; tail and head are uniform.
// Entries a queue of the window pass may hold (tests lower it to force the
// overflow path; the queue's own size is the ceiling).
__device__ uint32_t g_gather_queue_cap = 0xffffffffu;

// The queue has size slots (a power of two) and holds at most cap <= size
// entries.
__device__ inline uint32_t queue_push(uint16_t *q, uint32_t size, uint32_t cap, uint32_t head,
                                      uint32_t tail, bool want, uint32_t val, bool &miss) {
  const uint64_t b = __ballot(want);
  const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
  const uint32_t room = cap - (tail - head);
  if (want) {
    if (below < room) q[(tail + below) & (size - 1)] = (uint16_t)val;
    else miss = true;
  }
  const uint32_t n = (uint32_t)__popcll(b);
  return tail + (n < room ? n : room);
}

// The staged window pass's owners without conflict octants: gather_one_bfs
// over a wave at a time, with its rounds' work queued.  Lane l of wave w
// takes the owners a = w 64 + l + i 256 (as the per-lane loop does); the
// first round (the owner's two children) runs per lane, and the internal
// nodes it expands into go to the wave's queue as (window position, owner
// lane) in 16 bits, so that the later rounds spend a lane on each queued
// node rather than fixed 4 and 8 slots per owner (of which about 2.2 hold
// an internal node for uniform points).  A lane reads its entry's owner
// node offset and index from the owner's registers (a cross-lane read).
// The queues live in s_aidx, whose owner indices are read into registers
// first: STAGE / 4 entries per wave.  A node outside the window, or a queue
// full, sends its owner to the global-memory pass (s_miss), which rewrites
// the owner's nodes whole.
template <typename K, int DIM, int R>
__device__ inline void gather_owners_queued(uint32_t nact, uint32_t q0, int w0, int WN,
                                            uint32_t base, uint32_t lo, uint32_t a0,
                                            uint16_t *s_aidx, const uint32_t *s_pl,
                                            const uint32_t *s_lnk, int32_t *s_nodes,
                                            uint32_t *s_miss,
                                            const BrtNodeT<K> *__restrict__ brt, RunState *st) {
  constexpr int NC = 1 << DIM;
  constexpr int STRIDE = NC + 1;
  constexpr int32_t kFull = (1 << NC) - 1;
  constexpr uint32_t CAP = (uint32_t)(R * kBlock) / 4;   // queue entries per wave
  static_assert((CAP & (CAP - 1)) == 0, "queue indices wrap by mask");
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t e_own[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const uint32_t a = (uint32_t)(t + i * kBlock);
    e_own[i] = a < nact ? s_aidx[a] : 0xffffu;
  }
  __syncthreads();   // every owner index is read: s_aidx becomes the queues
  uint16_t *qw = s_aidx + w * CAP;
  const uint32_t cap = min(CAP, g_gather_queue_cap);
  const uint32_t nit = (nact + kBlock - 1) / kBlock;   // uniform
  for (uint32_t i = 0; i < nit; ++i) {
    uint32_t e = 0xffffu;
#pragma unroll
    for (int ii = 0; ii < R; ++ii)
      if ((uint32_t)ii == i) e = e_own[ii];
    const bool has = e != 0xffffu;
    const uint32_t q = q0 + (has ? e : 0u);
    const uint32_t excl_r = q ? s_pl[q - 1 - w0] >> kPlShift : 0u;
    const uint32_t lsq = (s_pl[q - w0] >> kPlShift) - excl_r;
    const uint32_t n0 = a0 + (excl_r - lo) * STRIDE;   // the owner's lowest staged node
    if (has) {
#pragma unroll
      for (int c = 0; c < NC; ++c) s_nodes[n0 + c] = -1;
    }
    const uint32_t lk = has ? s_lnk[q - w0] : 0u;
    bool miss = false;
    // round 1: the owner's two children
    bool want[4];
    uint32_t cand[4];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const uint32_t c = (lk & kSlotIdx) + side;
      const bool leaf = side ? (lk >> 31) != 0u : ((lk >> 30) & 1u) != 0u;
      const bool internal = has && !leaf;
      const unsigned k = (unsigned)((int)c - w0);
      const bool inwin = k < (unsigned)WN;
      miss |= internal && !inwin;
      const bool use = internal && inwin;
      const uint32_t pl = s_pl[use ? k : 0u];
      const uint32_t lx = pl & kPlLsx;
      const bool term = use && (lx & 0xffu) != 0;
      const bool expand = use && (lx & 0xffu) == 0;
      if (has) s_nodes[n0 + (term ? lx >> 8 : (uint32_t)NC)] = (int32_t)(base + (pl >> kPlShift) - 1);
      if (DIM == 1) {
        if (expand) atomicOr(&st->status, (uint32_t)ST_STACK);
        want[2 * side] = want[2 * side + 1] = false;
        cand[2 * side] = cand[2 * side + 1] = 0u;
      } else {
        const uint32_t v = s_lnk[use ? k : 0u];
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
          const uint32_t cc = (v & kSlotIdx) + ch;
          const bool cleaf = ch ? (v >> 31) != 0u : ((v >> 30) & 1u) != 0u;
          const bool cint = expand && !cleaf;
          const unsigned kk = (unsigned)((int)cc - w0);
          const bool cin = kk < (unsigned)WN;
          miss |= cint && !cin;
          want[2 * side + ch] = cint && cin;
          cand[2 * side + ch] = kk | ((uint32_t)lane << 10);
        }
      }
    }
    uint32_t head = 0, tail = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) tail = queue_push(qw, CAP, cap, head, tail, want[j], cand[j], miss);
    // rounds 2 .. DIM over the queued nodes, a generation at a time
    uint32_t glo = 0, ghi = tail;
#pragma unroll
    for (int r = 2; r <= DIM; ++r) {
      for (uint32_t b0 = glo; b0 < ghi; b0 += 64) {
        const uint32_t idx = b0 + (uint32_t)lane;
        const bool valid = idx < ghi;
        const uint32_t ent = valid ? (uint32_t)qw[idx & (CAP - 1)] : 0u;
        head = min(b0 + 64u, ghi);   // this chunk's entries are in registers
        const uint32_t k = ent & 1023u, ol = ent >> 10;
        const uint32_t n0o = (uint32_t)__shfl((int)n0, (int)ol, 64);
        const uint32_t eo = (uint32_t)__shfl((int)e, (int)ol, 64);
        const uint32_t pl = s_pl[k];
        const uint32_t lx = pl & kPlLsx;
        const bool term = valid && (lx & 0xffu) != 0;
        const bool expand = valid && (lx & 0xffu) == 0;
        if (valid) s_nodes[n0o + (term ? lx >> 8 : (uint32_t)NC)] = (int32_t)(base + (pl >> kPlShift) - 1);
        bool m2 = false;
        if (r == DIM) {
          // a well-formed tree has no same-level node this deep
          if (expand) atomicOr(&st->status, (uint32_t)ST_STACK);
        } else {
          const uint32_t v = s_lnk[k];
          bool wnt[2];
          uint32_t cd[2];
#pragma unroll
          for (int ch = 0; ch < 2; ++ch) {
            const uint32_t cc = (v & kSlotIdx) + ch;
            const bool cleaf = ch ? (v >> 31) != 0u : ((v >> 30) & 1u) != 0u;
            const bool cint = expand && !cleaf;
            const unsigned kk = (unsigned)((int)cc - w0);
            const bool cin = kk < (unsigned)WN;
            m2 |= cint && !cin;
            wnt[ch] = cint && cin;
            cd[ch] = kk | (ol << 10);
          }
#pragma unroll
          for (int ch = 0; ch < 2; ++ch)
            tail = queue_push(qw, CAP, cap, head, tail, wnt[ch], cd[ch], m2);
        }
        if (m2) atomicOr(&s_miss[eo >> 5], 1u << (eo & 31));
      }
      glo = ghi;
      ghi = tail;
    }
    if (!has) continue;
    if (miss) atomicOr(&s_miss[e >> 5], 1u << (e & 31));
    // the leaf mask from the children the rounds wrote, then the chain
    int32_t leafm = 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) leafm |= (s_nodes[n0 + c] == -1) << c;
    s_nodes[n0 + NC] = leafm;
    if (lsq > 1) {
      const BrtNodeT<K> nq = brt[q];
      const uint32_t excl = base + excl_r;
      for (uint32_t kq = 1; kq < lsq; ++kq) {
        int32_t *nd = s_nodes + n0 + kq * STRIDE;
#pragma unroll
        for (int c = 0; c < NC; ++c) nd[c] = -1;
        nd[NC] = kFull;
        if (q != 0) {
          const int o = quadrant<K, DIM>(nq, (int)kq - 1);
          nd[o] = (int32_t)(excl + kq - 1);
          nd[NC] = kFull & ~(1 << o);
        }
      }
    }
  }
}

// BRT entries (links, ps, lsx) staged either side of a workgroup's nodes.
constexpr int kGatherHalo = 256;
// BRT nodes per workgroup of the single-tree gather: kGatherRows rows of 256.
#ifndef PGVD_GATHER_ROWS
#define PGVD_GATHER_ROWS 2
#endif
constexpr int kGatherRows = PGVD_GATHER_ROWS;
// Staged octree nodes per workgroup, in percent of its BRT nodes: 58% of
// 512 (296 nodes, a uniform window owns ~240) keeps the LDS at 20.0 KB, so
// that eight workgroups fit a CU (62% gave 20.8 KB and seven).
#ifndef PGVD_GATHER_STAGE_PCT
#define PGVD_GATHER_STAGE_PCT 58
#endif
constexpr int kGatherStagePct = PGVD_GATHER_STAGE_PCT;

// Gather, first pass over the BRT nodes [q0, q0+R*256) of one tree with m
// keys (m >= 2).  The workgroup stages links, ps and lsx of BRT nodes
// [q0-H, q0+R*256+H) in LDS; a node whose descent leaves that window is
// handed to defer(true, q) and left to the global-memory pass (whose writes
// land after this kernel's, so the window pass's placeholder output for it
// is overwritten).  Every lane calls defer once per row, converged.  The
// octree nodes the workgroup's BRT nodes own are the contiguous range
// [ps[q0-1], ps[q_last]); when it holds at most R*256 nodes it is built in
// LDS and written out in coalesced runs.
// With chunk_excl (exclusive prefix sums of the local splits per chunk of
// 256 nodes), the window's prefix sums are computed here from lsx instead
// of read from ps, and the workgroup writes them to ps_out for its own
// nodes (and for node m-1 when that starts the next chunk).
template <typename K, int DIM, int R, bool CONF, typename Defer>
__device__ inline void gather_window(const BrtNodeT<K> *__restrict__ brt,
                                     const uint32_t *__restrict__ ps,
                                     const uint32_t *__restrict__ lsx,
                                     const uint32_t *__restrict__ chunk_excl,
                                     uint32_t *__restrict__ ps_out,
                                     const K *__restrict__ keys,
                                     const uint32_t *__restrict__ conflict_bits, int mbits,
                                     uint32_t m, uint32_t q0, int32_t *__restrict__ octree,
                                     uint32_t *__restrict__ node_conflict, RunState *st,
                                     Defer defer) {
  constexpr int NC = 1 << DIM;
  constexpr int STRIDE = NC + 1;
  constexpr int H = kGatherHalo, WR = R + 2, WN = WR * kBlock, STAGE = R * kBlock;
  // octree nodes staged in LDS: a window of STAGE BRT nodes owns ~0.47 STAGE
  // of them for uniform keys; a larger span is written straight to HBM
  constexpr int SN = STAGE * kGatherStagePct / 100;
  static_assert(H == kBlock, "the window is whole chunks of 256");
  // staged octree words, 16-byte aligned, plus up to 3 words in front so
  // that LDS and global words agree modulo 4 (16-byte copies out)
  __shared__ __align__(16) int32_t s_nodes[(SN * STRIDE + 3 + 3) & ~3];
  // per window position: packed links, and the inclusive prefix sum of
  // local splits (relative to the window's base) beside the packed local
  // split in one word, so that one read gets both
  __shared__ uint32_t s_lnk[WN], s_pl[WN];
  const int t = threadIdx.x;
  const int w0 = (int)q0 - H;
  // Every load of the window is issued before the first is used: addresses
  // outside [0, m) are clamped to entry 0 and the values dropped.
  uint32_t lx[WR], lk[WR], pv[WR];
#pragma unroll
  for (int r = 0; r < WR; ++r) {
    const int g = w0 + r * kBlock + t;
    const bool in = g >= 0 && g < (int)m;
    lx[r] = lsx[in ? g : 0];
    lk[r] = load_links(brt + (in && g < (int)m - 1 ? g : 0));
    if (!chunk_excl) pv[r] = ps[in ? g : 0];
  }
  uint32_t base;   // the prefix sum before the window
  if (chunk_excl) {
    // The window is whole chunks (q0 is a multiple of 256), so the base of
    // the first one plus a running scan gives every prefix sum.
    __shared__ uint32_t s_w[WR * 4];
    base = q0 ? chunk_excl[(q0 >> 8) - 1] : 0u;
    uint32_t v[WR], ex[WR];
#pragma unroll
    for (int r = 0; r < WR; ++r) {
      const int g = w0 + r * kBlock + t;
      if (g < 0 || g >= (int)m) lx[r] = 0;
      v[r] = lx[r] & 0xffu;
    }
    (void)striped_excl_scan<WR, uint32_t>(v, ex, s_w);
#pragma unroll
    for (int r = 0; r < WR; ++r) pv[r] = base + ex[r] + v[r];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const uint32_t q = q0 + j * kBlock + t;
      if (q < m) ps_out[q] = pv[1 + j];
    }
    if (t == 0 && q0 + STAGE == m - 1) ps_out[m - 1] = pv[R + 1];
  } else {
    base = w0 > 0 ? ps[w0 - 1] : 0u;
  }
#pragma unroll
  for (int r = 0; r < WR; ++r) {
    const int k = r * kBlock + t;
    s_pl[k] = (pv[r] - base) << kPlShift | (lx[r] & kPlLsx);
    s_lnk[k] = lk[r];
  }
  __syncthreads();
  const uint32_t qlast = min(q0 + STAGE - 1, m - 2);
  // prefix sums below are relative to base (they fit 21 bits: a window's
  // 1024 positions own at most 1024 * 42 octree nodes)
  const uint32_t lo = q0 ? s_pl[q0 - 1 - w0] >> kPlShift : 0u;
  const uint32_t span = (s_pl[qlast - w0] >> kPlShift) - lo;
  const bool staged = span <= (uint32_t)SN;   // uniform over the workgroup
  const uint32_t a0 = (uint32_t)(((uint64_t)(base + lo) * STRIDE) & 3u);
  // The nodes that own octree nodes (about half), compacted so that each
  // lane takes one: entry e = j * 256 + t of the rows.
  __shared__ uint16_t s_aidx[STAGE];
  __shared__ uint32_t s_miss[STAGE / 32], s_w4[4];
  if (t < STAGE / 32) s_miss[t] = 0;
  uint32_t amask = 0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const uint32_t q = q0 + j * kBlock + t;
    bool active = q <= qlast;
    if (active) {
      const uint32_t excl = q ? s_pl[q - 1 - w0] >> kPlShift : 0u;
      active = (q == 0) || (s_pl[q - w0] >> kPlShift) > excl;
    }
    amask |= (uint32_t)active << j;
  }
  uint32_t nact;
  uint32_t pos = block_excl_sum<uint32_t>((uint32_t)__popc(amask), s_w4, &nact);
#pragma unroll
  for (int j = 0; j < R; ++j)
    if ((amask >> j) & 1u) s_aidx[pos++] = (uint16_t)(j * kBlock + t);
  __syncthreads();
  if (!CONF && staged) {
    gather_owners_queued<K, DIM, R>(nact, q0, w0, WN, base, lo, a0, s_aidx, s_pl, s_lnk, s_nodes,
                                    s_miss, brt, st);
  } else {
  for (uint32_t a = t; a < nact; a += kBlock) {
    const uint32_t e = s_aidx[a];
    const uint32_t q = q0 + e;
    const uint32_t excl_r = q ? s_pl[q - 1 - w0] >> kPlShift : 0u;
    const uint32_t lsq = (s_pl[q - w0] >> kPlShift) - excl_r;
    const uint32_t excl = base + excl_r;
    bool ok;
    if (staged) {
      auto fetch = [&](uint32_t x, bool internal, bool need_links, uint32_t &psx, uint32_t &lx,
                       uint32_t &lnk) -> bool {
        const unsigned k = (unsigned)((int)x - w0);
        const bool inwin = k < (unsigned)WN;
        const unsigned kk = internal && inwin ? k : 0u;
        const uint32_t pl = s_pl[kk];   // prefix sum - base, packed local split
        psx = base + (pl >> kPlShift);
        lx = pl & kPlLsx;
        if (need_links) lnk = s_lnk[kk];
        return inwin;
      };
      ok = gather_one_bfs<K, DIM, CONF>(q, s_lnk[q - w0], brt, excl, lsq, fetch, keys,
                                        conflict_bits,
                                  mbits,
                                  s_nodes + a0 + (excl_r - lo) * STRIDE,
                                  node_conflict, st);
    } else {
      auto links = [&](uint32_t x, uint32_t &left, bool &ll, bool &rl) -> bool {
        const unsigned k = (unsigned)((int)x - w0);
        if (k >= (unsigned)WN) return false;
        const uint32_t v = s_lnk[k];
        left = v & 0x3fffffffu;
        ll = (v >> 30) & 1u;
        rl = v >> 31;
        return true;
      };
      auto child = [&](uint32_t c, uint32_t &psc, uint32_t &lxc) -> bool {
        const unsigned k = (unsigned)((int)c - w0);
        if (k >= (unsigned)WN) return false;
        psc = base + (s_pl[k] >> kPlShift);
        lxc = s_pl[k] & kPlLsx;
        return true;
      };
      auto put = [&](uint32_t idx, const int32_t *c) { write_node<DIM>(octree, idx, c); };
      ok = gather_one<K, DIM>(q, brt[q], excl, lsq, links, child, keys, conflict_bits, mbits,
                              put, node_conflict, st);
    }
    if (!ok) atomicOr(&s_miss[e >> 5], 1u << (e & 31));
  }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const uint32_t e = j * kBlock + t;
    defer(((s_miss[e >> 5] >> (e & 31)) & 1u) != 0u, q0 + e);
  }
  if (staged) {
    // LDS words [a0, a0 + words) -> octree words [lo STRIDE, ..): whole
    // 16-byte chunks as one store, the partial first and last chunks by word
    const uint32_t words = span * STRIDE;
    int32_t *out = octree + (size_t)(base + lo) * STRIDE - a0;   // out[i] <-> s_nodes[i]
    const uint32_t end = a0 + words, c0 = a0 >> 2, c1 = (end + 3) >> 2;
    for (uint32_t c = c0 + t; c < c1; c += kBlock) {
      const uint32_t i = c * 4;
      if (i >= a0 && i + 4 <= end) {
        *reinterpret_cast<int4 *>(out + i) = *reinterpret_cast<const int4 *>(s_nodes + i);
      } else {
        for (uint32_t k = max(i, a0); k < min(i + 4, end); ++k) out[k] = s_nodes[k];
      }
    }
  }
}

template <int DIM>
__device__ inline void single_node(int32_t *octree, uint32_t *node_conflict) {
  constexpr int NC = 1 << DIM;
  int32_t ch[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) ch[c] = -1;
  write_node<DIM>(octree, 0, ch);
  if (node_conflict) node_conflict[0] = 0;
}

// Nodes the window cannot resolve are marked in miss_words (one word per 64
// BRT nodes) for the global-memory pass.
// With chunk_excl the prefix sums ps[0..m) are written here (see
// gather_window); without it they are read.
// Waves per SIMD the window pass is compiled for (its register budget):
// eight, with the LDS of eight workgroups per CU, ran it 5% faster than
// seven although some SGPRs then spill to VGPR lanes.
#ifndef PGVD_GATHER_WPE
#define PGVD_GATHER_WPE 8
#endif
template <typename K, int DIM, bool CONF>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(PGVD_GATHER_WPE))) void
gather_kernel(
    const BrtNodeT<K> *__restrict__ brt, uint32_t *__restrict__ ps,
    const uint32_t *__restrict__ lsx, const uint32_t *__restrict__ chunk_excl,
    const K *__restrict__ keys, const uint32_t *__restrict__ conflict_bits, int mbits,
    int32_t *__restrict__ octree, uint32_t *__restrict__ node_conflict, int64_t capacity,
    uint64_t *__restrict__ miss_words, RunState *st) {
  if (st->malformed) return;
  const uint32_t m = st->m;
  const uint32_t q0 = xcd_segment(blockIdx.x, gridDim.x) * (kGatherRows * kBlock);
  if ((int64_t)st->octree_size > capacity) {
    if (q0 == 0 && threadIdx.x == 0) atomicOr(&st->status, (uint32_t)ST_CAPACITY);
    return;
  }
  if (m == 1) {             // single unique key: one node, no children
    if (q0 == 0 && threadIdx.x == 0) {
      single_node<DIM>(octree, node_conflict);
      if (chunk_excl) ps[0] = 1;
    }
    return;
  }
  if (q0 >= m - 1) return;
  gather_window<K, DIM, kGatherRows, CONF>(brt, ps, lsx, chunk_excl, ps, keys, conflict_bits,
                                           mbits, m,
                                     q0, octree,
                        node_conflict, st,
                        [&](bool miss, uint32_t q) { wave_mask(miss, miss_words, q >> 6); });
}

// Batched form: gather tiles of 256 key slots per instance; instance s writes
// its nodes at octree[oct_off[s]..) with instance-local child indices.
template <typename K, int DIM>
__global__ __launch_bounds__(kBlock) void gather_seg_kernel(
    const BrtNodeT<K> *__restrict__ brt, const uint32_t *__restrict__ ps,
    const uint32_t *__restrict__ lsx, const SegInfo *__restrict__ seg, TileTable tt,
    const SegState *__restrict__ ss, int32_t *__restrict__ octree, int64_t capacity,
    uint32_t *__restrict__ defer, RunState *st) {
  constexpr int STRIDE = (1 << DIM) + 1;
  if ((int64_t)st->octree_size > capacity) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&st->status, (uint32_t)ST_CAPACITY);
    return;
  }
  const uint32_t tile = xcd_segment(blockIdx.x, gridDim.x);
  const uint32_t s = tt.seg[tile];
  const SegState sst = ss[s];
  const uint32_t m = sst.m;
  const uint32_t q0 = (tile - tt.first[s]) * kBlock;
  int32_t *oct = octree + (size_t)sst.oct_off * STRIDE;
  if (m == 1) {
    if (q0 == 0 && threadIdx.x == 0) single_node<DIM>(oct, nullptr);
    return;
  }
  if (q0 >= m - 1) return;
  const uint32_t kofs = seg[s].kofs;
  gather_window<K, DIM, 1, false>(brt + kofs, ps + kofs, lsx + kofs, nullptr, nullptr,
                        (const K *)nullptr, nullptr, 0, m, q0,
                        oct, nullptr, st, [&](bool miss, uint32_t q) {
                          wave_append(miss, kofs + q, &st->counters[CTR_OCT_DEFER], defer);
                        });
}

// Gather, second pass: the BRT nodes marked in miss_words, reading global
// memory, breadth first (a round's loads are independent, so a node waits
// on DIM + 1 dependent loads, not one per trie node).
template <typename K, int DIM, bool CONF>
__global__ __launch_bounds__(kBlock) void gather_defer_kernel(
    const BrtNodeT<K> *__restrict__ brt, const uint32_t *__restrict__ ps,
    const uint32_t *__restrict__ lsx, const K *__restrict__ keys,
    const uint32_t *__restrict__ conflict_bits, int mbits, int32_t *__restrict__ octree,
    uint32_t *__restrict__ node_conflict, int64_t capacity,
    const uint64_t *__restrict__ miss_words, RunState *st) {
  constexpr int STRIDE = (1 << DIM) + 1;
  if (st->malformed || (int64_t)st->octree_size > capacity) return;
  const uint32_t m = st->m;
  const uint32_t nwords = m > 1 ? (m - 1 + 63) / 64u : 0u;
  // x of an internal slot is a BRT node (< m - 1); other slots read nothing
  auto fetch = [&](uint32_t x, bool internal, bool need_links, uint32_t &psx, uint32_t &lx,
                   uint32_t &lnk) -> bool {
    psx = 0;
    lx = 0;
    if (internal) {
      psx = ps[x];
      lx = lsx[x];
      if (need_links) lnk = load_links(brt + x);
    }
    return true;
  };
  for_each_masked(miss_words, nwords, [&](uint32_t q) {
    const uint32_t excl = q ? ps[q - 1] : 0u;
    (void)gather_one_bfs<K, DIM, CONF>(q, load_links(brt + q), brt, excl, ps[q] - excl, fetch,
                                       keys, conflict_bits, mbits,
                                       octree + (size_t)excl * STRIDE, node_conflict, st);
  });
}

template <typename K, int DIM>
__global__ __launch_bounds__(kBlock) void gather_seg_defer_kernel(
    const BrtNodeT<K> *__restrict__ brt, const uint32_t *__restrict__ ps,
    const uint32_t *__restrict__ lsx, const uint32_t *__restrict__ kofs_all, uint32_t B,
    const SegState *__restrict__ ss, int32_t *__restrict__ octree, int64_t capacity,
    const uint32_t *__restrict__ defer, RunState *st) {
  constexpr int STRIDE = (1 << DIM) + 1;
  if ((int64_t)st->octree_size > capacity) return;
  const uint32_t cnt = st->counters[CTR_OCT_DEFER];
  for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < cnt; k += gridDim.x * kBlock) {
    const uint32_t slot = defer[k];
    const uint32_t s = find_seg(kofs_all, B, slot);
    const uint32_t kofs = kofs_all[s];
    const uint32_t q = slot - kofs;
    const BrtNodeT<K> *b = brt + kofs;
    const uint32_t *p = ps + kofs;
    const uint32_t *lx = lsx + kofs;
    int32_t *oct = octree + (size_t)ss[s].oct_off * STRIDE;
    auto fetch = [&](uint32_t x, bool internal, bool need_links, uint32_t &psx, uint32_t &lxx,
                     uint32_t &lnk) -> bool {
      psx = 0;
      lxx = 0;
      if (internal) {
        psx = p[x];
        lxx = lx[x];
        if (need_links) lnk = load_links(b + x);
      }
      return true;
    };
    const uint32_t excl = q ? p[q - 1] : 0u;
    (void)gather_one_bfs<K, DIM, false>(q, load_links(b + q), b, excl, p[q] - excl, fetch,
                                 (const K *)nullptr, nullptr, 0, oct + (size_t)excl * STRIDE,
                                 nullptr, st);
  }
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void scatter_init_kernel(int32_t *__restrict__ octree,
                                                              uint32_t *__restrict__ node_conflict,
                                                              int64_t capacity, int force,
                                                              RunState *st) {
  if (!force && !st->malformed) return;
  constexpr int NC = 1 << DIM;
  const uint32_t O = st->octree_size;
  if ((int64_t)O > capacity) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&st->status, (uint32_t)ST_CAPACITY);
    return;
  }
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < O; i += gridDim.x * kBlock) {
    int32_t *o = octree + (size_t)i * (NC + 1);
#pragma unroll
    for (int c = 0; c < NC; ++c) o[c] = -1;
    o[NC] = (1 << NC) - 1;
    if (node_conflict) node_conflict[i] = 0;
  }
}

template <typename K, int DIM>
__device__ inline void scatter_one(uint32_t i, uint32_t m, const BrtNodeT<K> *__restrict__ brt,
                                   const int32_t *__restrict__ parent,
                                   const uint32_t *__restrict__ ps, const K *__restrict__ keys,
                                   const uint32_t *__restrict__ conflict_bits, int mbits,
                                   int32_t *__restrict__ octree,
                                   uint32_t *__restrict__ node_conflict, RunState *st);

// parent[] is in the reference form here (root -1, unclaimed 0).
template <typename K, int DIM>
__global__ __launch_bounds__(kBlock) void scatter_kernel(
    const BrtNodeT<K> *__restrict__ brt, const int32_t *__restrict__ parent,
    const uint32_t *__restrict__ ps, const K *__restrict__ keys,
    const uint32_t *__restrict__ conflict_bits, int mbits, int32_t *__restrict__ octree,
    uint32_t *__restrict__ node_conflict, int64_t capacity, int force, RunState *st) {
  if (!force && !st->malformed) return;
  const uint32_t m = st->m;
  if ((int64_t)st->octree_size > capacity) return;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i + 1 < m; i += gridDim.x * kBlock)
    scatter_one<K, DIM>(i, m, brt, parent, ps, keys, conflict_bits, mbits, octree,
                        node_conflict, st);
}

template <typename K, int DIM>
__device__ inline void scatter_one(uint32_t i, uint32_t m, const BrtNodeT<K> *__restrict__ brt,
                                   const int32_t *__restrict__ parent,
                                   const uint32_t *__restrict__ ps, const K *__restrict__ keys,
                                   const uint32_t *__restrict__ conflict_bits, int mbits,
                                   int32_t *__restrict__ octree,
                                   uint32_t *__restrict__ node_conflict, RunState *st) {
  constexpr int NC = 1 << DIM;
  const BrtNodeT<K> ni = brt[i];
  // Bound on every parent walk: a walk visits each BRT node at most once.
  const uint32_t bound = m;
  if (conflict_bits) {
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const bool leaf = side ? ni.right_leaf : ni.left_leaf;
      const uint32_t c = (uint32_t)ni.left + side;
      if (!leaf || !test_bit(conflict_bits, c)) continue;
      uint32_t q = i, steps = 0;
      while (q != 0 && ls_of(ps, q) == 0 && steps < bound) { q = (uint32_t)parent[q]; ++steps; }
      if (steps >= bound) { atomicOr(&st->status, (uint32_t)ST_WALK); continue; }
      const uint32_t T = q == 0 ? 0u : ps[q - 1];
      const int lvl = brt[q].lcp_length / DIM;
      atomicOr(&node_conflict[T], 1u << digit_at<K, DIM>(keys[c], lvl, mbits));
    }
  }
  if (i == 0) return;
  const uint32_t lsi = ps[i] - ps[i - 1];
  if (lsi == 0) return;
  uint32_t cur = ps[i - 1];
  for (uint32_t k = 0; k + 1 < lsi; ++k) {
    const uint32_t op = cur + 1;
    const int o = quadrant<K, DIM>(ni, (int)k);
    octree[(size_t)op * (NC + 1) + o] = (int32_t)cur;
    octree[(size_t)op * (NC + 1) + NC] &= ~(1 << o);
    cur = op;
  }
  uint32_t p = (uint32_t)parent[i], steps = 0;
  while (ls_of(ps, p) == 0 && steps < bound) { p = (uint32_t)parent[p]; ++steps; }
  if (steps >= bound) { atomicOr(&st->status, (uint32_t)ST_WALK); return; }
  const uint32_t T = p == 0 ? 0u : ps[p - 1];
  const int o = quadrant<K, DIM>(ni, (int)lsi - 1);
  atomicMax(&octree[(size_t)T * (NC + 1) + o], (int32_t)cur);
  atomicAnd(&octree[(size_t)T * (NC + 1) + NC], ~(1 << o));
}

template <typename K, int DIM>
void launch_dim(const BrtNodeT<K> *brt, const int32_t *parent, uint32_t *ps,
                const uint32_t *lsx, const uint32_t *chunk_excl, const K *keys,
                const uint32_t *conflict_bits, int64_t max_m, int mbits, int32_t *octree,
                uint32_t *node_conflict, int64_t capacity, int force_scatter, uint32_t *defer,
                RunState *st, hipStream_t s, hipEvent_t *gather_ev) {
  const int64_t nodes = max_m > 1 ? max_m - 1 : 1;
  const unsigned grid = (unsigned)((nodes + kBlock - 1) / kBlock);
  const unsigned sgrid = grid < 1024 ? grid : 1024;
  if (force_scatter != 1) {
    uint64_t *words = reinterpret_cast<uint64_t *>(defer);
    const unsigned ggrid =
        (unsigned)((nodes + kGatherRows * kBlock - 1) / (kGatherRows * kBlock));
    if (gather_ev) (void)hipEventRecord(gather_ev[0], s);
    // the conflict octants are compiled in only for builds with labels
    if (conflict_bits)
      gather_kernel<K, DIM, true><<<ggrid, kBlock, 0, s>>>(brt, ps, lsx, chunk_excl, keys,
                                                           conflict_bits, mbits, octree,
                                                           node_conflict, capacity, words, st);
    else
      gather_kernel<K, DIM, false><<<ggrid, kBlock, 0, s>>>(brt, ps, lsx, chunk_excl, keys,
                                                            nullptr, mbits, octree, nullptr,
                                                            capacity, words, st);
    if (gather_ev) (void)hipEventRecord(gather_ev[1], s);
    const int64_t g2 = ((nodes + 63) / 64 + kBlock - 1) / kBlock;
    const unsigned dgrid = (unsigned)(g2 < 4096 ? g2 : 4096);
    if (conflict_bits)
      gather_defer_kernel<K, DIM, true><<<dgrid, kBlock, 0, s>>>(brt, ps, lsx, keys, conflict_bits,
                                                                 mbits, octree, node_conflict,
                                                                 capacity, words, st);
    else
      gather_defer_kernel<K, DIM, false><<<dgrid, kBlock, 0, s>>>(brt, ps, lsx, keys, nullptr,
                                                                  mbits, octree, nullptr,
                                                                  capacity, words, st);
    if (force_scatter == 2) return;
  }
  // The scatter path is a no-op unless forced or st->malformed; its kernels
  // loop over their ranges, so a capped grid keeps the no-op cheap.
  scatter_init_kernel<DIM><<<sgrid, kBlock, 0, s>>>(octree, node_conflict, capacity,
                                                    force_scatter, st);
  scatter_kernel<K, DIM><<<sgrid, kBlock, 0, s>>>(brt, parent, ps, keys, conflict_bits, mbits,
                                                  octree, node_conflict, capacity, force_scatter,
                                                  st);
}

// ComputeLocalSplits (C/BuildOctree.c:17-27) as the reference runs it: node
// i writes its internal children's entries; when two nodes claim one child
// the serial loop's last (highest) writer wins, reproduced with a 64-bit
// atomicMax of ((i+1) << 32 | value).  tmp[] must be zero.
template <typename K>
__global__ __launch_bounds__(kBlock) void local_splits_claim_kernel(
    const BrtNodeT<K> *__restrict__ brt, int64_t m, int dim, unsigned long long *tmp) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= m - 1) return;
  const BrtNodeT<K> b = brt[i];
  const uint32_t local = b.lcp_length / dim;
  const unsigned long long tag = (unsigned long long)(i + 1) << 32;
  if (!b.left_leaf) {
    const uint32_t v = brt[b.left].lcp_length / dim - local;
    atomicMax(&tmp[b.left], tag | v);
  }
  if (!b.right_leaf) {
    const uint32_t v = brt[b.left + 1].lcp_length / dim - local;
    atomicMax(&tmp[b.left + 1], tag | v);
  }
}

template <typename K>
__global__ __launch_bounds__(kBlock) void local_splits_finish_kernel(
    const BrtNodeT<K> *__restrict__ brt, int64_t m, int dim, const unsigned long long *tmp,
    uint32_t *ls) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= m) return;
  if (c == 0) ls[0] = (m == 1) ? 1u : 1u + brt[0].lcp_length / dim;
  else ls[c] = (uint32_t)(tmp[c] & 0xFFFFFFFFull);
}

}  // namespace

// Test hook: the window pass's queues hold at most cap entries (0: their
// full size).  Synchronous; every device the process uses takes it.
int set_gather_queue_cap(uint32_t cap) {
  const uint32_t v = cap ? cap : 0xffffffffu;
  return hipMemcpyToSymbol(HIP_SYMBOL(g_gather_queue_cap), &v, sizeof v) == hipSuccess ? 0 : 1;
}


template <typename K>
void launch_local_splits(const BrtNodeT<K> *brt, int64_t m, int dim, uint32_t *ls, uint64_t *tmp,
                         hipStream_t s) {
  const unsigned grid = (unsigned)((m + kBlock - 1) / kBlock);
  unsigned long long *t = reinterpret_cast<unsigned long long *>(tmp);
  local_splits_claim_kernel<K><<<grid, kBlock, 0, s>>>(brt, m, dim, t);
  local_splits_finish_kernel<K><<<grid, kBlock, 0, s>>>(brt, m, dim, t, ls);
}

template <typename K>
void launch_octree(const BrtNodeT<K> *brt, const int32_t *parent, uint32_t *ps,
                   const uint32_t *lsx, const uint32_t *chunk_excl, const K *keys,
                   const uint32_t *conflict_bits, int64_t max_m, int dim, int mbits,
                   int32_t *octree, uint32_t *node_conflict, int64_t capacity,
                   int force_scatter, uint32_t *defer, RunState *st, hipStream_t s,
                   hipEvent_t *gather_ev) {
  if (dim == 2)
    launch_dim<K, 2>(brt, parent, ps, lsx, chunk_excl, keys, conflict_bits, max_m, mbits,
                     octree, node_conflict, capacity, force_scatter, defer, st, s, gather_ev);
  else
    launch_dim<K, 3>(brt, parent, ps, lsx, chunk_excl, keys, conflict_bits, max_m, mbits,
                     octree, node_conflict, capacity, force_scatter, defer, st, s, gather_ev);
}

template <typename K, int DIM>
void gather_seg_dim(const BrtNodeT<K> *brt, const uint32_t *ps, const uint32_t *lsx,
                    const BatchDev &b, int32_t *octree, int64_t capacity, uint32_t *defer,
                    RunState *st, hipStream_t s) {
  gather_seg_kernel<K, DIM><<<(unsigned)b.n_node, kBlock, 0, s>>>(brt, ps, lsx, b.seg, b.node, b.ss,
                                                                  octree, capacity, defer, st);
  gather_seg_defer_kernel<K, DIM><<<2048, kBlock, 0, s>>>(brt, ps, lsx, b.kofs, (uint32_t)b.B,
                                                          b.ss, octree, capacity, defer, st);
}

template <typename K>
void launch_gather_seg(const BrtNodeT<K> *brt, const uint32_t *ps, const uint32_t *lsx, int dim,
                       const BatchDev &b, int32_t *octree, int64_t capacity, uint32_t *defer,
                       RunState *st, hipStream_t s) {
  if (b.n_node <= 0) return;
  if (dim == 2) gather_seg_dim<K, 2>(brt, ps, lsx, b, octree, capacity, defer, st, s);
  else gather_seg_dim<K, 3>(brt, ps, lsx, b, octree, capacity, defer, st, s);
}

#define PGVD_INST(K)                                                                          \
  template void launch_octree<K>(const BrtNodeT<K> *, const int32_t *, uint32_t *,             \
                                 const uint32_t *, const uint32_t *, const K *,                \
                                 const uint32_t *, int64_t, int, int, int32_t *, uint32_t *,   \
                                 int64_t, int, uint32_t *, RunState *, hipStream_t,            \
                                 hipEvent_t *);                                                \
  template void launch_gather_seg<K>(const BrtNodeT<K> *, const uint32_t *, const uint32_t *,  \
                                     int, const BatchDev &, int32_t *, int64_t, uint32_t *,    \
                                     RunState *, hipStream_t);                                 \
  template void launch_local_splits<K>(const BrtNodeT<K> *, int64_t, int, uint32_t *,          \
                                       uint64_t *, hipStream_t);
PGVD_INST(uint32_t)
PGVD_INST(uint64_t)
PGVD_INST(Key128)
#undef PGVD_INST

}  // namespace pgvd
