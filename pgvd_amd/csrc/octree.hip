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

namespace pgvd {

namespace {

template <typename K, int DIM>
__device__ inline int quadrant(const BrtNodeT<K> &b, int i) {
  const int rshift = i * DIM + b.lcp_length % DIM;
  return (int)(klow32(kshr((K)b.lcp, rshift)) & ((1u << DIM) - 1));
}

// Octant of `key` at octree level lvl (the DIM bits right below a prefix of
// lvl digits).
template <typename K, int DIM>
__device__ inline int digit_at(K key, int lvl, int mbits) {
  const int sh = mbits - (lvl + 1) * DIM;
  return sh < 0 ? 0 : (int)(klow32(kshr(key, sh)) & ((1u << DIM) - 1));
}

__device__ inline bool test_bit(const uint32_t *bits, uint32_t i) {
  return (bits[i >> 5] >> (i & 31)) & 1u;
}

__device__ inline uint32_t ls_of(const uint32_t *ps, uint32_t c) {
  return c == 0 ? ps[0] : ps[c] - ps[c - 1];
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

// Octree nodes staged per workgroup: the nodes owned by BRT nodes
// [q0, q0+256) are the contiguous range [ps[q0-1], ps[q0+255]); when it fits
// here it is built in LDS and written out in coalesced runs.
constexpr int kStageNodes = 256;   // small: this kernel is latency bound

template <typename K, int DIM>
__global__ __launch_bounds__(kBlock) void gather_kernel(
    const BrtNodeT<K> *__restrict__ brt, const uint32_t *__restrict__ ps,
    const K *__restrict__ keys, const uint32_t *__restrict__ conflict_bits, int mbits,
    int32_t *__restrict__ octree, uint32_t *__restrict__ node_conflict, int64_t capacity,
    RunState *st) {
  if (st->malformed) return;
  constexpr int NC = 1 << DIM;
  constexpr int STRIDE = NC + 1;
  __shared__ int32_t s_nodes[kStageNodes * STRIDE];
  const uint32_t m = st->m;
  const uint32_t O = st->octree_size;
  const uint32_t q0 = blockIdx.x * kBlock;
  const uint32_t q = q0 + threadIdx.x;
  if ((int64_t)O > capacity) {
    if (q == 0) atomicOr(&st->status, (uint32_t)ST_CAPACITY);
    return;
  }
  int32_t ch[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) ch[c] = -1;
  if (m == 1) {             // single unique key: one node, no children
    if (q == 0) {
      write_node<DIM>(octree, 0, ch);
      if (node_conflict) node_conflict[0] = 0;
    }
    return;
  }
  if (q0 >= m - 1) return;
  const uint32_t qlast = min(q0 + kBlock - 1, m - 2);
  const uint32_t lo = q0 ? ps[q0 - 1] : 0u;
  const uint32_t span = ps[qlast] - lo;
  const bool staged = span <= (uint32_t)kStageNodes;   // uniform over the block
  auto put = [&](uint32_t idx, const int32_t *c) {
    if (staged) write_node<DIM>(s_nodes, idx - lo, c);
    else write_node<DIM>(octree, idx, c);
  };
  bool active = q <= qlast;
  uint32_t excl = 0, lsq = 0;
  if (active) {
    excl = q ? ps[q - 1] : 0u;
    lsq = ps[q] - excl;
    active = (q == 0) || lsq > 0;
  }
  if (active) {
  const BrtNodeT<K> nq = brt[q];
  const int lvl = nq.lcp_length / DIM;
  uint32_t cmask = 0;
  uint32_t stack[8];
  int sp = 0;
  stack[sp++] = q;
  // A well-formed tree visits at most 2^DIM - 1 nodes here; the cap only
  // guards against a malformed tree the BRT stage did not flag.
  int visits = 0;
  while (sp > 0) {
    if (++visits > 64) {
      atomicOr(&st->status, (uint32_t)ST_STACK);
      break;
    }
    const uint32_t x = stack[--sp];
    const BrtNodeT<K> nx = (x == q) ? nq : brt[x];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const uint32_t c = (uint32_t)nx.left + side;
      const bool leaf = side ? nx.right_leaf : nx.left_leaf;
      if (leaf) {
        if (conflict_bits && test_bit(conflict_bits, c))
          cmask |= 1u << digit_at<K, DIM>(keys[c], lvl, mbits);
        continue;
      }
      const uint32_t psc = ps[c];
      const uint32_t lsc = psc - ps[c - 1];
      if (lsc > 0) {
        ch[quadrant<K, DIM>(brt[c], (int)lsc - 1)] = (int32_t)(psc - 1);
      } else if (sp < 8) {
        stack[sp++] = c;
      } else {
        atomicOr(&st->status, (uint32_t)ST_STACK);
      }
    }
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
  }  // active
  if (staged) {
    __syncthreads();
    const uint32_t words = span * STRIDE;
    int32_t *out = octree + (size_t)lo * STRIDE;
    for (uint32_t w = threadIdx.x; w < words; w += kBlock) out[w] = s_nodes[w];
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
void launch_dim(const BrtNodeT<K> *brt, const int32_t *parent, const uint32_t *ps, const K *keys,
                const uint32_t *conflict_bits, int64_t max_m, int mbits, int32_t *octree,
                uint32_t *node_conflict, int64_t capacity, int force_scatter, RunState *st,
                hipStream_t s) {
  const int64_t nodes = max_m > 1 ? max_m - 1 : 1;
  const unsigned grid = (unsigned)((nodes + kBlock - 1) / kBlock);
  if (!force_scatter)
    gather_kernel<K, DIM><<<grid, kBlock, 0, s>>>(brt, ps, keys, conflict_bits, mbits, octree,
                                                  node_conflict, capacity, st);
  // The scatter path is a no-op unless forced or st->malformed; its kernels
  // loop over their ranges, so a capped grid keeps the no-op cheap.
  const unsigned sgrid = grid < 1024 ? grid : 1024;
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

template <typename K>
void launch_local_splits(const BrtNodeT<K> *brt, int64_t m, int dim, uint32_t *ls, uint64_t *tmp,
                         hipStream_t s) {
  const unsigned grid = (unsigned)((m + kBlock - 1) / kBlock);
  unsigned long long *t = reinterpret_cast<unsigned long long *>(tmp);
  local_splits_claim_kernel<K><<<grid, kBlock, 0, s>>>(brt, m, dim, t);
  local_splits_finish_kernel<K><<<grid, kBlock, 0, s>>>(brt, m, dim, t, ls);
}

template <typename K>
void launch_octree(const BrtNodeT<K> *brt, const int32_t *parent, const uint32_t *ps,
                   const K *keys, const uint32_t *conflict_bits, int64_t max_m, int dim,
                   int mbits, int32_t *octree, uint32_t *node_conflict, int64_t capacity,
                   int force_scatter, RunState *st, hipStream_t s) {
  if (dim == 2)
    launch_dim<K, 2>(brt, parent, ps, keys, conflict_bits, max_m, mbits, octree, node_conflict,
                     capacity, force_scatter, st, s);
  else
    launch_dim<K, 3>(brt, parent, ps, keys, conflict_bits, max_m, mbits, octree, node_conflict,
                     capacity, force_scatter, st, s);
}

#define PGVD_INST(K)                                                                          \
  template void launch_octree<K>(const BrtNodeT<K> *, const int32_t *, const uint32_t *,       \
                                 const K *, const uint32_t *, int64_t, int, int, int32_t *,    \
                                 uint32_t *, int64_t, int, RunState *, hipStream_t);   \
  template void launch_local_splits<K>(const BrtNodeT<K> *, int64_t, int, uint32_t *,          \
                                       uint64_t *, hipStream_t);
PGVD_INST(uint32_t)
PGVD_INST(uint64_t)
PGVD_INST(Key128)
#undef PGVD_INST

}  // namespace pgvd
