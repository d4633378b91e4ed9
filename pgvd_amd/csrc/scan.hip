// scan.hip -- single-pass decoupled look-back scans.  The unique and
// local-split scans give each thread 16 consecutive elements (16-byte vector
// loads and stores, a serial scan in registers, one workgroup scan of the
// per-thread totals); the order-exact splits scan and the stage scan use
// striped tiles (element r*256 + t in thread t).
//
//   unique        UniqueSorted (opencl/Kernels.cpp:198-217): UniquePredicate
//                 (C/ParallelAlgorithms.c:22-31) + scan + BUSingleCompact
//                 (:73-81) fused into one pass.  The reference pads the key
//                 array with zero keys (opencl/Kernels.cpp:62-75); here the
//                 padding is one virtual zero key in front of the sorted keys.
//   splits scan   ComputeLocalSplits (C/BuildOctree.c:17-40) fused into the
//                 inclusive scan of the local splits (opencl/Kernels.cpp:376),
//                 from parent claims -- the order-exact form used when the
//                 BRT is not a well-formed radix tree.
//   scan_u32      StreamScan_p / StreamScan_s (opencl/Kernels.cpp:102-159):
//                 inclusive scan of u32 values (mod 2^32).
#include "tilescan.hpp"

namespace pgvd {

namespace {

constexpr int kRows = kScanTile / kBlock;  // 16

template <typename K> __device__ inline K shfl_up1(K v);
template <> __device__ inline uint32_t shfl_up1<uint32_t>(uint32_t v) { return __shfl_up(v, 1, 64); }
template <> __device__ inline uint64_t shfl_up1<uint64_t>(uint64_t v) { return __shfl_up(v, 1, 64); }
template <> __device__ inline Key128 shfl_up1<Key128>(Key128 v) {
  return {__shfl_up(v.lo, 1, 64), __shfl_up(v.hi, 1, 64)};
}

// One unique tile: elements [base, base+4096) of kin[0..n) (sorted), thread t
// holding base + 16t .. +16.  Writes [zero?] + the distinct keys to
// kout[prepend + rank]; with labels, flags run r in conflict_bits when the
// labels of its keys are not all equal.  *m_out (last tile only) = m.
// kout may be kin - prepend (in place): the r-th distinct key then goes to
// kin[r], never past its own input slot, and a tile writes only after its
// look-back has resolved, i.e. after every earlier tile has loaded its
// input.  A tile whose keys are all distinct and already in place (no
// duplicate before it) writes nothing.
template <typename K, bool LAB>
__device__ inline void unique_tile(const K *__restrict__ kin, const uint32_t *__restrict__ lab,
                                   uint32_t n, uint32_t prepend, K *__restrict__ kout,
                                   uint32_t *__restrict__ conflict_bits, uint64_t *lookback,
                                   int tile, int lo, uint32_t base, uint32_t *m_out) {
  constexpr int R = kRows;
  __shared__ uint32_t s_w[4];
  const int t = threadIdx.x, lane = t & 63;
  const uint32_t e0 = base + (uint32_t)t * R;
  const uint32_t valid = e0 < n ? min((uint32_t)R, n - e0) : 0u;
  K k[R];
  load_run<K, R>(kin + e0, valid, k);
  K prev = shfl_up1(k[R - 1]);
  if (lane == 0 && valid > 0 && e0 > 0) prev = kin[e0 - 1];
  uint32_t heads = 0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const uint32_t e = e0 + j;
    const bool h = (j < (int)valid) &&
                   ((e == 0) ? (!prepend || !kzero(k[0])) : !keq(k[j], j ? k[j - 1] : prev));
    heads |= (uint32_t)h << j;
  }
  uint32_t total;
  const uint32_t texcl = block_excl_sum<uint32_t>((uint32_t)__popc(heads), s_w, &total);
  const uint64_t excl = tile_prefix(lookback, tile, total, lo);
  const uint32_t obase = (uint32_t)(prepend + excl + texcl);
  const uint32_t tile_n = n > base ? min((uint32_t)kScanTile, n - base) : 0u;
  const bool in_place = total == tile_n && kout + prepend + excl == kin + base;   // uniform
  if (in_place) {
    // every key of the tile is already where it goes
  } else if constexpr (sizeof(K) <= 8) {
    // Compacted keys are staged in LDS (index r + r/16 against bank
    // conflicts) so that the tile's output is written in coalesced runs.
    __shared__ K s_out[kScanTile + kScanTile / 16];
    uint32_t o = texcl;
#pragma unroll
    for (int j = 0; j < R; ++j)
      if ((heads >> j) & 1u) {
        s_out[o + (o >> 4)] = k[j];
        ++o;
      }
    __syncthreads();
    K *dst = kout + prepend + excl;
    for (uint32_t r = t; r < total; r += kBlock) dst[r] = s_out[r + (r >> 4)];
  } else {
    uint32_t o = obase;
#pragma unroll
    for (int j = 0; j < R; ++j)
      if ((heads >> j) & 1u) kout[o++] = k[j];
  }
  if (LAB) {
    uint32_t l[R];
    load_run<uint32_t, R>(lab + e0, valid, l);
    uint32_t lprev = __shfl_up(l[R - 1], 1, 64);
    if (lane == 0 && valid > 0 && e0 > 0) lprev = lab[e0 - 1];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const uint32_t e = e0 + j;
      if (j < (int)valid && e > 0 && !((heads >> j) & 1u) && l[j] != (j ? l[j - 1] : lprev)) {
        // not a head: same key as e-1, in run (heads up to e) - 1; the
        // virtual zero key carries no label, so e = 0 never compares.
        const uint32_t run = obase + (uint32_t)__popc(heads & ((1u << j) - 1u)) - 1u;
        atomicOr(&conflict_bits[run >> 5], 1u << (run & 31));
      }
    }
  }
  if (tile == lo && t == 0 && prepend) kout[0] = kzero_value<K>();
  if (m_out && t == 0) *m_out = (uint32_t)(prepend + excl + total);
}

template <typename K, bool LAB>
__global__ __launch_bounds__(kBlock) void unique_kernel(const K *__restrict__ keys,
                                                        const uint32_t *__restrict__ labels,
                                                        uint32_t n, int prepend_zero,
                                                        K *__restrict__ out,
                                                        uint32_t *__restrict__ conflict_bits,
                                                        uint64_t *lookback, RunState *st,
                                                        const RunState *skip) {
  if (skip && skip->msd_ok && !skip->msd_dups) {
    // distinct sorted keys, already where they go (see unique_tile)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      if (prepend_zero) out[0] = kzero_value<K>();
      st->m = (uint32_t)prepend_zero + n;
    }
    return;
  }
  __shared__ uint32_t s_tile;
  if (threadIdx.x == 0) s_tile = atomicAdd(&st->counters[CTR_UNIQUE], 1u);
  __syncthreads();
  const int tile = (int)s_tile;
  const int ntiles = n ? (int)((n + kScanTile - 1) / kScanTile) : 1;
  unique_tile<K, LAB>(keys, labels, n, (uint32_t)prepend_zero, out, conflict_bits, lookback, tile,
                      0, (uint32_t)tile * kScanTile, tile == ntiles - 1 ? &st->m : nullptr);
}

// Local split of BRT node c computed from its own parent claim
// (parent[c] = claimer + 1, 0 = unclaimed): C/BuildOctree.c:17-40 lets the
// last claimer write the child's entry; the claim array holds that claimer.
template <typename K>
__device__ inline uint32_t local_split(const BrtNodeT<K> *brt, const int32_t *parent, uint32_t c,
                                       uint32_t m, int dim) {
  if (c == 0) return (m == 1) ? 1u : 1u + brt[0].lcp_length / dim;
  if (c >= m - 1) return 0u;
  const int32_t pc = parent[c];
  if (pc == 0) return 0u;
  return (uint32_t)(brt[c].lcp_length / dim) - (uint32_t)(brt[pc - 1].lcp_length / dim);
}

template <typename K>
__global__ __launch_bounds__(kBlock) void splits_scan_kernel(const BrtNodeT<K> *__restrict__ brt,
                                                             int32_t *__restrict__ parent,
                                                             uint32_t *__restrict__ ls_out,
                                                             uint32_t *__restrict__ ps, int dim,
                                                             int only_if_malformed,
                                                             uint64_t *lookback, RunState *st) {
  if (only_if_malformed && !st->malformed) return;
  __shared__ uint32_t s_w[kRows * 4];
  __shared__ uint32_t s_tile;
  const int t = threadIdx.x;
  const uint32_t m = st->m;
  const int ntiles = (int)((m + kScanTile - 1) / kScanTile);
  for (;;) {
    __syncthreads();   // s_tile / s_w reuse across loop iterations
    if (t == 0) s_tile = atomicAdd(&st->counters[CTR_SPLITS], 1u);
    __syncthreads();
    const int tile = (int)s_tile;
    if (tile >= ntiles) return;
    const uint32_t base = (uint32_t)tile * kScanTile;
    uint32_t v[kRows], ex[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const uint32_t c = base + r * kBlock + t;
      v[r] = (c < m) ? local_split(brt, parent, c, m, dim) : 0u;
    }
    const uint32_t total = striped_excl_scan<kRows, uint32_t>(v, ex, s_w);
    const uint64_t excl = tile_prefix(lookback, tile, total);
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const uint32_t c = base + r * kBlock + t;
      if (c < m) {
        ps[c] = (uint32_t)(excl + ex[r] + v[r]);
        if (ls_out) ls_out[c] = v[r];
      }
    }
    // Reference form of the parent links: root -1, unclaimed 0
    // (opencl/Kernels.cpp:405 zero-initialises the BRT), else the claimer.
    // local_split above read parent[c] for this thread's own c only.
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const uint32_t c = base + r * kBlock + t;
      if (c + 1 < m) {
        const int32_t pc = parent[c];
        parent[c] = (c == 0) ? -1 : (pc ? pc - 1 : 0);
      }
    }
    if (tile == ntiles - 1 && t == 0) st->octree_size = (uint32_t)(excl + total);
  }
}

__global__ __launch_bounds__(kBlock) void scan_u32_kernel(const uint32_t *__restrict__ in,
                                                          uint32_t *__restrict__ out, uint32_t n,
                                                          uint64_t *lookback, RunState *st) {
  __shared__ uint64_t s_w[kRows * 4];
  __shared__ uint32_t s_tile;
  const int t = threadIdx.x;
  if (t == 0) s_tile = atomicAdd(&st->counters[CTR_SCAN], 1u);
  __syncthreads();
  const int tile = (int)s_tile;
  const uint32_t base = (uint32_t)tile * kScanTile;
  uint64_t v[kRows], ex[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const uint32_t e = base + r * kBlock + t;
    v[r] = (e < n) ? in[e] : 0u;
  }
  const uint64_t total = striped_excl_scan<kRows, uint64_t>(v, ex, s_w);
  const uint64_t excl = tile_prefix(lookback, tile, total);
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const uint32_t e = base + r * kBlock + t;
    if (e < n) out[e] = (uint32_t)(excl + ex[r] + v[r]);
  }
}

// One local-split scan tile: inclusive scan of (lsx & 0xff) over
// [base, base+4096) of [0, m) into ps, thread t holding base + 16t .. +16;
// *o_out (last tile only) = the octree size.
__device__ inline void ls_scan_tile(const uint32_t *__restrict__ lsx, uint32_t *__restrict__ ps,
                                    uint32_t m, uint64_t *lookback, int tile, int lo,
                                    uint32_t base, uint32_t *o_out) {
  constexpr int R = kRows;
  __shared__ uint32_t s_w[4];
  const uint32_t e0 = base + threadIdx.x * R;
  const uint32_t valid = e0 < m ? min((uint32_t)R, m - e0) : 0u;
  uint32_t v[R];
  load_run<uint32_t, R>(lsx + e0, valid, v);
  uint32_t run = 0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    run += (j < (int)valid) ? (v[j] & 0xffu) : 0u;
    v[j] = run;
  }
  uint32_t total;
  const uint32_t texcl = block_excl_sum<uint32_t>(run, s_w, &total);
  const uint64_t excl = tile_prefix(lookback, tile, total, lo);
  const uint32_t add = (uint32_t)excl + texcl;
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] += add;
  store_run<uint32_t, R>(ps + e0, valid, v);
  if (o_out && threadIdx.x == 0) *o_out = (uint32_t)(excl + total);
}

// Exclusive scan, in place, of the per-chunk local-split totals
// chunk[0..ceil(m/256)) written by the BRT pass; the grand total is the
// octree size (opencl/Kernels.cpp:376-378).  Each tile covers 4096 chunks,
// thread t holding 16 consecutive ones.
__global__ __launch_bounds__(kBlock) void chunk_scan_kernel(uint32_t *__restrict__ chunk,
                                                            uint64_t *lookback, RunState *st) {
  constexpr int R = kRows;
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_w[4];
  const uint32_t m = st->m;
  const uint32_t nc = (m + kBlock - 1) / kBlock;
  const int ntiles = (int)((nc + kScanTile - 1) / kScanTile);
  if (threadIdx.x == 0) s_tile = atomicAdd(&st->counters[CTR_LS], 1u);
  __syncthreads();
  const int tile = (int)s_tile;
  if (tile >= ntiles) return;
  const uint32_t e0 = (uint32_t)tile * kScanTile + threadIdx.x * R;
  const uint32_t valid = e0 < nc ? min((uint32_t)R, nc - e0) : 0u;
  uint32_t v[R];
  load_run<uint32_t, R>(chunk + e0, valid, v);
  uint64_t run = 0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const uint32_t x = (j < (int)valid) ? v[j] : 0u;
    v[j] = (uint32_t)run;
    run += x;
  }
  uint64_t total;
  const uint64_t texcl = block_excl_sum<uint64_t>(run, s_w, &total);
  const uint64_t excl = tile_prefix(lookback, tile, total);
  const uint32_t add = (uint32_t)(excl + texcl);
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] += add;
  store_run<uint32_t, R>(chunk + e0, valid, v);
  if (tile == ntiles - 1 && threadIdx.x == 0) {
    const uint64_t all = excl + total;
    if (all > 0xFFFFFFFFull) atomicOr(&st->status, (uint32_t)ST_TOO_LARGE);
    st->octree_size = (uint32_t)all;
  }
}

// ---- batched form ----
// Per-instance unique: instance s reads keys[kofs .. kofs+n) (sorted) and
// writes [zero?] + its distinct keys to out[kofs..); scan tiles cover n+1
// slots per instance, the last one writes ss[s].m.
template <typename K>
__global__ __launch_bounds__(kBlock) void unique_seg_kernel(const K *__restrict__ keys,
                                                            K *__restrict__ out,
                                                            const SegInfo *__restrict__ seg,
                                                            TileTable tt, SegState *ss,
                                                            uint64_t *lookback, RunState *st) {
  __shared__ uint32_t s_tile;
  if (threadIdx.x == 0) s_tile = atomicAdd(&st->counters[CTR_UNIQUE], 1u);
  __syncthreads();
  const int tile = (int)s_tile;
  const uint32_t s = tt.seg[tile];
  const int lo = (int)tt.first[s];
  const SegInfo si = seg[s];
  unique_tile<K, false>(keys + si.kofs, nullptr, si.n, si.prepend, out + si.kofs, nullptr,
                        lookback, tile, lo, (uint32_t)(tile - lo) * kScanTile,
                        tile == (int)tt.first[s + 1] - 1 ? &ss[s].m : nullptr);
}

// Per-instance inclusive scan of the local splits (low byte of lsx) over
// [0, m_s); the last tile in use writes ss[s].octree_size.
__global__ __launch_bounds__(kBlock) void ls_scan_seg_kernel(const uint32_t *__restrict__ lsx,
                                                             uint32_t *__restrict__ ps,
                                                             const SegInfo *__restrict__ seg,
                                                             TileTable tt, SegState *ss,
                                                             uint64_t *lookback, RunState *st) {
  __shared__ uint32_t s_tile;
  if (threadIdx.x == 0) s_tile = atomicAdd(&st->counters[CTR_LS], 1u);
  __syncthreads();
  const int tile = (int)s_tile;
  const uint32_t s = tt.seg[tile];
  const int lo = (int)tt.first[s];
  const uint32_t m = ss[s].m;
  const int used = (int)((m + kScanTile - 1) / kScanTile);
  if (tile - lo >= used) return;    // no later tile of this instance depends on it
  const uint32_t kofs = seg[s].kofs;
  ls_scan_tile(lsx + kofs, ps + kofs, m, lookback, tile, lo, (uint32_t)(tile - lo) * kScanTile,
               tile - lo == used - 1 ? &ss[s].octree_size : nullptr);
}

// One workgroup: oct_off = exclusive scan of the instances' octree sizes.
__global__ __launch_bounds__(kBlock) void oct_offsets_kernel(SegState *ss, int64_t B,
                                                             RunState *st) {
  __shared__ uint32_t s_w[kRows * 4];
  uint64_t carry = 0;
  for (int64_t b0 = 0; b0 < B; b0 += kScanTile) {
    uint32_t v[kRows], ex[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int64_t i = b0 + r * kBlock + threadIdx.x;
      v[r] = (i < B) ? ss[i].octree_size : 0u;
    }
    const uint32_t total = striped_excl_scan<kRows, uint32_t>(v, ex, s_w);
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int64_t i = b0 + r * kBlock + threadIdx.x;
      if (i < B) ss[i].oct_off = (uint32_t)(carry + ex[r]);
    }
    carry += total;
    __syncthreads();    // s_w is reused by the next chunk
  }
  if (threadIdx.x == 0) {
    if (carry > 0xFFFFFFFFull) atomicOr(&st->status, (uint32_t)ST_TOO_LARGE);
    st->octree_size = (uint32_t)carry;
  }
}

}  // namespace

void launch_chunk_scan(uint32_t *chunk_tot, int64_t max_m, uint64_t *lookback, RunState *st,
                       hipStream_t s) {
  const int64_t chunks = ((max_m > 0 ? max_m : 1) + kBlock - 1) / kBlock;
  chunk_scan_kernel<<<(unsigned)scan_tiles(chunks), kBlock, 0, s>>>(chunk_tot, lookback, st);
}

template <typename K>
void launch_unique(const K *keys, const uint32_t *labels, int64_t n, int prepend_zero, K *out,
                   uint32_t *conflict_bits, uint64_t *lookback, RunState *st, hipStream_t s,
                   const RunState *skip_if_msd) {
  const int64_t tiles = scan_tiles(n > 0 ? n : 1);
  if (labels)
    unique_kernel<K, true><<<(unsigned)tiles, kBlock, 0, s>>>(keys, labels, (uint32_t)n,
                                                              prepend_zero, out, conflict_bits,
                                                              lookback, st, skip_if_msd);
  else
    unique_kernel<K, false><<<(unsigned)tiles, kBlock, 0, s>>>(keys, nullptr, (uint32_t)n,
                                                               prepend_zero, out, nullptr,
                                                               lookback, st, skip_if_msd);
}

template <typename K>
void launch_splits_scan(const BrtNodeT<K> *brt, int32_t *parent, uint32_t *ls, uint32_t *ps,
                        int64_t max_m, int dim, int only_if_malformed, uint64_t *lookback,
                        RunState *st, hipStream_t s) {
  int64_t grid = scan_tiles(max_m > 0 ? max_m : 1);
  if (grid > 2048) grid = 2048;      // persistent: each workgroup loops over tiles
  splits_scan_kernel<K><<<(unsigned)grid, kBlock, 0, s>>>(brt, parent, ls, ps, dim,
                                                          only_if_malformed, lookback, st);
}

void launch_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, uint64_t *lookback,
                     RunState *st, hipStream_t s) {
  const int64_t tiles = scan_tiles(n > 0 ? n : 1);
  scan_u32_kernel<<<(unsigned)tiles, kBlock, 0, s>>>(in, out, (uint32_t)n, lookback, st);
}

template <typename K>
void launch_unique_seg(const K *keys, K *out, const BatchDev &b, uint64_t *lookback,
                       RunState *st, hipStream_t s) {
  if (b.n_scan <= 0) return;
  unique_seg_kernel<K><<<(unsigned)b.n_scan, kBlock, 0, s>>>(keys, out, b.seg, b.scan, b.ss,
                                                            lookback, st);
}

void launch_ls_scan_seg(const uint32_t *lsx, uint32_t *ps, const BatchDev &b,
                        uint64_t *lookback, RunState *st, hipStream_t s) {
  if (b.n_scan <= 0) return;
  ls_scan_seg_kernel<<<(unsigned)b.n_scan, kBlock, 0, s>>>(lsx, ps, b.seg, b.scan, b.ss,
                                                          lookback, st);
}

void launch_oct_offsets(const BatchDev &b, RunState *st, hipStream_t s) {
  oct_offsets_kernel<<<1, kBlock, 0, s>>>(b.ss, b.B, st);
}

#define PGVD_INST(K)                                                                          \
  template void launch_unique<K>(const K *, const uint32_t *, int64_t, int, K *, uint32_t *,   \
                                 uint64_t *, RunState *, hipStream_t, const RunState *);       \
  template void launch_splits_scan<K>(const BrtNodeT<K> *, int32_t *, uint32_t *, uint32_t *,  \
                                      int64_t, int, int, uint64_t *, RunState *, hipStream_t);  \
  template void launch_unique_seg<K>(const K *, K *, const BatchDev &, uint64_t *, RunState *, \
                                     hipStream_t);
PGVD_INST(uint32_t)
PGVD_INST(uint64_t)
PGVD_INST(Key128)
#undef PGVD_INST

}  // namespace pgvd
