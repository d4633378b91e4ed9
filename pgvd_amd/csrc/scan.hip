// scan.hip -- single-pass decoupled look-back scans.
//
//   unique        UniqueSorted (opencl/Kernels.cpp:198-217): UniquePredicate
//                 (C/ParallelAlgorithms.c:22-31) + scan + BUSingleCompact
//                 (:73-81) fused into one pass.  The reference pads the key
//                 array with zero keys (opencl/Kernels.cpp:62-75); here the
//                 padding is one virtual zero key in front of the sorted keys.
//   splits scan   ComputeLocalSplits (C/BuildOctree.c:17-40) fused into the
//                 inclusive scan of the local splits (opencl/Kernels.cpp:376).
//   scan_u32      StreamScan_p / StreamScan_s (opencl/Kernels.cpp:102-159):
//                 inclusive scan of u32 values (mod 2^32).
//
// Each workgroup claims a tile id from an atomic counter, reduces its tile,
// publishes the aggregate, and one wave resolves the exclusive prefix by
// looking back over up to 64 predecessor tiles at a time.
#include "kernels.hpp"

namespace pgvd {

namespace {

constexpr uint64_t kSAgg = 1ull << 62;
constexpr uint64_t kSIncl = 2ull << 62;
constexpr uint64_t kSVal = (1ull << 62) - 1;
constexpr int kIPT = kScanTile / kBlock;  // 16

enum { CTR_UNIQUE = kMaxPasses, CTR_SPLITS = kMaxPasses + 1, CTR_SCAN = kMaxPasses + 2 };

// Inclusive scan of one value per thread across the block; returns the
// thread's inclusive prefix and (via *total) the block total.
__device__ inline uint64_t block_incl_scan64(uint64_t x, uint64_t *s_wsum, uint64_t *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t incl = wave_incl_scan_add(x);
  if (lane == 63) s_wsum[w] = incl;
  __syncthreads();
  uint64_t off = 0, tot = 0;
  for (int j = 0; j < kBlock / 64; ++j) {
    if (j < w) off += s_wsum[j];
    tot += s_wsum[j];
  }
  *total = tot;
  return off + incl;
}

// Called by every lane of wave 0 of the tile's workgroup after the tile's
// aggregate has been published.  Returns the exclusive prefix of the tile.
__device__ inline uint64_t lookback_excl(uint64_t *status, int tile) {
  const int lane = threadIdx.x & 63;
  uint64_t excl = 0;
  int base = tile - 1;
  while (base >= 0) {
    const int k = base - lane;
    uint64_t wv = kSIncl;  // before tile 0: an inclusive zero
    if (k >= 0) {
      for (;;) {
        wv = ld_acq_u64(status + k);
        if (wv & ~kSVal) break;
        __builtin_amdgcn_s_sleep(1);
      }
    }
    const uint64_t incl = __ballot((wv & ~kSVal) == kSIncl);
    uint64_t val = wv & kSVal;
    if (incl) {
      const int first = __ffsll((long long)incl) - 1;
      if (lane > first) val = 0;
      excl += wave_reduce_add(val);
      break;
    }
    excl += wave_reduce_add(val);
    base -= 64;
  }
  return excl;
}

// Publishes the tile aggregate, resolves the exclusive prefix and publishes
// the inclusive value.  Must be called by all threads of the block.
__device__ inline uint64_t tile_prefix(uint64_t *status, int tile, uint64_t total,
                                       uint64_t *s_excl) {
  if (threadIdx.x == 0) st_rel_u64(status + tile, (tile == 0 ? kSIncl : kSAgg) | total);
  if (tile > 0 && threadIdx.x < 64) {
    const uint64_t e = lookback_excl(status, tile);
    if (threadIdx.x == 0) {
      st_rel_u64(status + tile, kSIncl | (e + total));
      *s_excl = e;
    }
  } else if (threadIdx.x == 0) {
    *s_excl = 0;
  }
  __syncthreads();
  return *s_excl;
}

template <typename K, bool LAB>
__global__ __launch_bounds__(kBlock) void unique_kernel(const K *__restrict__ keys,
                                                        const uint32_t *__restrict__ labels,
                                                        uint32_t n, int prepend_zero,
                                                        K *__restrict__ out,
                                                        uint32_t *__restrict__ conflict_bits,
                                                        uint64_t *lookback, RunState *st) {
  __shared__ K s_k[kScanTile];
  __shared__ uint64_t s_wsum[kBlock / 64];
  __shared__ uint64_t s_excl;
  __shared__ uint32_t s_tile;
  const int t = threadIdx.x;
  if (t == 0) s_tile = atomicAdd(&st->counters[CTR_UNIQUE], 1u);
  __syncthreads();
  const int tile = (int)s_tile;
  const int ntiles = n ? (int)((n + kScanTile - 1) / kScanTile) : 1;
  const uint32_t base = (uint32_t)tile * kScanTile;
#pragma unroll
  for (int j = 0; j < kIPT; ++j) {
    const uint32_t e = base + j * kBlock + t;
    if (e < n) s_k[j * kBlock + t] = keys[e];
  }
  __syncthreads();

  // Thread t owns elements [t*16, t*16+16) of the tile.
  const uint32_t e0 = base + t * kIPT;
  K prev = kzero_value<K>();
  if (e0 > 0 && e0 < n) prev = (t > 0) ? s_k[t * kIPT - 1] : keys[e0 - 1];
  uint32_t heads = 0, cnt = 0;
#pragma unroll
  for (int j = 0; j < kIPT; ++j) {
    const uint32_t e = e0 + j;
    if (e < n) {
      const K cur = s_k[t * kIPT + j];
      const bool h = (e == 0) ? (!prepend_zero || !kzero(cur)) : !keq(cur, prev);
      heads |= (uint32_t)h << j;
      cnt += h;
      prev = cur;
    }
  }
  uint64_t total;
  const uint64_t incl = block_incl_scan64(cnt, s_wsum, &total);
  const uint64_t excl = tile_prefix(lookback, tile, total, &s_excl);

  uint32_t pos = (uint32_t)(prepend_zero + excl + incl - cnt);
#pragma unroll
  for (int j = 0; j < kIPT; ++j) {
    const uint32_t e = e0 + j;
    if (e >= n) continue;
    if (heads & (1u << j)) {
      out[pos] = s_k[t * kIPT + j];
      ++pos;
    } else if (LAB && e > 0) {
      // Not a head: same key as element e-1 (the virtual zero has no label).
      if (labels[e] != labels[e - 1]) {
        const uint32_t r = pos - 1;
        atomicOr(&conflict_bits[r >> 5], 1u << (r & 31));
      }
    }
  }
  if (tile == 0 && t == 0 && prepend_zero) out[0] = kzero_value<K>();
  if (tile == ntiles - 1 && t == 0) st->m = (uint32_t)(prepend_zero + excl + total);
}

// Local split of BRT node c (C/BuildOctree.c:17-40 / opencl/Kernels.cpp:303-313),
// computed from the node's own parent claim instead of being written by the
// parent.  parent[c] holds claimer+1 (0 = unclaimed).
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
                                                             uint64_t *lookback, RunState *st) {
  __shared__ uint64_t s_wsum[kBlock / 64];
  __shared__ uint64_t s_excl;
  __shared__ uint32_t s_tile;
  const int t = threadIdx.x;
  const uint32_t m = st->m;
  const int ntiles = (int)((m + kScanTile - 1) / kScanTile);
  if (t == 0) s_tile = atomicAdd(&st->counters[CTR_SPLITS], 1u);
  __syncthreads();
  const int tile = (int)s_tile;
  if (tile >= ntiles) return;
  const uint32_t e0 = (uint32_t)tile * kScanTile + t * kIPT;
  uint32_t v[kIPT];
  uint32_t sum = 0;
#pragma unroll
  for (int j = 0; j < kIPT; ++j) {
    const uint32_t c = e0 + j;
    v[j] = (c < m) ? local_split(brt, parent, c, m, dim) : 0u;
    sum += v[j];
  }
  uint64_t total;
  const uint64_t incl = block_incl_scan64(sum, s_wsum, &total);
  const uint64_t excl = tile_prefix(lookback, tile, total, &s_excl);
  uint32_t run = (uint32_t)(excl + incl - sum);
#pragma unroll
  for (int j = 0; j < kIPT; ++j) {
    const uint32_t c = e0 + j;
    if (c < m) {
      run += v[j];
      ps[c] = run;
      if (ls_out) ls_out[c] = v[j];
    }
  }
  // Reference form of the parent links: root -1, unclaimed 0
  // (opencl/Kernels.cpp:405 zero-initialises the BRT), else the claimer.
  // Every local_split read of parent[] in this tile happened above the
  // block barrier inside tile_prefix, and other tiles read only their own
  // elements.
#pragma unroll
  for (int j = 0; j < kIPT; ++j) {
    const uint32_t c = e0 + j;
    if (c + 1 < m) {
      const int32_t pc = parent[c];
      parent[c] = (c == 0) ? -1 : (pc ? pc - 1 : 0);
    }
  }
  if (tile == ntiles - 1 && t == 0) st->octree_size = (uint32_t)(excl + total);
}

__global__ __launch_bounds__(kBlock) void scan_u32_kernel(const uint32_t *__restrict__ in,
                                                          uint32_t *__restrict__ out, uint32_t n,
                                                          uint64_t *lookback, RunState *st) {
  __shared__ uint32_t s_v[kScanTile];
  __shared__ uint64_t s_wsum[kBlock / 64];
  __shared__ uint64_t s_excl;
  __shared__ uint32_t s_tile;
  const int t = threadIdx.x;
  if (t == 0) s_tile = atomicAdd(&st->counters[CTR_SCAN], 1u);
  __syncthreads();
  const int tile = (int)s_tile;
  const uint32_t base = (uint32_t)tile * kScanTile;
#pragma unroll
  for (int j = 0; j < kIPT; ++j) {
    const uint32_t e = base + j * kBlock + t;
    s_v[j * kBlock + t] = (e < n) ? in[e] : 0u;
  }
  __syncthreads();
  uint32_t v[kIPT];
  uint64_t sum = 0;
#pragma unroll
  for (int j = 0; j < kIPT; ++j) { v[j] = s_v[t * kIPT + j]; sum += v[j]; }
  uint64_t total;
  const uint64_t incl = block_incl_scan64(sum, s_wsum, &total);
  const uint64_t excl = tile_prefix(lookback, tile, total, &s_excl);
  uint32_t run = (uint32_t)(excl + incl - sum);
#pragma unroll
  for (int j = 0; j < kIPT; ++j) { run += v[j]; s_v[t * kIPT + j] = run; }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kIPT; ++j) {
    const uint32_t e = base + j * kBlock + t;
    if (e < n) out[e] = s_v[j * kBlock + t];
  }
}

}  // namespace

template <typename K>
void launch_unique(const K *keys, const uint32_t *labels, int64_t n, int prepend_zero, K *out,
                   uint32_t *conflict_bits, uint64_t *lookback, RunState *st, hipStream_t s) {
  const int64_t tiles = scan_tiles(n > 0 ? n : 1);
  if (labels)
    unique_kernel<K, true><<<(unsigned)tiles, kBlock, 0, s>>>(keys, labels, (uint32_t)n,
                                                              prepend_zero, out, conflict_bits,
                                                              lookback, st);
  else
    unique_kernel<K, false><<<(unsigned)tiles, kBlock, 0, s>>>(keys, nullptr, (uint32_t)n,
                                                               prepend_zero, out, nullptr,
                                                               lookback, st);
}

template <typename K>
void launch_splits_scan(const BrtNodeT<K> *brt, int32_t *parent, uint32_t *ls, uint32_t *ps,
                        int64_t max_m, int dim, uint64_t *lookback, RunState *st,
                        hipStream_t s) {
  const int64_t tiles = scan_tiles(max_m > 0 ? max_m : 1);
  splits_scan_kernel<K><<<(unsigned)tiles, kBlock, 0, s>>>(brt, parent, ls, ps, dim, lookback, st);
}

void launch_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, uint64_t *lookback,
                     RunState *st, hipStream_t s) {
  const int64_t tiles = scan_tiles(n > 0 ? n : 1);
  scan_u32_kernel<<<(unsigned)tiles, kBlock, 0, s>>>(in, out, (uint32_t)n, lookback, st);
}

#define PGVD_INST(K)                                                                          \
  template void launch_unique<K>(const K *, const uint32_t *, int64_t, int, K *, uint32_t *,   \
                                 uint64_t *, RunState *, hipStream_t);                         \
  template void launch_splits_scan<K>(const BrtNodeT<K> *, int32_t *, uint32_t *, uint32_t *,  \
                                      int64_t, int, uint64_t *, RunState *, hipStream_t);
PGVD_INST(uint32_t)
PGVD_INST(uint64_t)
PGVD_INST(Key128)
#undef PGVD_INST

}  // namespace pgvd
