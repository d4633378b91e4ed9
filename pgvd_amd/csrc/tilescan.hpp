// tilescan.hpp -- single-pass decoupled look-back building blocks shared by
// the scan, unique and fused BRT kernels (256-thread workgroups, wave64).
//
// A tile's status word is 64 bits: flag in bits 62-63 (1 = aggregate only,
// 2 = inclusive prefix), value in bits 0-61.  Status arrays must be zero
// before the kernel runs.  Tile ids come from an atomic counter, so every
// tile a workgroup looks back at was claimed by a workgroup already running.
#pragma once

#include "kernels.hpp"

namespace pgvd {

constexpr uint64_t kSAgg = 1ull << 62;
constexpr uint64_t kSIncl = 2ull << 62;
constexpr uint64_t kSVal = (1ull << 62) - 1;

// Called by every lane of one wave after the tile's aggregate is published.
// Returns the exclusive prefix of `tile` within the run of tiles starting at
// `lo`, resolving up to 64 predecessors per step.
__device__ inline uint64_t lookback_excl(uint64_t *status, int tile, int lo = 0) {
  const int lane = threadIdx.x & 63;
  uint64_t excl = 0;
  int base = tile - 1;
  while (base >= lo) {
    const int k = base - lane;
    uint64_t wv = kSIncl;  // before the run's first tile: an inclusive zero
    if (k >= lo) {
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

// Publishes the tile aggregate, resolves and publishes the inclusive
// prefix; returns the tile's exclusive prefix (within the run of tiles
// starting at `lo`) to every thread.  All threads of the block must call it.
__device__ inline uint64_t tile_prefix(uint64_t *status, int tile, uint64_t total,
                                       uint64_t *s_excl, int lo = 0) {
  if (threadIdx.x == 0) st_rel_u64(status + tile, (tile == lo ? kSIncl : kSAgg) | total);
  if (tile > lo && threadIdx.x < 64) {
    const uint64_t e = lookback_excl(status, tile, lo);
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

// Exclusive scan of a striped tile: thread t holds v[r] for element
// r*256 + t.  Writes each element's exclusive prefix within the tile into
// ex[r] and returns the tile total.  s_w: RPT*4 words of LDS.
template <int RPT, typename T>
__device__ inline T striped_excl_scan(const T (&v)[RPT], T (&ex)[RPT], T *s_w) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T incl[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    incl[r] = wave_incl_scan_add(v[r]);
    if (lane == 63) s_w[r * 4 + w] = incl[r];
  }
  __syncthreads();
  T running = 0;
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    T before = 0, row = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const T x = s_w[r * 4 + j];
      if (j < w) before += x;
      row += x;
    }
    ex[r] = running + before + incl[r] - v[r];
    running += row;
  }
  return running;
}

// Same for 0/1 flags via ballots: returns the tile total; ex[r] is the
// number of set flags before element r*256 + t in the tile.
template <int RPT>
__device__ inline uint32_t striped_flag_scan(const bool (&f)[RPT], uint32_t (&ex)[RPT],
                                             uint32_t *s_w) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t ltm = lanemask_lt();
  uint32_t inwave[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const uint64_t b = __ballot(f[r]);
    inwave[r] = (uint32_t)__popcll(b & ltm);
    if (lane == 0) s_w[r * 4 + w] = (uint32_t)__popcll(b);
  }
  __syncthreads();
  uint32_t running = 0;
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    uint32_t before = 0, row = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t x = s_w[r * 4 + j];
      if (j < w) before += x;
      row += x;
    }
    ex[r] = running + before + inwave[r];
    running += row;
  }
  return running;
}

}  // namespace pgvd
