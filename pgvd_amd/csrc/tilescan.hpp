// tilescan.hpp -- single-pass decoupled look-back building blocks shared by
// the scan and unique kernels (wave64; 256-thread workgroups unless a
// template parameter says otherwise).
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

// Publishes the tile aggregate, resolves the tile's exclusive prefix within
// the run of tiles starting at `lo` and publishes its inclusive prefix;
// returns the exclusive prefix to every thread.  All BS threads of the
// workgroup call it.  Each round reads the status of the BS nearest
// unresolved predecessors at once (one per thread) and stops at the nearest
// inclusive one, so a tile k tiles past the last resolved one waits about
// k/BS round trips.
template <int BS = kBlock>
__device__ inline uint64_t tile_prefix(uint64_t *status, int tile, uint64_t total, int lo = 0) {
  constexpr int NW = BS / 64;
  __shared__ int s_first[NW];
  __shared__ uint64_t s_part[NW];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) st_rel_u64(status + tile, (tile == lo ? kSIncl : kSAgg) | total);
  uint64_t excl = 0;
  if (tile > lo) {
    int base = tile - 1;
    for (;;) {
      const int k = base - t;
      uint64_t wv = kSIncl;  // before the run's first tile: an inclusive zero
      if (k >= lo) {
        for (;;) {
          wv = ld_acq_u64(status + k);
          if (wv & ~kSVal) break;
          __builtin_amdgcn_s_sleep(1);
        }
      }
      const uint64_t b = __ballot((wv & ~kSVal) == kSIncl);
      if (lane == 0) s_first[w] = b ? __ffsll((long long)b) - 1 : 64;
      __syncthreads();
      int first = BS;
      if constexpr (NW <= 4) {
        for (int j = NW - 1; j >= 0; --j)
          if (s_first[j] < 64) first = j * 64 + s_first[j];
      } else {
        const int f = lane < NW ? s_first[lane] : 64;
        first = f < 64 ? lane * 64 + f : BS;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) first = min(first, __shfl_xor(first, d, 64));
      }
      uint64_t val = (t <= first) ? (wv & kSVal) : 0;
      val = wave_reduce_add(val);
      if (lane == 0) s_part[w] = val;
      __syncthreads();
      if constexpr (NW <= 4) {
        for (int j = 0; j < NW; ++j) excl += s_part[j];
      } else {
        excl += wave_reduce_add<uint64_t>(lane < NW ? s_part[lane] : 0ull);
      }
      __syncthreads();  // s_first / s_part are reused by the next round
      if (first < BS) break;
      base -= BS;
    }
    if (t == 0) st_rel_u64(status + tile, kSIncl | (excl + total));
  }
  return excl;
}

// Exclusive scan of one value per thread over the 256-thread workgroup;
// *total receives the workgroup sum.  s_w: NW words of LDS, NW = waves per
// workgroup.
template <typename T, int NW = kBlock / 64>
__device__ inline T block_excl_sum(T x, T *s_w, T *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const T incl = wave_incl_scan_add(x);
  if (lane == 63) s_w[w] = incl;
  __syncthreads();
  T before = 0, all = 0;
  if constexpr (NW <= 4) {
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      const T y = s_w[j];
      if (j < w) before += y;
      all += y;
    }
  } else {
    // one wave total per lane, scanned across the wave (few registers)
    const T y = wave_incl_scan_add(lane < NW ? s_w[lane] : T(0));
    all = __shfl(y, NW - 1, 64);
    const T b = __shfl(y, w > 0 ? w - 1 : 0, 64);
    before = w > 0 ? b : T(0);
  }
  *total = all;
  return before + incl - x;
}

// Calls f(index) for every bit set in words[0..nwords), index = word * 64 +
// bit.  A workgroup takes 256 words at a time and spreads their set bits
// over its threads (entry e goes to thread e % 256), so a word with many
// bits does not hold one lane for all of them.  Uniform over the workgroup;
// all 256 threads call it.
template <typename F>
__device__ inline void for_each_masked(const uint64_t *__restrict__ words, uint32_t nwords, F f) {
  __shared__ uint64_t s_word[kBlock];
  __shared__ uint32_t s_ofs[kBlock + 1];
  __shared__ uint32_t s_w4[4];
  const uint32_t t = threadIdx.x;
  for (uint32_t base = blockIdx.x * kBlock; base < nwords; base += gridDim.x * kBlock) {
    const uint32_t wd = base + t;
    const uint64_t b = wd < nwords ? words[wd] : 0ull;
    uint32_t total;
    const uint32_t ex = block_excl_sum<uint32_t>((uint32_t)__popcll(b), s_w4, &total);
    s_word[t] = b;
    s_ofs[t] = ex;
    if (t == 0) s_ofs[kBlock] = total;
    __syncthreads();
    for (uint32_t e = t; e < total; e += kBlock) {
      // the word holding entry e: s_ofs[lo] <= e < s_ofs[lo + 1]
      uint32_t lo = 0, hi = kBlock;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_ofs[mid] <= e) lo = mid; else hi = mid;
      }
      uint64_t x = s_word[lo];
      for (uint32_t k = e - s_ofs[lo]; k; --k) x &= x - 1;   // drop the lower set bits
      f((base + lo) * 64u + (uint32_t)(__ffsll((unsigned long long)x) - 1));
    }
    __syncthreads();   // s_word, s_ofs and s_w4 are refilled next round
  }
}

// R consecutive elements starting at p (valid of them in range): 16-byte
// vector loads when the run is whole and aligned, else element loads.
template <typename T, int R>
__device__ inline void load_run(const T *__restrict__ p, uint32_t valid, T (&v)[R]) {
  constexpr int NV = (int)(sizeof(T) * R / 16);
  if ((sizeof(T) * R) % 16 == 0 && valid >= (uint32_t)R && ((uintptr_t)p & 15) == 0) {
    uint4 q[NV];
    const uint4 *src = reinterpret_cast<const uint4 *>(p);
#pragma unroll
    for (int i = 0; i < NV; ++i) q[i] = src[i];
    __builtin_memcpy(&v[0], &q[0], sizeof(v));
  } else {
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = (j < (int)valid) ? p[j] : T{};
  }
}

template <typename T, int R>
__device__ inline void store_run(T *__restrict__ p, uint32_t valid, const T (&v)[R]) {
  constexpr int NV = (int)(sizeof(T) * R / 16);
  if ((sizeof(T) * R) % 16 == 0 && valid >= (uint32_t)R && ((uintptr_t)p & 15) == 0) {
    uint4 q[NV];
    __builtin_memcpy(&q[0], &v[0], sizeof(v));
    uint4 *dst = reinterpret_cast<uint4 *>(p);
#pragma unroll
    for (int i = 0; i < NV; ++i) dst[i] = q[i];
  } else {
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (j < (int)valid) p[j] = v[j];
  }
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
