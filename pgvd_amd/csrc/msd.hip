// msd.hip -- bucket sort for 64-bit keys of large single builds: the keys
// are split by their top b bits (b = 10..12) into 2^b buckets in two
// scatter passes of b/2 and b - b/2 bits, then every bucket is sorted on its
// remaining low bits inside one workgroup's LDS.
//
// The output is the sorted key sequence, the same one the reference's LSD
// split sort (RadixSortBigUnsigned, opencl/Kernels.cpp:219-247) produces:
// keys without payload, so equal keys are indistinguishable and the order of
// the scatters within a bucket does not matter.  Every kernel here does
// nothing unless st->msd_ok, which the plan kernel sets when every bucket
// holds at most kMsdCap keys; otherwise the pipeline's LSD passes, launched
// beside these and doing nothing when st->msd_ok, sort instead.  For
// uniform-ish keys a bucket is ~4K keys: the LSD passes over HBM become
// passes inside LDS.
#include "kernels.hpp"
#include "tilescan.hpp"

namespace pgvd {

namespace {

constexpr int kSplitIpt = 16;                  // keys per thread of a scatter tile
constexpr int kSplitTile = kBlock * kSplitIpt; // 4096
constexpr int kSplitMaxDigits = 64;            // 2^(b/2) and 2^(b - b/2) for b <= 12
static_assert(kMsdCap == kBlock * 20, "a bucket sort holds 20 keys per thread at most");

// Words of msd_cursor: cursor[0..2^b) for the second scatter (then the
// list of buckets left to the fallback kernel), then the first scatter's
// 2^(b/2) cursors kMsdCStride words apart, then the second scatter's
// 2^(b/2) + 1 tile starts.
constexpr int kCursor1 = kMsdMaxBuckets;
constexpr int kTileStart = kMsdMaxBuckets + kSplitMaxDigits * kMsdCStride;
static_assert(kTileStart + kSplitMaxDigits + 1 <= kMsdCursorWords, "msd_cursor holds them");

// Plan, one workgroup: exclusive scan of the bucket histogram into
// start[0..nb] and the second scatter's cursors; the first scatter's
// cursors (the start of every group of nb2 buckets) and, per group, the
// start of its tiles in the second scatter; msd_ok = 1 when every bucket
// fits one workgroup.  nb is a power of two in [1024, 4096].
__global__ __launch_bounds__(kBlock) void msd_plan_kernel(const uint32_t *__restrict__ hist,
                                                          uint32_t n, int nb, int nb2,
                                                          uint32_t *__restrict__ start,
                                                          uint32_t *__restrict__ cursor,
                                                          RunState *st) {
  __shared__ uint32_t s_w[4], s_mx[4];
  const int t = threadIdx.x;
  const int per = nb / kBlock;                  // 4..16 consecutive buckets per thread
  uint32_t v[16];
  uint32_t run = 0, mx = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    v[i] = i < per ? hist[t * per + i] : 0u;
    run += v[i];
    mx = max(mx, v[i]);
  }
  uint32_t total;
  uint32_t ex = block_excl_sum<uint32_t>(run, s_w, &total);
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (i < per) {
      start[t * per + i] = ex;
      cursor[t * per + i] = ex;
      ex += v[i];
    }
  if (t == 0) start[nb] = total;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, d, 64));
  if ((t & 63) == 0) s_mx[t >> 6] = mx;
  __syncthreads();   // start[] is complete
  if (t == 0) {
    const uint32_t m4 = max(max(s_mx[0], s_mx[1]), max(s_mx[2], s_mx[3]));
    // (a histogram that does not count every key leaves the sort to the LSD
    // passes)
    st->msd_ok = m4 <= (uint32_t)kMsdCap && total == n ? 1u : 0u;
    st->msd_dups = 0;
  }
  const int nb1 = nb / nb2;   // <= 64: wave 0, lane g = group g
  if (t < 64) {
    uint32_t lo = 0, tiles = 0;
    if (t < nb1) {
      lo = start[t * nb2];
      tiles = (start[(t + 1) * nb2] - lo + kSplitTile - 1) / kSplitTile;
    }
    const uint32_t incl = wave_incl_scan_add(tiles);
    if (t < nb1) {
      cursor[kCursor1 + t * kMsdCStride] = lo;
      cursor[kTileStart + t] = incl - tiles;
    }
    if (t == nb1 - 1) cursor[kTileStart + nb1] = incl;
  }
}

// One scatter tile of a BS-thread workgroup: keys in[lo .. hi) (0 < hi - lo
// <= 16 BS) go to out in runs by digit ld = (key >> dshift) & (nd - 1), the
// run of digit ld starting at an atomic claim on gcursor[gbase + ld *
// cstride].  Keys are staged in LDS (s_keys, 16 BS words) in digit order so
// that each run is written by consecutive lanes.
template <int BS>
__device__ inline void split_tile(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                  uint32_t n, uint32_t lo, uint32_t hi, int dshift, int nd,
                                  uint32_t *__restrict__ gcursor, uint32_t gbase,
                                  int cstride, uint64_t *s_keys) {
  __shared__ uint32_t s_cnt[kSplitMaxDigits], s_lstart[kSplitMaxDigits],
      s_gstart[kSplitMaxDigits];
  const int t = threadIdx.x;
  const uint32_t dmask = (uint32_t)nd - 1;
  if (t < nd) s_cnt[t] = 0;
  uint64_t k[kSplitIpt];
  uint32_t dr[kSplitIpt];
#pragma unroll
  for (int j = 0; j < kSplitIpt; ++j) {
    const uint32_t i = lo + j * BS + t;
    k[j] = in[i < hi ? i : lo];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSplitIpt; ++j) {
    const uint32_t i = lo + j * BS + t;
    dr[j] = 0;
    if (i < hi) {
      const uint32_t d = (uint32_t)(k[j] >> dshift) & dmask;
      dr[j] = atomicAdd(&s_cnt[d], 1u) | (d << 16);
    }
  }
  __syncthreads();
  if (t < 64) {   // wave 0: digit starts in the tile and the runs' claims
    const uint32_t c = t < nd ? s_cnt[t] : 0u;
    const uint32_t incl = wave_incl_scan_add(c);
    if (t < nd) {
      s_lstart[t] = incl - c;
      s_gstart[t] = c ? atomicAdd(&gcursor[gbase + t * cstride], c) : 0u;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSplitIpt; ++j) {
    const uint32_t i = lo + j * BS + t;
    if (i < hi) s_keys[s_lstart[dr[j] >> 16] + (dr[j] & 0xffffu)] = k[j];
  }
  __syncthreads();
  const uint32_t cnt = hi - lo;
  for (uint32_t p = t; p < cnt; p += BS) {
    const uint64_t key = s_keys[p];
    const uint32_t d = (uint32_t)(key >> dshift) & dmask;
    const uint32_t pos = s_gstart[d] + (p - s_lstart[d]);
    if (pos < n) out[pos] = key;   // always true when the histogram matches the keys
  }
}

// First scatter: contiguous tiles of the whole array, by the top b1 bits,
// in 512-thread workgroups of 8192 keys: every tile claims on the same 2^b1
// cursors, so fewer, larger tiles wait less on them (two such workgroups
// per CU hold as many waves as four of 256 threads).
#ifndef PGVD_SPLIT1_BLOCK
#define PGVD_SPLIT1_BLOCK 512
#endif
constexpr int kSplit1Block = PGVD_SPLIT1_BLOCK;
constexpr int kSplit1Tile = kSplit1Block * kSplitIpt;
__global__ __launch_bounds__(kSplit1Block) void msd_split1_kernel(
    const uint64_t *__restrict__ in, uint64_t *__restrict__ out, uint32_t n, int shift1, int nb1,
    uint32_t *__restrict__ cursor, const RunState *st) {
  if (!st->msd_ok) return;
  __shared__ uint64_t s_keys[kSplit1Tile];
  const uint32_t lo = blockIdx.x * (uint32_t)kSplit1Tile;
  const uint32_t hi = min(lo + (uint32_t)kSplit1Tile, n);
  split_tile<kSplit1Block>(in, out, n, lo, hi, shift1, nb1, cursor + kCursor1, 0u, kMsdCStride,
                           s_keys);
}

// Second scatter: tiles within each first-level group g (keys
// [start[g nb2], start[(g+1) nb2])), by the next b2 bits into the group's
// nb2 buckets.
__global__ __launch_bounds__(kBlock) void msd_split2_kernel(const uint64_t *__restrict__ in,
                                                            uint64_t *__restrict__ out,
                                                            uint32_t n, int shift, int nb2,
                                                            int nb1,
                                                            const uint32_t *__restrict__ start,
                                                            uint32_t *__restrict__ cursor,
                                                            const RunState *st) {
  if (!st->msd_ok) return;
  __shared__ uint64_t s_keys[kSplitTile];
  const uint32_t tile = blockIdx.x;
  const uint32_t *ts = cursor + kTileStart;
  if (tile >= ts[nb1]) return;
  int g = 0;   // the group holding this tile: ts[g] <= tile < ts[g + 1]
  for (int step = kSplitMaxDigits / 2; step >= 1; step >>= 1)
    if (g + step < nb1 && ts[g + step] <= tile) g += step;
  const uint32_t glo = start[g * nb2], ghi = start[(g + 1) * nb2];
  const uint32_t lo = glo + (tile - ts[g]) * (uint32_t)kSplitTile;
  const uint32_t hi = min(lo + (uint32_t)kSplitTile, ghi);
  split_tile<kBlock>(in, out, n, lo, hi, shift, nb2, cursor, (uint32_t)(g * nb2), 1, s_keys);
}

// Stable LSD passes, in LDS, over the BS IPT keys k[] of a workgroup: key
// index le = 64 IPT w + 64 j + lane lives in wave w, register j, and each
// pass ranks by the ballot matching of the onesweep tile, so it is stable.
// On return s_keys[0 .. BS IPT) holds the keys in order of bits
// [0, 8 (passes - 1) + last_bits).  The caller fills the slots past the
// bucket with ~0 at the highest indices: ~0 then stays behind every real key
// (a larger or equal digit in every pass, and later in the stable order).
// s_cnt: NW 256 words, s_dig: 256 words, s_w: NW words.
template <int BS, int IPT>
__device__ inline void lds_radix_passes(uint64_t (&k)[IPT], int passes, int last_bits,
                                        uint64_t *s_keys, uint32_t *s_cnt, uint32_t *s_dig,
                                        uint32_t *s_w) {
  constexpr int NW = BS / 64;
  static_assert(64 * IPT < (1 << 16), "a rank fits 16 bits");
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint32_t wofs = (uint32_t)w * 64u * IPT;
  for (int p = 0; p < passes; ++p) {
    const int sh = 8 * p;
    const int nbits = p == passes - 1 ? last_bits : 8;   // uniform
    const uint32_t dmask = (1u << nbits) - 1;
    for (int i = t; i < NW * 256; i += BS) s_cnt[i] = 0;
    __syncthreads();
    uint32_t rk[IPT];   // rank within (wave, digit) | digit << 16
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t d = (uint32_t)(k[j] >> sh) & dmask;
      uint32_t plo = ~0u, phi = ~0u;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        if (b < nbits) {
          const int32_t sb = __builtin_amdgcn_sbfe((int32_t)d, b, 1);
          const uint64_t bb = __ballot(sb < 0);
          plo = __builtin_amdgcn_bitop3_b32(plo, (uint32_t)bb, (uint32_t)sb, 0x90);
          phi = __builtin_amdgcn_bitop3_b32(phi, (uint32_t)(bb >> 32), (uint32_t)sb, 0x90);
        }
      }
      const uint32_t c = s_cnt[w * 256 + d];
      const uint32_t below = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
      rk[j] = (c + below) | (d << 16);
      // every lane of the group stores the same new count
      s_cnt[w * 256 + d] = c + (uint32_t)(__popc(plo) + __popc(phi));
    }
    __syncthreads();
    // thread t < 256 owns digit t: wave offsets, then the digit's start
    uint32_t tot = 0;
    if (t < 256) {
#pragma unroll
      for (int ww = 0; ww < NW; ++ww) {
        const uint32_t c = s_cnt[ww * 256 + t];
        s_cnt[ww * 256 + t] = tot;
        tot += c;
      }
    }
    uint32_t all;
    const uint32_t dex = block_excl_sum<uint32_t, NW>(tot, s_w, &all);
    if (t < 256) s_dig[t] = dex;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t d = rk[j] >> 16;
      s_keys[(rk[j] & 0xffffu) + s_dig[d] + s_cnt[w * 256 + d]] = k[j];
    }
    __syncthreads();
    if (p + 1 < passes) {
#pragma unroll
      for (int j = 0; j < IPT; ++j) k[j] = s_keys[wofs + j * 64 + lane];
      __syncthreads();   // s_keys and s_cnt are rewritten by the next pass
    }
  }
}

// Sub-buckets of the counting pass: the 12 bits below a bucket's own.
constexpr int kSubBits = 12;
constexpr int kSubBins = 1 << kSubBits;
// Largest sub-bucket the counting pass puts in order by ranking its keys.
constexpr uint32_t kSubMaxGroup = 32;
// Threads of a bucket workgroup.  The LDS of one bucket (~58 KB) allows two
// workgroups per CU, so the waves that hide the loads' latency come from
// large workgroups (and registers are capped so that two fit: BS/128 waves
// per SIMD).
#ifndef PGVD_MSD_BLOCK
#define PGVD_MSD_BLOCK 1024
#endif
constexpr int kBucketBlock = PGVD_MSD_BLOCK;
#ifndef PGVD_MSD_WPE
#define PGVD_MSD_WPE (PGVD_MSD_BLOCK / 128)
#endif

// One workgroup per bucket: bucket b holds cnt <= kMsdCap keys in[s0 ..),
// all sharing the bits from `shift` up; they are sorted in LDS and written
// to out[s0 ..).  Equal keys share a bucket (and a sub-bucket), so the
// workgroup also sees every duplicate key: when it finds one -- or the
// zero key at the front while a virtual zero key is prepended (prepend),
// which UniqueSorted would merge -- it sets st->msd_dups, and only then does
// unique_kernel run over the sorted keys.
//
// Sorting: every key gets a slot in its sub-bucket (its next 12 bits, bits
// shift-12 .. shift-1) from an LDS atomic, in any order -- keys without
// payload need no stable order -- and then its place from its rank on the
// whole key within the sub-bucket (~1 key for uniform keys).  A bucket
// with a sub-bucket of more than kSubMaxGroup keys (clustered keys) is
// sorted by stable LDS radix passes on its low `shift` bits instead.
template <int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(PGVD_MSD_WPE))) void
msd_bucket_kernel(const uint64_t *__restrict__ in,
                                                        uint64_t *__restrict__ out,
                                                        const uint32_t *__restrict__ start,
                                                        int shift, int passes, int last_bits,
                                                        uint32_t prepend, RunState *st) {
  constexpr int IPT = kMsdCap / BS;
  static_assert(IPT * BS == kMsdCap, "the workgroup holds a full bucket");
  constexpr int NW = BS / 64;
  constexpr int PER = kSubBins / BS;   // consecutive sub-buckets per thread
  static_assert(NW * 256 <= kSubBins, "the radix counters fit the sub-bucket space");
  if (!st->msd_ok) return;
  __shared__ uint64_t s_keys[kMsdCap];
  __shared__ uint32_t s_sub[kSubBins + 1];   // counts, then starts; radix counters
  __shared__ uint32_t s_dig[256], s_w[NW], s_mx[NW];
  const int t = threadIdx.x;
  const uint32_t s0 = start[blockIdx.x];
  const uint32_t cnt = min(start[blockIdx.x + 1] - s0, (uint32_t)kMsdCap);
  if (cnt == 0) return;
  for (int i = t; i < kSubBins; i += BS) s_sub[i] = 0;
  __syncthreads();
  const int sub_shift = shift - kSubBits;
  uint64_t k[IPT];
  uint32_t rs[IPT];   // slot in sub-bucket | sub-bucket << 16
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = j * BS + t;
    k[j] = i < cnt ? in[s0 + i] : 0ull;
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = j * BS + t;
    rs[j] = 0;
    if (i < cnt) {
      const uint32_t sb = (uint32_t)(k[j] >> sub_shift) & (kSubBins - 1);
      rs[j] = atomicAdd(&s_sub[sb], 1u) | (sb << 16);
    }
  }
  __syncthreads();
  // starts of thread t's PER consecutive sub-buckets, and the largest one
  uint32_t cp[PER], run = 0, mx = 0;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    cp[i] = s_sub[t * PER + i];
    run += cp[i];
    mx = max(mx, cp[i]);
  }
  uint32_t total;
  uint32_t ex = block_excl_sum<uint32_t, NW>(run, s_w, &total);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, d, 64));
  if ((t & 63) == 0) s_mx[t >> 6] = mx;
  __syncthreads();   // every thread has read its counts
  uint32_t gmax = (t & 63) < NW ? s_mx[t & 63] : 0u;   // uniform after the reduction
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) gmax = max(gmax, (uint32_t)__shfl_xor((int)gmax, d, 64));
  if (gmax <= kSubMaxGroup) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      s_sub[t * PER + i] = ex;
      ex += cp[i];
    }
    if (t == 0) s_sub[kSubBins] = total;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t i = j * BS + t;
      if (i < cnt) s_keys[s_sub[rs[j] >> 16] + (rs[j] & 0xffffu)] = k[j];
    }
    __syncthreads();
    // every key's rank within its sub-bucket [b0, b1) (at most kSubMaxGroup
    // keys; ties by position), then one permutation in LDS
    uint64_t v[IPT];
    uint32_t dst[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t i = j * BS + t;
      dst[j] = i;
      if (i < cnt) {
        v[j] = s_keys[i];
        const uint32_t sb = (uint32_t)(v[j] >> sub_shift) & (kSubBins - 1);
        const uint32_t b0 = s_sub[sb], b1 = s_sub[sb + 1];
        uint32_t r = b0;
        for (uint32_t x = b0; x < b1; ++x) {
          const uint64_t w = s_keys[x];
          r += (w < v[j] || (w == v[j] && x < i)) ? 1u : 0u;
        }
        dst[j] = r;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j)
      if (j * BS + t < cnt) s_keys[dst[j]] = v[j];
  } else {
    // keys in the radix passes' layout, the slots past the bucket last
    const int lane = t & 63, w = t >> 6;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t le = (uint32_t)w * 64u * IPT + j * 64 + lane;
      k[j] = le < cnt ? in[s0 + le] : ~0ull;
    }
    lds_radix_passes<BS, IPT>(k, passes, last_bits, s_keys, s_sub, s_dig, s_w);
  }
  __syncthreads();
  // write out, looking for equal neighbours
  bool dup = false;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = j * BS + t;
    if (i < cnt) {
      const uint64_t v = s_keys[i];
      dup |= i > 0 ? v == s_keys[i - 1] : (s0 == 0 && prepend && v == 0ull);
      out[s0 + i] = v;
    }
  }
  if (__any(dup) && (t & 63) == 0)
    __hip_atomic_store(&st->msd_dups, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

int msd_bits_for(int64_t n) {
  if (n < ((int64_t)1 << 23)) return 0;
  for (int b = 10; b <= 12; ++b)
    if ((double)n <= 0.9 * kMsdCap * (double)(1 << b)) return b;
  return 0;
}

void launch_msd_sort(uint64_t *keys, uint64_t *alt, uint64_t *out, int prepend, int64_t n,
                     int mbits, int b, const uint32_t *hist, uint32_t *start, uint32_t *cursor,
                     RunState *st, hipStream_t s) {
  const int nb = 1 << b, b1 = b / 2, b2 = b - b1;
  const int nb1 = 1 << b1, nb2 = 1 << b2;
  const int shift = mbits - b;             // bits each bucket is sorted on
  msd_plan_kernel<<<1, kBlock, 0, s>>>(hist, (uint32_t)n, nb, nb2, start, cursor, st);
  const unsigned tiles1 = (unsigned)((n + kSplit1Tile - 1) / kSplit1Tile);
  msd_split1_kernel<<<tiles1, kSplit1Block, 0, s>>>(keys, alt, (uint32_t)n, mbits - b1, nb1,
                                                    cursor, st);
  // the second scatter's tiles (kSplitTile keys) within each group: the sum
  // of ceil(len / tile) over the groups is at most tiles2 + nb1
  const unsigned tiles2 = (unsigned)((n + kSplitTile - 1) / kSplitTile);
  msd_split2_kernel<<<tiles2 + (unsigned)nb1, kBlock, 0, s>>>(alt, keys, (uint32_t)n, shift, nb2,
                                                              nb1, start, cursor, st);
  const int passes = (shift + 7) / 8;
  const int last_bits = shift - 8 * (passes - 1);
  msd_bucket_kernel<kBucketBlock><<<(unsigned)nb, kBucketBlock, 0, s>>>(
      keys, out, start, shift, passes, last_bits, (uint32_t)prepend, st);
}

}  // namespace pgvd
