// bucket.hip -- the whole path's sort: a most-significant-digit bucket sort
// for 32-, 64- and 128-bit keys, with or without a u32 label per key.
//
// It produces the ascending key sequence the reference's LSD split sort
// (RadixSortBigUnsigned, opencl/Kernels.cpp:219-247) produces.  Equal keys
// are indistinguishable, and no output of the whole path depends on the
// order of labels among equal keys (a run's conflict flag only asks whether
// its labels differ), so the sort need not be stable.
//
// Level 0 splits the keys by their top b bits (b chosen so that a bucket
// averages well under CAP keys) in one or two LDS-staged scatters; every
// bucket of at most CAP keys is then sorted inside one workgroup's LDS.
// Buckets above CAP (clustered keys) go to further levels: a level first ORs
// and ANDs each segment's keys, so it knows the segment's highest varying
// bit hb (of OR ^ AND) and splits it on the 8 bits below hb -- never on bits
// all its keys share.  Each level is that pass, a histogram, a per-segment
// scan and a scatter; sub-buckets of at most CAP keys become
// LDS sort items (consecutive small ones merged up to CAP), larger ones the
// next level's segments.  A level consumes at least 8 varying bits, so a
// 128-bit key needs at most 16 levels; a segment whose keys are all equal
// (OR == AND) is copied, not split.  The host reads the segment count after
// the level-0 plan and after each level's scan while the device keeps
// working, and stops launching levels when none is left.
#include "kernels.hpp"
#include "tilescan.hpp"

namespace pgvd {

namespace {

// ---- per key width ---------------------------------------------------------
// CAP keys (and labels) of one LDS sort item; BS threads of IPT keys each;
// scatter tiles of kBlock threads x TIPT keys.
template <typename K> struct Bk;
template <> struct Bk<uint32_t> { static constexpr int CAP = 8192, BS = 1024, IPT = 8, TIPT = 16; };
template <> struct Bk<uint64_t> { static constexpr int CAP = 5120, BS = 1024, IPT = 5, TIPT = 16; };
template <> struct Bk<Key128> { static constexpr int CAP = 2560, BS = 512, IPT = 5, TIPT = 8; };
// Waves per SIMD the LDS sort is compiled for: 8 (two 1024-thread
// workgroups per CU, 64 VGPRs) for level-0 buckets of 64-bit keys without
// labels, where the kernel fits them without spilling; 4 (one workgroup, or
// two of 512) where it needs more registers.
template <typename K, bool LAB, bool MATCH> constexpr int sort_wpe() {
  return (sizeof(K) == 8 && !LAB && !MATCH) ? 8 : 4;
}

// 64-bit halves of a key (OR / AND accumulation)
__device__ inline uint64_t klo(uint32_t a) { return a; }
__device__ inline uint64_t klo(uint64_t a) { return a; }
__device__ inline uint64_t klo(Key128 a) { return a.lo; }
__device__ inline uint64_t khi(uint32_t) { return 0; }
__device__ inline uint64_t khi(uint64_t) { return 0; }
__device__ inline uint64_t khi(Key128 a) { return a.hi; }

__device__ inline bool kless(uint32_t a, uint32_t b) { return a < b; }
__device__ inline bool kless(uint64_t a, uint64_t b) { return a < b; }
__device__ inline bool kless(Key128 a, Key128 b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }

// nbits (<= 14) bits of a key from bit `shift` up
template <typename K>
__device__ inline uint32_t kbits(K a, int shift, uint32_t mask) {
  return klow32(kshr(a, shift)) & mask;
}

// all-ones key
template <typename K> __device__ inline K kones();
template <> __device__ inline uint32_t kones<uint32_t>() { return ~0u; }
template <> __device__ inline uint64_t kones<uint64_t>() { return ~0ull; }
template <> __device__ inline Key128 kones<Key128>() { return {~0ull, ~0ull}; }

// number of low bits that vary: bitlen(or ^ and)
__device__ inline int varying_bits(uint64_t olo, uint64_t ohi, uint64_t alo, uint64_t ahi) {
  const uint64_t xh = ohi ^ ahi, xl = olo ^ alo;
  return xh ? 128 - __builtin_clzll(xh) : (xl ? 64 - __builtin_clzll(xl) : 0);
}

// A new segment's accumulators: OR 0, AND all ones -- in the high half only
// for 128-bit keys (wide), whose OR / AND pass accumulates it; narrower keys
// keep both high halves 0, so they never count as varying.
__device__ inline void init_seg(BkSeg *sg, uint32_t start, uint32_t count, bool wide) {
  sg->start = start;
  sg->count = count;
  sg->tile_base = 0;
  sg->shift = 0;
  sg->nbits = 0;
  sg->or_lo = 0; sg->or_hi = 0;
  sg->and_lo = ~0ull; sg->and_hi = wide ? ~0ull : 0ull;
}

// ---- level-0 plan ------------------------------------------------------------
// One workgroup of 1024 threads: bucket starts from the histogram of the top
// b bits (nb = 2^b buckets, at most kBkMaxBuckets), both scatters' cursors
// and the second scatter's tile starts, and the level-1 segments (buckets
// above CAP) with their accumulators.
constexpr int kPlanBlock = 1024;
// The LSD passes take over when more than 1 / PGVD_BK_LSD_SHARE of the keys
// are left to the levels (0: never).
#ifndef PGVD_BK_LSD_SHARE
#define PGVD_BK_LSD_SHARE 4
#endif
__global__ __launch_bounds__(kPlanBlock) void bk_plan_kernel(
    const uint32_t *__restrict__ hist, uint32_t n, int nb, int nb2, int tile2, uint32_t cap,
    int wide, int lsd_ok, uint32_t *__restrict__ start, uint32_t *__restrict__ cursor,
    BkSeg *__restrict__ segs, uint32_t *__restrict__ hrows, RunState *st, RunState *host) {
  __shared__ uint32_t s_w[kPlanBlock / 64];
  __shared__ uint32_t s_nseg, s_bigkeys;
  const int t = threadIdx.x;
  if (t == 0) {
    s_nseg = 0;
    s_bigkeys = 0;
  }
  const int per = (nb + kPlanBlock - 1) / kPlanBlock;   // <= 16 consecutive buckets per thread
  uint32_t v[16];
  uint32_t run = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int b = t * per + i;
    v[i] = (i < per && b < nb) ? hist[b] : 0u;
    run += v[i];
  }
  uint32_t total;
  uint32_t ex = block_excl_sum<uint32_t, kPlanBlock / 64>(run, s_w, &total);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int b = t * per + i;
    if (i < per && b < nb) {
      start[b] = ex;
      cursor[b] = ex;
      if (v[i] > cap) {
        const uint32_t sid = atomicAdd(&s_nseg, 1u);
        init_seg(&segs[sid], ex, v[i], wide != 0);
        atomicAdd(&s_bigkeys, v[i]);
      }
      ex += v[i];
    }
  }
  if (t == 0) start[nb] = total;
  __syncthreads();   // start[] complete, s_nseg final
  const uint32_t nseg = s_nseg;
  // the level-1 segments' histogram rows start at zero
  for (uint32_t i = t; i < nseg * 256u; i += kPlanBlock) hrows[i] = 0u;
  if (t == 0) {
    // Keys that clustered so much that a quarter of them are left to the
    // levels sort faster by the LSD passes where those are few (keys of at
    // most 64 bits): then the scatters and bucket sorts do nothing, and the
    // host launches the passes over the keys as encoded.
    const bool lsd = lsd_ok && PGVD_BK_LSD_SHARE > 0 &&
                     (uint64_t)s_bigkeys * PGVD_BK_LSD_SHARE > (uint64_t)n;
    st->bk_lsd = lsd ? 1u : 0u;
    st->bk_segs = lsd ? 0u : nseg;
    st->bk_keys = lsd ? 0u : s_bigkeys;
    // the same three words straight into the host's pinned mirror, which
    // reads them once the plan's completion event has fired (no copy
    // launch between the plan and the first scatter)
    host->bk_lsd = lsd ? 1u : 0u;
    host->bk_segs = lsd ? 0u : nseg;
    host->bk_keys = lsd ? 0u : s_bigkeys;
    __threadfence_system();
    // (a histogram that does not count every key leaves the keys unsorted:
    // then the status says so and nothing downstream trusts the order)
    st->msd_ok = (total == n && !lsd) ? 1u : 0u;
    if (total != n) atomicOr(&st->status, (uint32_t)ST_TOO_LARGE);
    st->msd_dups = 0;
  }
  const int nb1 = nb / nb2;   // first-scatter groups (<= 128)
  if (t < 64) {
    // wave 0, groups 2 lane and 2 lane + 1: an exclusive scan of the groups'
    // tile counts in the second scatter
    uint32_t lo[2], tiles[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int g = 2 * t + h;
      lo[h] = 0;
      tiles[h] = 0;
      if (g < nb1) {
        lo[h] = start[g * nb2];
        tiles[h] = (start[(g + 1) * nb2] - lo[h] + tile2 - 1) / tile2;
      }
    }
    const uint32_t pair = tiles[0] + tiles[1];
    const uint32_t incl = wave_incl_scan_add(pair);
    uint32_t ex = incl - pair;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int g = 2 * t + h;
      if (g < nb1) {
        cursor[kBkCursor1 + g * kBkCStride] = lo[h];
        cursor[kBkTileStart + g] = ex;
      }
      ex += tiles[h];
    }
    if (t == 63) cursor[kBkTileStart + nb1] = incl;
  }
}

// ---- LDS-staged scatter -------------------------------------------------------
// Lanes of the wave whose NB-bit value v equals this lane's, among the lanes
// with `active` set (peer matching by one ballot per bit, as the LSD ranking
// does): clustered keys put many lanes on one counter, and one LDS atomic
// per distinct value and wave replaces one per lane.
template <int NB>
__device__ inline uint64_t wave_peers(uint32_t v, bool active) {
  const uint64_t act = __ballot(active);
  uint32_t plo = (uint32_t)act, phi = (uint32_t)(act >> 32);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int32_t sb = __builtin_amdgcn_sbfe((int32_t)v, b, 1);
    const uint64_t bb = __ballot(sb < 0);
    plo = __builtin_amdgcn_bitop3_b32(plo, (uint32_t)bb, (uint32_t)sb, 0x90);
    phi = __builtin_amdgcn_bitop3_b32(phi, (uint32_t)(bb >> 32), (uint32_t)sb, 0x90);
  }
  return active ? (((uint64_t)phi << 32) | plo) : 0ull;
}

// A slot for this lane in counter[v]: the wave's lanes of equal v take
// consecutive slots from one atomic of their lowest lane.  Call with the
// whole wave converged.
template <int NB>
__device__ inline uint32_t matched_slot(uint32_t *counter, uint32_t v, bool active) {
  const uint64_t pm = wave_peers<NB>(v, active);
  const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(pm >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)pm, 0u));
  uint32_t base = 0;
  if (active && below == 0) base = atomicAdd(&counter[v], (uint32_t)__popcll(pm));
  const int leader = pm ? __ffsll((unsigned long long)pm) - 1 : 0;
  base = (uint32_t)__shfl((int)base, leader, 64);
  return base + below;
}

// One tile of a kBlock-thread workgroup: keys (and labels) in[lo .. hi) go to
// out in runs by digit d = (key >> dshift) & (nd - 1) (nd <= 256), the run of
// digit d starting at an atomic claim on gcur[d * cstride].  MATCH: the
// tile's digit counts by peer matching (clustered keys).
template <typename K, bool LAB>
struct TileIO {
  const K *in;
  const uint32_t *vin;
  K *out;
  uint32_t *vout;
};

template <typename K, bool LAB, int TIPT, bool MATCH>
__device__ inline void scatter_tile(const TileIO<K, LAB> &io, uint32_t n, uint32_t lo, uint32_t hi,
                                    int dshift, int nd, uint32_t *gcur, int cstride, K *s_keys,
                                    uint32_t *s_vals) {
  __shared__ uint32_t s_cnt[256], s_lstart[256], s_gstart[256], s_w4[4];
  const int t = threadIdx.x;
  const uint32_t dmask = (uint32_t)nd - 1;
  if (t < nd) s_cnt[t] = 0;
  K k[TIPT];
  uint32_t lv[TIPT];
  uint32_t dr[TIPT];
#pragma unroll
  for (int j = 0; j < TIPT; ++j) {
    const uint32_t i = lo + j * kBlock + t;
    k[j] = io.in[i < hi ? i : lo];
    if (LAB) lv[j] = io.vin[i < hi ? i : lo];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < TIPT; ++j) {
    const uint32_t i = lo + j * kBlock + t;
    const uint32_t d = kbits(k[j], dshift, dmask);
    if (MATCH) {
      dr[j] = matched_slot<8>(s_cnt, d, i < hi) | (d << 16);
    } else {
      dr[j] = 0;
      if (i < hi) dr[j] = atomicAdd(&s_cnt[d], 1u) | (d << 16);
    }
  }
  __syncthreads();
  {
    // thread d: the digit's start in the tile and its run's claim
    const uint32_t c = t < nd ? s_cnt[t] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_sum<uint32_t>(c, s_w4, &tot);
    if (t < nd) {
      s_lstart[t] = ex;
      s_gstart[t] = c ? atomicAdd(&gcur[t * cstride], c) : 0u;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < TIPT; ++j) {
    const uint32_t i = lo + j * kBlock + t;
    if (i < hi) {
      const uint32_t p = s_lstart[dr[j] >> 16] + (dr[j] & 0xffffu);
      s_keys[p] = k[j];
      if (LAB) s_vals[p] = lv[j];
    }
  }
  __syncthreads();
  const uint32_t cnt = hi - lo;
  for (uint32_t p = t; p < cnt; p += kBlock) {
    const K key = s_keys[p];
    const uint32_t d = kbits(key, dshift, dmask);
    const uint32_t pos = s_gstart[d] + (p - s_lstart[d]);
    if (pos < n) {   // always true when the histogram matches the keys
      io.out[pos] = key;
      if (LAB) io.vout[pos] = s_vals[p];
    }
  }
}

template <typename K, bool LAB>
struct StageMem {
  static constexpr int TILE = kBlock * Bk<K>::TIPT;
  K keys[TILE];
  uint32_t vals[LAB ? TILE : 1];
};

// First scatter: contiguous tiles of the whole array by the top b1 bits
// (digits nd1 <= 128).
template <typename K, bool LAB>
__global__ __launch_bounds__(kBlock) void bk_split1_kernel(TileIO<K, LAB> io, uint32_t n,
                                                           int shift1, int nd1,
                                                           uint32_t *__restrict__ cursor,
                                                           int cstride, const RunState *st) {
  __shared__ StageMem<K, LAB> sm;
  constexpr int TILE = StageMem<K, LAB>::TILE;
  if (st->bk_lsd) return;
  const uint32_t lo = blockIdx.x * (uint32_t)TILE;
  if (lo >= n) return;
  const uint32_t hi = min(lo + (uint32_t)TILE, n);
  scatter_tile<K, LAB, Bk<K>::TIPT, false>(io, n, lo, hi, shift1, nd1, cursor, cstride, sm.keys,
                                           sm.vals);
}

// Second scatter: tiles within each first-level group g (keys
// [start[g nb2], start[(g+1) nb2])) by the next b2 bits into the group's
// nb2 buckets.
template <typename K, bool LAB>
__global__ __launch_bounds__(kBlock) void bk_split2_kernel(TileIO<K, LAB> io, uint32_t n,
                                                           int shift, int nb2, int nb1,
                                                           const uint32_t *__restrict__ start,
                                                           uint32_t *__restrict__ cursor,
                                                           const RunState *st) {
  __shared__ StageMem<K, LAB> sm;
  constexpr int TILE = StageMem<K, LAB>::TILE;
  if (st->bk_lsd) return;
  const uint32_t tile = blockIdx.x;
  const uint32_t *ts = cursor + kBkTileStart;
  if (tile >= ts[nb1]) return;
  int g = 0;   // the group holding this tile: ts[g] <= tile < ts[g + 1]
  for (int step = 64; step >= 1; step >>= 1)
    if (g + step < nb1 && ts[g + step] <= tile) g += step;
  const uint32_t glo = start[g * nb2], ghi = start[(g + 1) * nb2];
  const uint32_t lo = glo + (tile - ts[g]) * (uint32_t)TILE;
  const uint32_t hi = min(lo + (uint32_t)TILE, ghi);
  scatter_tile<K, LAB, Bk<K>::TIPT, false>(io, n, lo, hi, shift, nb2, cursor + g * nb2, 1,
                                           sm.keys, sm.vals);
}

// ---- LDS sort of one item ---------------------------------------------------------
// Stable LSD passes in LDS over the BS IPT keys k[] (and labels v[]) of a
// workgroup; key index le = 64 IPT w + 64 j + lane lives in wave w,
// register j; ranks by ballot matching.  On return s_keys / s_vals hold the
// pairs in order of bits [0, 8 (passes - 1) + last_bits).  Slots past the
// item hold all-ones keys at the highest indices, which stay behind every
// real key.  s_cnt: NW 256 words, s_dig 256, s_w NW.
template <typename K, bool LAB, int BS, int IPT>
__device__ inline void lds_radix(K (&k)[IPT], uint32_t (&v)[IPT], int passes, int last_bits,
                                 K *s_keys, uint32_t *s_vals, uint32_t *s_cnt, uint32_t *s_dig,
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
      const uint32_t d = kdigit(k[j], sh) & dmask;
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
      const uint32_t q = (rk[j] & 0xffffu) + s_dig[d] + s_cnt[w * 256 + d];
      s_keys[q] = k[j];
      if (LAB) s_vals[q] = v[j];
    }
    __syncthreads();
    if (p + 1 < passes) {
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        k[j] = s_keys[wofs + j * 64 + lane];
        if (LAB) v[j] = s_vals[wofs + j * 64 + lane];
      }
      __syncthreads();   // s_keys and s_cnt are rewritten by the next pass
    }
  }
}

// ---- merge sort of one item in LDS -----------------------------------------------
// (key, position) pairs in a strict total order: equal keys by position, so
// the sort is stable and slots past the item (all-ones keys, positions >=
// cnt) stay behind every real key.
template <typename K>
__device__ inline bool pair_lt(const K &a, uint32_t ia, const K &b, uint32_t ib) {
  return kless(a, b) || (keq(a, b) && ia < ib);
}

// The item's cnt keys (and labels) from in[s0 ..] sorted by merging: each
// thread sorts IPT consecutive keys in registers, then log2(BS) rounds merge
// runs pairwise, every thread producing IPT outputs of its run pair from a
// merge-path split found by binary search.  The cost does not depend on how
// the keys are spread, which suits wide keys that no few digits separate
// (sparse 128-bit keys on curves).  Positions travel in s_idx (16 bits);
// labels, staged in s_vals in input order, are gathered by position at the
// end.  On return s_keys[0 .. cnt) (and s_vals) hold the item in order.
template <typename K, bool LAB, int BS, int IPT>
__device__ inline void lds_merge_sort(const K *__restrict__ in, const uint32_t *__restrict__ vin,
                                      uint32_t s0, uint32_t cnt, K *s_keys, uint32_t *s_vals,
                                      uint16_t *s_idx) {
  constexpr uint32_t CAP = (uint32_t)BS * IPT;
  static_assert(CAP <= 65536, "positions fit 16 bits");
  const uint32_t t = threadIdx.x;
  const uint32_t p0 = t * IPT;
  K k[IPT];
  uint32_t x[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t p = p0 + j;
    k[j] = p < cnt ? in[s0 + p] : kones<K>();
    x[j] = p;
    if (LAB && p < cnt) s_vals[p] = vin[s0 + p];
  }
  // odd-even transposition sort of the thread's IPT pairs
#pragma unroll
  for (int r = 0; r < IPT; ++r) {
#pragma unroll
    for (int j = r & 1; j + 1 < IPT; j += 2) {
      if (pair_lt(k[j + 1], x[j + 1], k[j], x[j])) {
        const K tk = k[j]; k[j] = k[j + 1]; k[j + 1] = tk;
        const uint32_t tx = x[j]; x[j] = x[j + 1]; x[j + 1] = tx;
      }
    }
  }
  for (uint32_t w = IPT; w < CAP; w *= 2) {
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      s_keys[p0 + j] = k[j];
      s_idx[p0 + j] = (uint16_t)x[j];
    }
    __syncthreads();
    const uint32_t base = p0 / (2 * w) * (2 * w);
    const uint32_t d = p0 - base;   // this thread's first output in the merged pair
    const uint32_t A = base, B = base + w;
    // i keys from run A and d - i from run B come first
    uint32_t lo = d > w ? d - w : 0u, hi = d < w ? d : w;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      const uint32_t bj = B + d - 1 - mid;
      if (pair_lt(s_keys[bj], (uint32_t)s_idx[bj], s_keys[A + mid], (uint32_t)s_idx[A + mid]))
        hi = mid;
      else
        lo = mid + 1;
    }
    uint32_t ai = A + lo, bi = B + (d - lo);
    const uint32_t ae = A + w, be = B + w;
    K ka = ai < ae ? s_keys[ai] : kones<K>(), kb = bi < be ? s_keys[bi] : kones<K>();
    uint32_t xa = ai < ae ? s_idx[ai] : 0xffffffffu, xb = bi < be ? s_idx[bi] : 0xffffffffu;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const bool take_a = pair_lt(ka, xa, kb, xb);   // an exhausted run compares last
      k[j] = take_a ? ka : kb;
      x[j] = take_a ? xa : xb;
      if (take_a) {
        ++ai;
        ka = ai < ae ? s_keys[ai] : kones<K>();
        xa = ai < ae ? s_idx[ai] : 0xffffffffu;
      } else {
        ++bi;
        kb = bi < be ? s_keys[bi] : kones<K>();
        xb = bi < be ? s_idx[bi] : 0xffffffffu;
      }
    }
    __syncthreads();   // every read of this round is done before the next writes
  }
  uint32_t lab[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) lab[j] = (LAB && x[j] < cnt) ? s_vals[x[j]] : 0u;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    s_keys[p0 + j] = k[j];
    if (LAB) s_vals[p0 + j] = lab[j];
  }
}

constexpr int kSubBits = 12;                 // sub-buckets of the counting pass
constexpr int kSubBins = 1 << kSubBits;
// Largest sub-bucket ranked in place (rank = keys below it in the
// sub-bucket, one LDS read each, so the cost grows with its size): level-0
// buckets of uniform keys hold ~1 key per sub-bucket; the levels' items
// (PGVD_BK_RANK_MAX) may hold keys on curves and surfaces, whose top 12
// varying bits leave larger sub-buckets, sorted past this size by LDS radix
// passes or the merge sort.
#ifndef PGVD_BK_RANK_MAX
#define PGVD_BK_RANK_MAX 32
#endif
constexpr uint32_t kSubMaxGroup = 32;
constexpr uint32_t kSubMaxGroupList = PGVD_BK_RANK_MAX;
// Past ranking: LDS radix passes over at most this many varying bits (three
// passes), a merge sort over more.
constexpr int kRadixMaxBits = 24;

template <typename K, bool LAB>
struct SortArgs {
  const K *kin[2];          // the two key buffers (list items name theirs)
  const uint32_t *vin[2];   // their labels
  const K *k0;              // level 0: where the buckets are
  const uint32_t *v0;
  K *out;
  uint32_t *vout;
  const uint32_t *start;    // level 0: bucket starts (nb + 1)
  const BkItem *items;      // list mode
  uint32_t prepend;
};

// One item of at most CAP keys per workgroup, sorted into out at its own
// positions.  Level-0 mode (list == 0): item b is bucket b, its keys sharing
// their bits from sbits0 = mbits - b up (buckets above CAP are left to the
// levels).  List mode: item blockIdx.x of items[] (count st->bk_items), its
// sbits lowered to the bits that vary among its keys.  An item with sbits 0
// holds equal keys and is copied.  Otherwise a counting pass on the 12 bits
// below sbits, then each key's rank within its sub-bucket -- or, when a
// sub-bucket holds more than 32 keys, LSD passes in LDS over the bits below
// sbits (at most 24 of them) or a merge sort (more).  Equal
// keys share an item, so the workgroup sees every duplicate: it raises
// st->msd_dups for equal neighbours, or for the zero key at the front when
// a zero key is prepended (UniqueSorted would merge them); only then does
// unique_kernel run over the sorted keys.  MATCH: the counting pass by peer
// matching (items of the levels hold clustered keys).
template <typename K, bool LAB, int BS, int IPT, int WPE, bool MATCH>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WPE))) void
bk_sort_kernel(SortArgs<K, LAB> a, int list, int sbits0, RunState *st) {
  constexpr int CAP = BS * IPT;
  constexpr int NW = BS / 64;
  constexpr int PER = kSubBins / BS;   // consecutive sub-buckets per thread
  static_assert(NW * 256 <= kSubBins, "the radix counters fit the sub-bucket space");
  __shared__ K s_keys[CAP];
  __shared__ uint32_t s_vals[LAB ? CAP : 1];
  __shared__ uint32_t s_sub[kSubBins + 1];   // counts, then starts; radix counters
  __shared__ uint32_t s_dig[256], s_w[NW], s_mx[NW];
  __shared__ unsigned long long s_red[MATCH ? NW : 1][4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t s0, cnt;
  int sbits;
  const K *in;
  const uint32_t *vin;
  if (list) {
    if (blockIdx.x >= st->bk_items) return;
    const BkItem it = a.items[blockIdx.x];
    s0 = it.start;
    cnt = min(it.count, (uint32_t)CAP);
    sbits = (int)it.sbits;
    // (selects, not an index into the argument struct, which would move it
    // to scratch memory)
    in = (it.parity & 1) ? a.kin[1] : a.kin[0];
    vin = (it.parity & 1) ? a.vin[1] : a.vin[0];
  } else {
    if (st->bk_lsd) return;
    s0 = a.start[blockIdx.x];
    cnt = a.start[blockIdx.x + 1] - s0;
    if (cnt > (uint32_t)CAP) return;   // a level's segment
    sbits = sbits0;
    in = a.k0;
    vin = a.v0;
  }
  if (cnt == 0) return;
  for (int i = t; i < kSubBins; i += BS) s_sub[i] = 0;
  K k[IPT];
  uint32_t v[IPT];
  uint32_t rs[IPT];   // slot in sub-bucket | sub-bucket << 16
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = j * BS + t;
    k[j] = i < cnt ? in[s0 + i] : kzero_value<K>();
    v[j] = (LAB && i < cnt) ? vin[s0 + i] : 0u;
  }
  if constexpr (MATCH) {
    // items of the levels: sort on the bits that vary in the item (OR ^
    // AND of its keys), so that the counting pass splits clustered keys by
    // their top 12 varying bits
    uint64_t olo = 0, ohi = 0, alo = ~0ull, ahi = ~0ull;
#pragma unroll
    for (int j = 0; j < IPT; ++j)
      if (j * BS + t < cnt) {
        olo |= klo(k[j]); ohi |= khi(k[j]);
        alo &= klo(k[j]); ahi &= khi(k[j]);
      }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      olo |= (uint64_t)__shfl_xor((unsigned long long)olo, d, 64);
      ohi |= (uint64_t)__shfl_xor((unsigned long long)ohi, d, 64);
      alo &= (uint64_t)__shfl_xor((unsigned long long)alo, d, 64);
      ahi &= (uint64_t)__shfl_xor((unsigned long long)ahi, d, 64);
    }
    if (lane == 0) {
      s_red[w][0] = olo; s_red[w][1] = ohi; s_red[w][2] = alo; s_red[w][3] = ahi;
    }
    __syncthreads();
    olo = 0; ohi = 0; alo = ~0ull; ahi = ~0ull;
#pragma unroll
    for (int ww = 0; ww < NW; ++ww) {
      olo |= s_red[ww][0]; ohi |= s_red[ww][1]; alo &= s_red[ww][2]; ahi &= s_red[ww][3];
    }
    sbits = min(sbits, varying_bits(olo, ohi, alo, ahi));   // uniform
  }
  __syncthreads();
  if (sbits == 0) {
    // equal keys: written back as they are
    bool dup = false;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t i = j * BS + t;
      if (i < cnt) {
        dup |= i > 0 || (s0 == 0 && a.prepend && kzero(k[j]));
        a.out[s0 + i] = k[j];
        if (LAB) a.vout[s0 + i] = v[j];
      }
    }
    if (__any(dup) && lane == 0)
      raise_flag(&st->msd_dups);
    return;
  }
  const int sub_shift = sbits > kSubBits ? sbits - kSubBits : 0;
  const uint32_t sub_mask = (1u << (sbits - sub_shift)) - 1u;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = j * BS + t;
    const uint32_t sb = kbits(k[j], sub_shift, sub_mask);
    if (MATCH) {
      rs[j] = matched_slot<kSubBits>(s_sub, sb, i < cnt) | (sb << 16);
    } else {
      rs[j] = 0;
      if (i < cnt) rs[j] = atomicAdd(&s_sub[sb], 1u) | (sb << 16);
    }
  }
  __syncthreads();
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
  if (lane == 0) s_mx[w] = mx;
  __syncthreads();   // every thread has read its counts
  uint32_t gmax = lane < NW ? s_mx[lane] : 0u;   // uniform after the reduction
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) gmax = max(gmax, (uint32_t)__shfl_xor((int)gmax, d, 64));
  // sbits <= 12: a sub-bucket holds equal keys, the counting pass places them
  if (sbits <= kSubBits || gmax <= (MATCH ? kSubMaxGroupList : kSubMaxGroup)) {
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
      if (i < cnt) {
        const uint32_t q = s_sub[rs[j] >> 16] + (rs[j] & 0xffffu);
        s_keys[q] = k[j];
        if (LAB) s_vals[q] = v[j];
      }
    }
    if (sbits > kSubBits) {
      __syncthreads();
      // every key's rank within its sub-bucket [b0, b1) (at most 32 keys;
      // ties by position), then one permutation in LDS
      uint32_t dst[IPT];
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t i = j * BS + t;
        dst[j] = i;
        if (i < cnt) {
          k[j] = s_keys[i];
          if (LAB) v[j] = s_vals[i];
          const uint32_t sb = kbits(k[j], sub_shift, sub_mask);
          const uint32_t b0 = s_sub[sb], b1 = s_sub[sb + 1];
          uint32_t r = b0;
          for (uint32_t x = b0; x < b1; ++x) {
            const K y = s_keys[x];
            r += (kless(y, k[j]) || (keq(y, k[j]) && x < i)) ? 1u : 0u;
          }
          dst[j] = r;
        }
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < IPT; ++j)
        if (j * BS + t < cnt) {
          s_keys[dst[j]] = k[j];
          if (LAB) s_vals[dst[j]] = v[j];
        }
    }
  } else if (sbits <= kRadixMaxBits || WPE > 4) {
    // (the merge sort is compiled where the kernel has the registers for
    // it: at most 4 waves per SIMD)
    // keys in the radix passes' layout, the slots past the item last; the
    // passes cover the item's bits below sbits
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t le = (uint32_t)w * 64u * IPT + j * 64 + lane;
      k[j] = le < cnt ? in[s0 + le] : kones<K>();
      if (LAB) v[j] = le < cnt ? vin[s0 + le] : 0u;
    }
    const int eff = sbits;
    const int passes = (eff + 7) / 8;
    lds_radix<K, LAB, BS, IPT>(k, v, passes, eff - 8 * (passes - 1), s_keys, s_vals, s_sub, s_dig,
                               s_w);
  } else if constexpr (WPE <= 4) {
    // many varying bits and crowded sub-buckets: merge sort (the counters
    // are no longer needed; their space holds the positions)
    lds_merge_sort<K, LAB, BS, IPT>(in, vin, s0, cnt, s_keys, s_vals,
                                    reinterpret_cast<uint16_t *>(s_sub));
  }
  __syncthreads();
  // write out, looking for equal neighbours
  bool dup = false;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = j * BS + t;
    if (i < cnt) {
      const K x = s_keys[i];
      dup |= i > 0 ? keq(x, s_keys[i - 1]) : (s0 == 0 && a.prepend && kzero(x));
      a.out[s0 + i] = x;
      if (LAB) a.vout[s0 + i] = s_vals[i];
    }
  }
  if (__any(dup) && lane == 0)
    raise_flag(&st->msd_dups);
}

// ---- levels for segments above CAP -------------------------------------------------------
// A level over its segments (more than CAP keys each, in one buffer):
//   lprep     one workgroup: the tiles of every segment (kBlock x TIPT keys)
//   lorand    per tile: OR and AND of the keys into the segment
//   lhist     per tile: counts of the segment's digit, the 8 bits below its
//             highest varying bit (segments whose keys are all equal skip)
//   lscan     per segment: digit starts (cursors); digits above CAP become
//             the next level's segments, the others LDS sort items (small
//             consecutive ones merged up to CAP); an all-equal segment
//             becomes copy items
//   lscatter  per tile: the keys to their digits' runs in the other buffer
// Counting in lhist and lscatter is by peer matching: a segment's keys are
// clustered.
template <typename K> constexpr int level_tile() { return kBlock * Bk<K>::TIPT; }

// The segment holding tile `tile` of a level (segs[i].tile_base ascending).
__device__ inline uint32_t seg_of_tile(const BkSeg *segs, uint32_t nseg, uint32_t tile) {
  uint32_t lo = 0, hi = nseg;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (segs[mid].tile_base <= tile) lo = mid; else hi = mid;
  }
  return lo;
}

// A segment's digit from its accumulated OR / AND: the top 8 (or fewer)
// varying bits.  nbits 0: all keys equal.
__device__ inline void seg_digit(const BkSeg &sg, uint32_t &shift, uint32_t &nbits) {
  const int eff = varying_bits(sg.or_lo, sg.or_hi, sg.and_lo, sg.and_hi);
  shift = eff > 8 ? (uint32_t)(eff - 8) : 0u;
  nbits = (uint32_t)eff - shift;
}

__global__ __launch_bounds__(kPlanBlock) void bk_lprep_kernel(BkSeg *__restrict__ segs, int tile,
                                                             RunState *st) {
  __shared__ uint32_t s_w[kPlanBlock / 64];
  __shared__ uint32_t s_carry;
  const uint32_t nseg = st->bk_segs;
  const int t = threadIdx.x;
  if (t == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < nseg; base += kPlanBlock) {
    const uint32_t i = base + t;
    const uint32_t tiles = i < nseg ? (segs[i].count + tile - 1) / tile : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_sum<uint32_t, kPlanBlock / 64>(tiles, s_w, &total);
    if (i < nseg) segs[i].tile_base = s_carry + ex;
    __syncthreads();
    if (t == 0) s_carry += total;
    __syncthreads();
  }
  if (t == 0) st->bk_tiles = s_carry;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void bk_lorand_kernel(const K *__restrict__ in,
                                                          BkSeg *__restrict__ segs,
                                                          const RunState *st) {
  constexpr int T = level_tile<K>();
  __shared__ unsigned long long s_r[4][4];
  const uint32_t tile = blockIdx.x;
  if (tile >= st->bk_tiles) return;
  const uint32_t si = seg_of_tile(segs, st->bk_segs, tile);
  const uint32_t start = segs[si].start, count = segs[si].count, tb = segs[si].tile_base;
  const uint32_t lo = start + (tile - tb) * (uint32_t)T;
  const uint32_t hi = min(lo + (uint32_t)T, start + count);
  uint64_t olo = 0, ohi = 0, alo = ~0ull, ahi = ~0ull;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) {
    const K x = in[i];
    olo |= klo(x); ohi |= khi(x);
    alo &= klo(x); ahi &= khi(x);
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    olo |= (uint64_t)__shfl_xor((unsigned long long)olo, d, 64);
    ohi |= (uint64_t)__shfl_xor((unsigned long long)ohi, d, 64);
    alo &= (uint64_t)__shfl_xor((unsigned long long)alo, d, 64);
    ahi &= (uint64_t)__shfl_xor((unsigned long long)ahi, d, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_r[w][0] = olo; s_r[w][1] = ohi; s_r[w][2] = alo; s_r[w][3] = ahi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int ww = 1; ww < 4; ++ww) {
      olo |= s_r[ww][0]; ohi |= s_r[ww][1]; alo &= s_r[ww][2]; ahi &= s_r[ww][3];
    }
    BkSeg *sg = &segs[si];
    atomicOr(&sg->or_lo, (unsigned long long)olo);
    atomicAnd(&sg->and_lo, (unsigned long long)alo);
    if (sizeof(K) > 8) {
      atomicOr(&sg->or_hi, (unsigned long long)ohi);
      atomicAnd(&sg->and_hi, (unsigned long long)ahi);
    }
  }
}

// Histogram of every tile's digits into its segment's row of 256 counters.
template <typename K>
__global__ __launch_bounds__(kBlock) void bk_lhist_kernel(const K *__restrict__ in,
                                                         const BkSeg *__restrict__ segs,
                                                         uint32_t *__restrict__ hrows,
                                                         const RunState *st) {
  constexpr int T = level_tile<K>();
  __shared__ uint32_t s_h[256];
  const uint32_t tile = blockIdx.x;
  if (tile >= st->bk_tiles) return;
  const uint32_t si = seg_of_tile(segs, st->bk_segs, tile);
  const BkSeg sg = segs[si];
  uint32_t shift, nbits;
  seg_digit(sg, shift, nbits);
  if (nbits == 0) return;   // all equal: copied
  const uint32_t lo = sg.start + (tile - sg.tile_base) * (uint32_t)T;
  const uint32_t hi = min(lo + (uint32_t)T, sg.start + sg.count);
  const uint32_t dmask = (1u << nbits) - 1u;
  s_h[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t b = lo; b < hi; b += kBlock) {   // uniform trip count: matching needs the wave
    const uint32_t i = b + threadIdx.x;
    const uint32_t d = i < hi ? kbits(in[i], (int)shift, dmask) : 0u;
    const uint64_t pm = wave_peers<8>(d, i < hi);
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(pm >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)pm, 0u));
    if (i < hi && below == 0) atomicAdd(&s_h[d], (uint32_t)__popcll(pm));
  }
  __syncthreads();
  const uint32_t c = s_h[threadIdx.x];
  if (c) atomicAdd(&hrows[si * 256u + threadIdx.x], c);
}

__global__ __launch_bounds__(kBlock) void bk_lscan_kernel(
    BkSeg *__restrict__ segs, const uint32_t *__restrict__ hrows, uint32_t *__restrict__ cursor,
    BkSeg *__restrict__ nsegs, uint32_t *__restrict__ nhrows, uint32_t max_segs, uint32_t cap,
    int wide, int parity, BkItem *__restrict__ items, uint32_t max_items, RunState *st) {
  __shared__ uint32_t s_w4[4], s_c[256], s_x[256], s_nid[256], s_nnew;
  const uint32_t si = blockIdx.x;
  if (si >= st->bk_segs) return;
  const int t = threadIdx.x;
  BkSeg sg = segs[si];
  uint32_t shift, nbits;
  seg_digit(sg, shift, nbits);
  if (nbits == 0) {
    // all equal: copy items of at most CAP keys, in the buffer holding them
    if (t == 0) {
      segs[si].shift = 0;
      segs[si].nbits = 0;
      for (uint32_t c = 0; c < sg.count; c += cap) {
        const uint32_t slot = atomicAdd(&st->bk_items, 1u);
        if (slot < max_items) items[slot] = BkItem{sg.start + c, min(cap, sg.count - c), (uint32_t)(parity ^ 1), 0u};
        else atomicOr(&st->status, (uint32_t)ST_TOO_LARGE);
      }
    }
    return;
  }
  if (t == 0) {
    segs[si].shift = shift;   // for the scatter
    segs[si].nbits = nbits;
  }
  const uint32_t c = hrows[si * 256u + t];
  uint32_t total;
  const uint32_t ex = block_excl_sum<uint32_t>(c, s_w4, &total);
  cursor[si * 256u + t] = sg.start + ex;
  s_c[t] = c;
  s_x[t] = ex;
  uint32_t nid = ~0u;
  if (c > cap) {
    nid = atomicAdd(&st->bk_next_segs, 1u);
    if (nid < max_segs) {
      init_seg(&nsegs[nid], sg.start + ex, c, wide != 0);
      atomicAdd(&st->bk_next_keys, c);
    } else {
      atomicOr(&st->status, (uint32_t)ST_TOO_LARGE);
      nid = ~0u;
    }
  }
  // the new segments' histogram rows start at zero (listed in LDS: few)
  if (t == 0) s_nnew = 0;
  __syncthreads();
  if (nid != ~0u) s_nid[atomicAdd(&s_nnew, 1u)] = nid;
  __syncthreads();
  for (uint32_t k = 0; k < s_nnew; ++k) nhrows[s_nid[k] * 256u + t] = 0u;
  // items hold the keys the scatter writes to the other buffer; their keys
  // share the bits from shift + nbits up.  Thread 0 merges the digits into
  // items in LDS (at most one per digit); one atomic claims their slots.
  __shared__ uint32_t s_is[256], s_in[256], s_ni, s_base;
  if (t == 0) {
    uint32_t ni = 0, cur_s = 0, cur_n = 0;
    for (int d = 0; d < 256; ++d) {
      const uint32_t cd = s_c[d];
      if (cd == 0) continue;
      if (cd > cap || cur_n + cd > cap) {
        if (cur_n) { s_is[ni] = cur_s; s_in[ni] = cur_n; ++ni; }
        cur_n = 0;
        if (cd > cap) continue;
      }
      if (cur_n == 0) cur_s = s_x[d];
      cur_n += cd;
    }
    if (cur_n) { s_is[ni] = cur_s; s_in[ni] = cur_n; ++ni; }
    s_ni = ni;
    s_base = ni ? atomicAdd(&st->bk_items, ni) : 0u;
  }
  __syncthreads();
  if ((uint32_t)t < s_ni) {
    const uint32_t slot = s_base + t;
    if (slot < max_items)
      items[slot] = BkItem{sg.start + s_is[t], s_in[t], (uint32_t)parity, shift + nbits};
    else
      atomicOr(&st->status, (uint32_t)ST_TOO_LARGE);
  }
}

// The level's scatter: tiles within segments by their digit, claims on
// the segment's cursors.
template <typename K, bool LAB>
__global__ __launch_bounds__(kBlock) void bk_lscatter_kernel(TileIO<K, LAB> io, uint32_t n,
                                                            const BkSeg *__restrict__ segs,
                                                            uint32_t *__restrict__ cursor,
                                                            const RunState *st) {
  __shared__ StageMem<K, LAB> sm;
  constexpr int T = level_tile<K>();
  const uint32_t tile = blockIdx.x;
  if (tile >= st->bk_tiles) return;
  const uint32_t si = seg_of_tile(segs, st->bk_segs, tile);
  const BkSeg sg = segs[si];
  if (sg.nbits == 0) return;   // all equal: copied from where it is
  const uint32_t lo = sg.start + (tile - sg.tile_base) * (uint32_t)T;
  const uint32_t hi = min(lo + (uint32_t)T, sg.start + sg.count);
  scatter_tile<K, LAB, Bk<K>::TIPT, true>(io, n, lo, hi, (int)sg.shift, 1 << sg.nbits,
                                          cursor + si * 256u, 1, sm.keys, sm.vals);
}

__global__ void bk_next_level_kernel(RunState *st, uint32_t max_segs) {
  // the next level's segments become the current ones
  st->bk_segs = min(st->bk_next_segs, max_segs);
  st->bk_keys = st->bk_next_keys;
  st->bk_next_segs = 0;
  st->bk_next_keys = 0;
  st->bk_tiles = 0;
}

}  // namespace

// ---- host side ------------------------------------------------------------------------
// Waits for ev by polling: the device is still busy with the work queued
// after it, and a sleeping wait wakes up too late to keep it fed.
static bool spin_wait(hipEvent_t ev) {
  for (;;) {
    const hipError_t e = hipEventQuery(ev);
    if (e == hipSuccess) return true;
    if (e != hipErrorNotReady) return false;
  }
}

template <typename K>
int bucket_cap() { return Bk<K>::CAP; }

template <typename K>
int bucket_bits_for(int64_t n) {
  const double cap = 0.9 * Bk<K>::CAP;
  int b = 0;
  while (b < kBkMaxBits && (double)n > cap * (double)(1 << b)) ++b;
  return b;
}

template <typename K>
size_t bucket_max_segs(int64_t n) {
  return (size_t)(n / Bk<K>::CAP) + 2;
}

template <typename K>
size_t bucket_max_items(int64_t n) {
  // per level at most 3 count / CAP + 1 items per segment (consecutive
  // items hold more than CAP keys together unless a big digit separates
  // them), plus the copies of equal segments; levels <= key bits / 8
  const size_t per = 4 * (size_t)(n / Bk<K>::CAP) + 4;
  return per * (size_t)(sizeof(K) + 1) + 2 * (size_t)(n / Bk<K>::CAP) + 16;
}

template <typename K>
size_t bucket_ws_bytes(int64_t n) {
  const size_t S = bucket_max_segs<K>(n);
  size_t b = 0;
  b += align256_dev(((size_t)kBkMaxBuckets + 1) * 4);   // start
  b += align256_dev((size_t)kBkCursorWords * 4);        // cursors
  b += 2 * align256_dev(S * sizeof(BkSeg));             // segment lists (ping-pong)
  b += 2 * align256_dev(S * 256 * 4);                   // histogram rows (ping-pong)
  b += align256_dev(S * 256 * 4);                       // level cursors
  b += align256_dev(bucket_max_items<K>(n) * sizeof(BkItem));
  return b;
}

template <typename K>
void bucket_carve(void *ws, int64_t n, BucketMem &m) {
  const size_t S = bucket_max_segs<K>(n);
  char *p = static_cast<char *>(ws);
  auto take = [&](size_t bytes) { char *q = p; p += align256_dev(bytes); return q; };
  m.start = reinterpret_cast<uint32_t *>(take(((size_t)kBkMaxBuckets + 1) * 4));
  m.cursor = reinterpret_cast<uint32_t *>(take((size_t)kBkCursorWords * 4));
  m.segs[0] = reinterpret_cast<BkSeg *>(take(S * sizeof(BkSeg)));
  m.segs[1] = reinterpret_cast<BkSeg *>(take(S * sizeof(BkSeg)));
  m.hrows[0] = reinterpret_cast<uint32_t *>(take(S * 256 * 4));
  m.hrows[1] = reinterpret_cast<uint32_t *>(take(S * 256 * 4));
  m.lcursor = reinterpret_cast<uint32_t *>(take(S * 256 * 4));
  m.items = reinterpret_cast<BkItem *>(take(bucket_max_items<K>(n) * sizeof(BkItem)));
  m.max_segs = (uint32_t)S;
  m.max_items = (uint32_t)bucket_max_items<K>(n);
}

template <typename K>
int launch_bucket_sort(const BucketArgs<K> &a, RunState *st, RunState *st_host,
                       hipEvent_t ev, hipStream_t s) {
  using C = Bk<K>;
  const int64_t n = a.n;
  const int b = a.b;
  const int nb = 1 << b;
  const bool lab = a.labels != nullptr;
  const BucketMem &m = a.mem;
  // level 0: plan, scatters, LDS sort of the buckets that fit
  int nb1 = nb, nb2 = 1, b1 = b;
  const int splits = b == 0 ? 0 : (b <= 7 ? 1 : 2);
  if (splits == 2) {
    b1 = b / 2;
    nb1 = 1 << b1;
    nb2 = 1 << (b - b1);
  }
  constexpr int TILE = kBlock * C::TIPT;
  constexpr int wide = sizeof(K) > 8 ? 1 : 0;
  bk_plan_kernel<<<1, kPlanBlock, 0, s>>>(a.hist, (uint32_t)n, nb, nb2, TILE, (uint32_t)C::CAP,
                                          wide, sizeof(K) <= 8 ? 1 : 0, m.start, m.cursor,
                                          m.segs[0], m.hrows[0], st, st_host);
  if (hipGetLastError() != hipSuccess) return 1;
  // the host learns the level-1 segment count (written by the plan into its
  // pinned mirror) while the scatters run
  if (hipEventRecord(ev, s) != hipSuccess) return 1;
  K *kA = a.kA, *kB = a.kB;
  uint32_t *vA = a.vA, *vB = a.vB;
  const unsigned tiles1 = (unsigned)((n + TILE - 1) / TILE);
  int p0 = 0;   // buffer holding the level-0 buckets
  const K *k0 = kA;
  const uint32_t *v0 = lab ? a.labels : nullptr;
  if (splits >= 1) {
    TileIO<K, true> io1{kA, a.labels, kB, vB};
    TileIO<K, false> io1n{kA, nullptr, kB, nullptr};
    const int cstride = splits == 2 ? kBkCStride : 1;
    uint32_t *cur1 = splits == 2 ? m.cursor + kBkCursor1 : m.cursor;
    if (lab)
      bk_split1_kernel<K, true><<<tiles1, kBlock, 0, s>>>(io1, (uint32_t)n, a.mbits - b1, nb1, cur1,
                                                          cstride, st);
    else
      bk_split1_kernel<K, false><<<tiles1, kBlock, 0, s>>>(io1n, (uint32_t)n, a.mbits - b1, nb1,
                                                           cur1, cstride, st);
    p0 = 1;
    k0 = kB;
    v0 = vB;
  }
  if (splits == 2) {
    TileIO<K, true> io2{kB, vB, kA, vA};
    TileIO<K, false> io2n{kB, nullptr, kA, nullptr};
    const unsigned grid2 = tiles1 + (unsigned)nb1;
    if (lab)
      bk_split2_kernel<K, true><<<grid2, kBlock, 0, s>>>(io2, (uint32_t)n, a.mbits - b, nb2, nb1,
                                                         m.start, m.cursor, st);
    else
      bk_split2_kernel<K, false><<<grid2, kBlock, 0, s>>>(io2n, (uint32_t)n, a.mbits - b, nb2, nb1,
                                                          m.start, m.cursor, st);
    p0 = 0;
    k0 = kA;
    v0 = vA;
  }
  SortArgs<K, true> sa{{kA, kB}, {vA, vB}, k0, v0, kA, vA, m.start, m.items, (uint32_t)a.prepend};
  SortArgs<K, false> san{{kA, kB}, {nullptr, nullptr}, k0, nullptr, kA, nullptr, m.start, m.items,
                         (uint32_t)a.prepend};
  if (lab)
    bk_sort_kernel<K, true, C::BS, C::IPT, sort_wpe<K, true, false>(), false>
        <<<(unsigned)nb, C::BS, 0, s>>>(sa, 0, a.mbits - b, st);
  else
    bk_sort_kernel<K, false, C::BS, C::IPT, sort_wpe<K, false, false>(), false>
        <<<(unsigned)nb, C::BS, 0, s>>>(san, 0, a.mbits - b, st);
  if (hipGetLastError() != hipSuccess) return 1;
  if (!spin_wait(ev)) return 1;
  if (st_host->bk_lsd) return 3;   // the LSD passes sort instead
  uint32_t nseg = st_host->bk_segs;
  uint64_t nkeys = st_host->bk_keys;
  int levels = 0;
  int parity = p0;
  int cur = 0;
  const int max_levels = (int)sizeof(K) + 1;   // each level consumes 8 varying bits
  bool items = false;
  while (nseg > 0 && levels < max_levels) {
    const unsigned grid = (unsigned)(nkeys / (uint64_t)TILE + nseg + 1);
    const K *kin = parity ? kB : kA;
    bk_lprep_kernel<<<1, kPlanBlock, 0, s>>>(m.segs[cur], TILE, st);
    bk_lorand_kernel<K><<<grid, kBlock, 0, s>>>(kin, m.segs[cur], st);
    bk_lhist_kernel<K><<<grid, kBlock, 0, s>>>(kin, m.segs[cur], m.hrows[cur], st);
    bk_lscan_kernel<<<nseg, kBlock, 0, s>>>(m.segs[cur], m.hrows[cur], m.lcursor, m.segs[cur ^ 1],
                                            m.hrows[cur ^ 1], m.max_segs, (uint32_t)C::CAP, wide,
                                            parity ^ 1, m.items, m.max_items, st);
    if (hipGetLastError() != hipSuccess) return 1;
    if (hipMemcpyAsync(st_host, st, sizeof(RunState), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipEventRecord(ev, s) != hipSuccess)
      return 1;
    if (lab) {
      TileIO<K, true> io{parity ? kB : kA, parity ? vB : vA, parity ? kA : kB, parity ? vA : vB};
      bk_lscatter_kernel<K, true><<<grid, kBlock, 0, s>>>(io, (uint32_t)n, m.segs[cur], m.lcursor,
                                                          st);
    } else {
      TileIO<K, false> io{parity ? kB : kA, nullptr, parity ? kA : kB, nullptr};
      bk_lscatter_kernel<K, false><<<grid, kBlock, 0, s>>>(io, (uint32_t)n, m.segs[cur], m.lcursor,
                                                           st);
    }
    bk_next_level_kernel<<<1, 1, 0, s>>>(st, m.max_segs);
    if (hipGetLastError() != hipSuccess) return 1;
    if (!spin_wait(ev)) return 1;
    nseg = min(st_host->bk_next_segs, m.max_segs);
    nkeys = st_host->bk_next_keys;
    items = true;
    parity ^= 1;
    cur ^= 1;
    ++levels;
  }
  if (nseg > 0) {
    // a segment left after the last possible level cannot exist (each
    // level consumes 8 varying bits); report it rather than leave it unsorted
    return 2;
  }
  if (items) {
    const uint32_t ni = min(st_host->bk_items, m.max_items);
    if (ni) {
      if (lab)
        bk_sort_kernel<K, true, C::BS, C::IPT, sort_wpe<K, true, true>(), true>
            <<<ni, C::BS, 0, s>>>(sa, 1, 0, st);
      else
        bk_sort_kernel<K, false, C::BS, C::IPT, sort_wpe<K, false, true>(), true>
            <<<ni, C::BS, 0, s>>>(san, 1, 0, st);
      if (hipGetLastError() != hipSuccess) return 1;
    }
  }
  return 0;
}

#define PGVD_BK_INST(K)                                                                   \
  template int bucket_cap<K>();                                                           \
  template int bucket_bits_for<K>(int64_t);                                               \
  template size_t bucket_ws_bytes<K>(int64_t);                                            \
  template void bucket_carve<K>(void *, int64_t, BucketMem &);                            \
  template int launch_bucket_sort<K>(const BucketArgs<K> &, RunState *, RunState *,       \
                                     hipEvent_t, hipStream_t);
PGVD_BK_INST(uint32_t)
PGVD_BK_INST(uint64_t)
PGVD_BK_INST(Key128)
#undef PGVD_BK_INST

}  // namespace pgvd
