// sort.hip -- onesweep LSD radix sort (8-bit digits) for 32/64/128-bit keys
// with an optional u32 payload.
//
// Replaces the reference's bit-serial split sort (RadixSortBigUnsigned,
// opencl/Kernels.cpp:219-247: BitPredicate + StreamScan + BUCompact per
// bit).  Both are stable LSD sorts, so the sorted key sequence -- the only
// thing the path consumes -- is identical.
//
// One kernel launch per digit.  Each workgroup claims a tile id from an
// atomic counter (so every tile it waits on was claimed by a workgroup that
// is already running), ranks its keys with wave64 ballot matching and
// per-wave LDS digit counters, publishes its per-digit counts, and resolves
// its global digit offsets with a per-digit decoupled look-back.  Keys are
// then staged in LDS in digit order and written out in coalesced runs.
#include "kernels.hpp"

namespace pgvd {

namespace {

constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagIncl = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1;
constexpr int kWaves = kBlock / 64;

// Keys per thread of a tile.  The digit-ordered staging buffer reuses the
// per-wave counters' space; payload tiles are kept near 30 KB of LDS.
#ifndef PGVD_SORT_IPT64
#define PGVD_SORT_IPT64 24   // keys per thread for 32/64-bit keys without payload
#endif
// Waves per SIMD the sort kernels are compiled for (register budget); LDS
// allows three 256-thread tiles of 24 64-bit keys per CU.
#ifndef PGVD_SORT_WAVES
#define PGVD_SORT_WAVES 3
#endif
template <typename K, bool PAY> struct SortCfg { static constexpr int IPT = PGVD_SORT_IPT64; };
template <> struct SortCfg<uint64_t, true> { static constexpr int IPT = 10; };
template <> struct SortCfg<Key128, false> { static constexpr int IPT = 7; };
template <> struct SortCfg<Key128, true> { static constexpr int IPT = 6; };

// Exclusive scan of one u32 per thread over the 256-thread block.
__device__ inline uint32_t block_excl_scan(uint32_t x, uint32_t *s_wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t incl = wave_incl_scan_add(x);
  if (lane == 63) s_wsum[w] = incl;
  __syncthreads();
  uint32_t off = 0;
  for (int j = 0; j < w; ++j) off += s_wsum[j];
  return off + incl - x;
}

// LDS of one onesweep tile.  The per-wave digit counters (ranking) and the
// digit-ordered staging of keys and payload (write-out) are never live
// together: one area.  s_dig holds the digit starts within the tile, then
// (once every key's slot is known) each digit's global base.
template <typename K, bool PAY, int IPT>
struct OnesweepLds {
  static constexpr int kTile = kBlock * IPT;
  static constexpr int kStageBytes = kTile * (int)sizeof(K) + (PAY ? kTile * 4 : 0);
  static constexpr int kCntBytes = kWaves * 256 * 4;
  __attribute__((aligned(16))) unsigned char raw[kStageBytes > kCntBytes ? kStageBytes : kCntBytes];
  uint32_t dig[256];
  uint32_t wsum_a[kWaves], wsum_b[kWaves];
};

// One instance per <K, PAY, IPT>, shared by the whole-tile and partial-tile
// forms of onesweep_tile.
template <typename K, bool PAY, int IPT>
__device__ inline OnesweepLds<K, PAY, IPT> &onesweep_lds() {
  __shared__ OnesweepLds<K, PAY, IPT> lds;
  return lds;
}

// One onesweep tile: the nvalid (<= IPT*256) keys at kin[base..) of a run
// of tiles [lo, ...) whose keys go to kout[seg0..) in digit order.  `tile` is
// the tile's index in the look-back array (tiles lo..tile-1 precede it in the
// same run); ghist is the run's digit histogram for this pass.  FULL: nvalid
// is the whole tile, so no key needs a range test.
// DBITS: the digit width (<= 8); the tile's 256 digit slots hold 2^DBITS
// digits, and slots above them stay empty.
template <typename K, bool PAY, int IPT, bool FULL, int DBITS = 8>
__device__ __attribute__((always_inline)) inline void onesweep_tile(const K *__restrict__ kin, K *__restrict__ kout,
                                     const uint32_t *__restrict__ vin, uint32_t *__restrict__ vout,
                                     uint32_t base, uint32_t nvalid, uint32_t tile, uint32_t lo,
                                     uint32_t seg0, int shift, const uint32_t *__restrict__ ghist,
                                     uint32_t *__restrict__ lookback) {
  constexpr int TILE = kBlock * IPT;
  constexpr uint32_t DMASK = (1u << DBITS) - 1u;
  OnesweepLds<K, PAY, IPT> &L = onesweep_lds<K, PAY, IPT>();
  uint32_t *s_cnt = reinterpret_cast<uint32_t *>(L.raw);
  K *s_keys = reinterpret_cast<K *>(L.raw);
  uint32_t *s_vals = reinterpret_cast<uint32_t *>(L.raw + TILE * sizeof(K));
  uint32_t *s_dig = L.dig;
  uint32_t *s_wsum_a = L.wsum_a, *s_wsum_b = L.wsum_b;

  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int i = t; i < kWaves * 256; i += kBlock) s_cnt[i] = 0;
  __syncthreads();
  const uint32_t wofs = (uint32_t)w * 64u * IPT;

  K k[IPT];
  uint32_t v[IPT];
  uint32_t rk[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t le = wofs + j * 64 + lane;
    k[j] = (FULL || le < nvalid) ? kin[base + le] : kzero_value<K>();
    if (PAY) v[j] = (FULL || le < nvalid) ? vin[base + le] : 0u;
  }

  // Rank: keys with equal digits within a wave found by 8 ballots; the
  // lowest lane of each group bumps the wave's LDS counter for that digit.
  // Items are processed in index order, so ranks are stable.  Per digit bit
  // b, s = -1 where the bit is set, else 0, and a lane keeps the lanes whose
  // ballot bit equals its own: peers &= ~(ballot ^ s), in two 32-bit halves,
  // one v_bitop3 each (truth table 0x90: operand 0 is the table index's
  // high bit, so the result is 1 for (1,0,0) and (1,1,1)).
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t le = wofs + j * 64 + lane;
    const bool valid = FULL || le < nvalid;
    const uint32_t d = DBITS == 8 ? kdigit(k[j], shift) : kdigit_w(k[j], shift, DMASK);
    uint32_t plo = ~0u, phi = ~0u;
    if (!FULL) {
      const uint64_t vb = __ballot(valid);
      plo = (uint32_t)vb;
      phi = (uint32_t)(vb >> 32);
    }
#pragma unroll
    for (int b = 0; b < DBITS; ++b) {
      const int32_t sb = __builtin_amdgcn_sbfe((int32_t)d, b, 1);
      const uint64_t bb = __ballot(sb < 0);
      plo = __builtin_amdgcn_bitop3_b32(plo, (uint32_t)bb, (uint32_t)sb, 0x90);
      phi = __builtin_amdgcn_bitop3_b32(phi, (uint32_t)(bb >> 32), (uint32_t)sb, 0x90);
    }
    if (valid) {
      const uint32_t cnt = s_cnt[w * 256 + d];
      // peers in lower lanes, counted by mbcnt
      const uint32_t below = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
      rk[j] = cnt + below;
      if (below == 0) s_cnt[w * 256 + d] = cnt + (uint32_t)(__popc(plo) + __popc(phi));
    }
  }
  __syncthreads();

  // Thread t owns digit t: per-wave exclusive offsets and the tile total.
  uint32_t total = 0;
#pragma unroll
  for (int ww = 0; ww < kWaves; ++ww) {
    const uint32_t c = s_cnt[ww * 256 + t];
    s_cnt[ww * 256 + t] = total;
    total += c;
  }
  uint32_t *lb = lookback + (size_t)tile * 256;
  st_rel_u32(lb + t, (tile == lo ? kFlagIncl : kFlagAgg) | total);

  const uint32_t dstart = block_excl_scan(total, s_wsum_a);
  s_dig[t] = dstart;
  const uint32_t gexcl = block_excl_scan(ghist[t], s_wsum_b);

  // Look-back for digit t: read up to kBatch predecessor tiles per round
  // trip, consume them nearest first until an inclusive prefix; a tile that
  // has not published yet is re-read on the next round.
  uint32_t excl = 0;
  if (tile > lo && (DBITS == 8 || (uint32_t)t <= DMASK)) {   // slots above the digits stay 0
    constexpr int kBatch = 8;
    int kk = (int)tile - 1;
    for (;;) {
      uint32_t wv[kBatch];
#pragma unroll
      for (int b = 0; b < kBatch; ++b)
        wv[b] = (kk - b >= (int)lo) ? ld_acq_u32(lookback + (size_t)(kk - b) * 256 + t) : kFlagIncl;
      int used = 0;
      bool done = false;
#pragma unroll
      for (int b = 0; b < kBatch; ++b) {
        if (done || used != b) continue;
        const uint32_t f = wv[b] & ~kValMask;
        if (f == 0) continue;            // not published: stop consuming here
        excl += wv[b] & kValMask;
        used = b + 1;
        done = (f == kFlagIncl);
      }
      if (done) break;
      kk -= used;
      if (used < kBatch) __builtin_amdgcn_s_sleep(1);
    }
    st_rel_u32(lb + t, kFlagIncl | (excl + total));
  }
  __syncthreads();

  // Each key's slot in the digit-ordered staging, taken from the counters
  // and digit starts before their space is reused.
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t le = wofs + j * 64 + lane;
    if (FULL || le < nvalid) {
      const uint32_t d = DBITS == 8 ? kdigit(k[j], shift) : kdigit_w(k[j], shift, DMASK);
      rk[j] += s_dig[d] + s_cnt[w * 256 + d];
    }
  }
  __syncthreads();
  s_dig[t] = seg0 + gexcl + excl - dstart;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t le = wofs + j * 64 + lane;
    if (FULL || le < nvalid) {
      s_keys[rk[j]] = k[j];
      if (PAY) s_vals[rk[j]] = v[j];
    }
  }
  __syncthreads();

#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t p = j * kBlock + t;
    if (FULL || p < nvalid) {
      const K key = s_keys[p];
      const uint32_t g =
          s_dig[DBITS == 8 ? kdigit(key, shift) : kdigit_w(key, shift, DMASK)] + p;
      kout[g] = key;
      if (PAY) vout[g] = s_vals[p];
    }
  }
}

// Tile ids come from an atomic counter, so every tile a workgroup looks
// back at was claimed by a workgroup that is already running.
__device__ inline uint32_t claim_tile(uint32_t *counter) {
  __shared__ uint32_t s_tile;
  if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
  __syncthreads();
  return s_tile;
}

template <typename K, bool PAY, int IPT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(PGVD_SORT_WAVES))) void onesweep_kernel(
    const K *__restrict__ kin, K *__restrict__ kout, const uint32_t *__restrict__ vin,
    uint32_t *__restrict__ vout, uint32_t n, int shift, const uint32_t *__restrict__ ghist,
    uint32_t *__restrict__ lookback, uint32_t *tile_counter, const RunState *skip) {
  if (skip && skip->msd_ok) return;   // uniform: the bucket sort has sorted the keys
  constexpr uint32_t TILE = kBlock * IPT;
  const uint32_t tile = claim_tile(tile_counter);
  const uint32_t base = tile * TILE;
  const uint32_t nvalid = (n - base) < TILE ? (n - base) : TILE;
  if (nvalid == TILE)
    onesweep_tile<K, PAY, IPT, true>(kin, kout, vin, vout, base, nvalid, tile, 0u, 0u, shift,
                                     ghist, lookback);
  else
    onesweep_tile<K, PAY, IPT, false>(kin, kout, vin, vout, base, nvalid, tile, 0u, 0u, shift,
                                      ghist, lookback);
}

// Segmented form: instance s owns keys [seg.kofs, +seg.n) and sort tiles
// [tt.first[s], tt.first[s+1]); each instance is sorted on its own, in one
// launch per digit for the whole batch.
template <typename K, int IPT, int DBITS, int WAVES>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES))) void onesweep_seg_kernel(
    const K *__restrict__ kin, K *__restrict__ kout, const SegInfo *__restrict__ seg,
    TileTable tt, int passes, int pass, const uint32_t *__restrict__ bhist,
    uint32_t *__restrict__ lookback, uint32_t *tile_counter) {
  constexpr uint32_t TILE = kBlock * IPT;
  const uint32_t tile = claim_tile(tile_counter);
  const uint32_t s = tt.seg[tile];
  const uint32_t lo = tt.first[s];
  const SegInfo si = seg[s];
  const uint32_t lofs = (tile - lo) * TILE;
  const uint32_t nvalid = (si.n - lofs) < TILE ? (si.n - lofs) : TILE;
  const uint32_t *gh = bhist + ((size_t)s * passes + pass) * 256;
  if (nvalid == TILE)
    onesweep_tile<K, false, IPT, true, DBITS>(kin, kout, nullptr, nullptr, si.kofs + lofs, nvalid,
                                              tile, lo, si.kofs, DBITS * pass, gh, lookback);
  else
    onesweep_tile<K, false, IPT, false, DBITS>(kin, kout, nullptr, nullptr, si.kofs + lofs, nvalid,
                                               tile, lo, si.kofs, DBITS * pass, gh, lookback);
}

}  // namespace

int onesweep_tile_size(int key_bytes, bool payload) {
  int ipt;
  if (key_bytes == 16) ipt = payload ? SortCfg<Key128, true>::IPT : SortCfg<Key128, false>::IPT;
  else if (key_bytes == 8) ipt = payload ? SortCfg<uint64_t, true>::IPT : SortCfg<uint64_t, false>::IPT;
  else ipt = payload ? SortCfg<uint32_t, true>::IPT : SortCfg<uint32_t, false>::IPT;
  return kBlock * ipt;
}

int64_t onesweep_tiles(int key_bytes, bool payload, int64_t n) {
  const int ts = onesweep_tile_size(key_bytes, payload);
  return (n + ts - 1) / ts;
}

template <typename K>
int launch_onesweep(K *keys, K *keys_alt, const uint32_t *vals_in, uint32_t *vals,
                    uint32_t *vals_alt, int64_t n, int passes, const uint32_t *hist,
                    uint32_t *lookback, RunState *st, hipStream_t s, hipEvent_t *pass_events,
                    const RunState *skip_if_msd) {
  if (n <= 1 || passes <= 0) return 0;
  const bool pay = vals_in != nullptr;
  const int64_t tiles = onesweep_tiles((int)sizeof(K), pay, n);

  K *src = keys, *dst = keys_alt;
  const uint32_t *vs = vals_in;
  uint32_t *vd = vals_alt, *vother = vals;
  for (int r = 0; r < passes; ++r) {
    uint32_t *lb = lookback + (size_t)r * (size_t)tiles * 256;
    uint32_t *ctr = &st->counters[r];
    if (pass_events) (void)hipEventRecord(pass_events[2 * r], s);
    if (pay) {
      constexpr int IPT = SortCfg<K, true>::IPT;
      onesweep_kernel<K, true, IPT><<<(unsigned)tiles, kBlock, 0, s>>>(
          src, dst, vs, vd, (uint32_t)n, 8 * r, hist + 256 * r, lb, ctr, skip_if_msd);
    } else {
      constexpr int IPT = SortCfg<K, false>::IPT;
      onesweep_kernel<K, false, IPT><<<(unsigned)tiles, kBlock, 0, s>>>(
          src, dst, nullptr, nullptr, (uint32_t)n, 8 * r, hist + 256 * r, lb, ctr, skip_if_msd);
    }
    if (pass_events) (void)hipEventRecord(pass_events[2 * r + 1], s);
    K *tk = src; src = dst; dst = tk;
    // payload: the buffer just written becomes the next source; the next
    // destination is the other work buffer.
    uint32_t *written = vd;
    vd = vother;
    vother = written;
    vs = written;
  }
  return passes & 1;
}

bool onesweep_seg_ipt_ok(int key_bytes, int ipt) {
  return key_bytes == 16 ? (ipt == 4 || ipt == 7 || ipt == 8)
                         : (ipt == 8 || ipt == 12 || ipt == 15 || ipt == 16);
}

bool onesweep_seg_digit_ok(int key_bytes, int digit_bits) {
  return key_bytes == 16 ? digit_bits == 8 : (digit_bits >= 6 && digit_bits <= 8);
}

bool onesweep_seg_waves_ok(int waves) { return waves >= 2 && waves <= 4; }

template <typename K, int IPT, int DBITS, int WAVES>
void seg_passes(K *&src, K *&dst, const BatchDev &b, int passes, const uint32_t *bhist,
                uint32_t *lookback, RunState *st, hipStream_t s) {
  for (int r = 0; r < passes; ++r) {
    uint32_t *lb = lookback + (size_t)r * (size_t)b.n_sort * 256;
    onesweep_seg_kernel<K, IPT, DBITS, WAVES><<<(unsigned)b.n_sort, kBlock, 0, s>>>(
        src, dst, b.seg, b.sort, passes, r, bhist, lb, &st->counters[r]);
    K *tk = src; src = dst; dst = tk;
  }
}

template <typename K, int IPT, int DBITS>
void seg_passes_w(K *&src, K *&dst, const BatchDev &b, int passes, const uint32_t *bhist,
                  uint32_t *lookback, RunState *st, hipStream_t s) {
  if (b.sort_waves == 2) seg_passes<K, IPT, DBITS, 2>(src, dst, b, passes, bhist, lookback, st, s);
  else if (b.sort_waves == 4) seg_passes<K, IPT, DBITS, 4>(src, dst, b, passes, bhist, lookback, st, s);
  else seg_passes<K, IPT, DBITS, 3>(src, dst, b, passes, bhist, lookback, st, s);
}

template <typename K, int IPT>
void seg_passes_d(K *&src, K *&dst, const BatchDev &b, int passes, const uint32_t *bhist,
                  uint32_t *lookback, RunState *st, hipStream_t s) {
  if constexpr (sizeof(K) == 16) {
    seg_passes_w<K, IPT, 8>(src, dst, b, passes, bhist, lookback, st, s);
  } else {
    if (b.digit_bits == 6) seg_passes_w<K, IPT, 6>(src, dst, b, passes, bhist, lookback, st, s);
    else if (b.digit_bits == 7) seg_passes_w<K, IPT, 7>(src, dst, b, passes, bhist, lookback, st, s);
    else seg_passes_w<K, IPT, 8>(src, dst, b, passes, bhist, lookback, st, s);
  }
}

template <typename K>
int launch_onesweep_seg(K *keys, K *keys_alt, const BatchDev &b, int passes,
                        const uint32_t *bhist, uint32_t *lookback, RunState *st, hipStream_t s) {
  if (passes <= 0 || b.n_sort <= 0) return 0;
  K *src = keys, *dst = keys_alt;
  if constexpr (sizeof(K) == 16) {
    if (b.sort_ipt == 4) seg_passes_d<K, 4>(src, dst, b, passes, bhist, lookback, st, s);
    else if (b.sort_ipt == 7) seg_passes_d<K, 7>(src, dst, b, passes, bhist, lookback, st, s);
    else seg_passes_d<K, 8>(src, dst, b, passes, bhist, lookback, st, s);
  } else {
    if (b.sort_ipt == 8) seg_passes_d<K, 8>(src, dst, b, passes, bhist, lookback, st, s);
    else if (b.sort_ipt == 12) seg_passes_d<K, 12>(src, dst, b, passes, bhist, lookback, st, s);
    else if (b.sort_ipt == 15) seg_passes_d<K, 15>(src, dst, b, passes, bhist, lookback, st, s);
    else seg_passes_d<K, 16>(src, dst, b, passes, bhist, lookback, st, s);
  }
  return passes & 1;
}

#define PGVD_INST(K)                                                                          \
  template int launch_onesweep<K>(K *, K *, const uint32_t *, uint32_t *, uint32_t *, int64_t,  \
                                  int, const uint32_t *, uint32_t *, RunState *, hipStream_t,   \
                                  hipEvent_t *, const RunState *);                            \
  template int launch_onesweep_seg<K>(K *, K *, const BatchDev &, int, const uint32_t *,       \
                                      uint32_t *, RunState *, hipStream_t);
PGVD_INST(uint32_t)
PGVD_INST(uint64_t)
PGVD_INST(Key128)
#undef PGVD_INST

}  // namespace pgvd
