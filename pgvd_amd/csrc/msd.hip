// msd.hip -- bucket sort for 64-bit keys of large single builds: one scatter
// by the top b bits (b = 10..12) into 2^b buckets, then every bucket sorted
// on its remaining low bits inside one workgroup's LDS.
//
// The output is the sorted key sequence, the same one the reference's LSD
// split sort (RadixSortBigUnsigned, opencl/Kernels.cpp:219-247) produces:
// keys without payload, so equal keys are indistinguishable and the order of
// the scatter within a bucket does not matter.  Every kernel here does
// nothing unless st->msd_ok, which the plan kernel sets when every bucket
// holds at most kMsdCap keys; otherwise the pipeline's LSD passes, launched
// beside these and doing nothing when st->msd_ok, sort instead.  For
// uniform-ish keys a bucket is ~4K keys: the LSD passes over HBM become
// seven passes inside LDS.
#include "kernels.hpp"
#include "tilescan.hpp"

namespace pgvd {

namespace {

constexpr int kMsdIpt = 24;                   // keys per thread of a bucket sort
constexpr int kMsdScatterIpt = 16;            // keys per thread of the scatter
constexpr int kMsdScatterTile = kBlock * kMsdScatterIpt;
static_assert(kMsdCap == kBlock * kMsdIpt, "bucket capacity is one workgroup's keys");

// Exclusive scan of the bucket histogram into start[0..nb] and the scatter
// cursors; msd_ok = 1 when every bucket fits one workgroup.  One workgroup;
// nb is a power of two in [256, 4096].
__global__ __launch_bounds__(kBlock) void msd_plan_kernel(const uint32_t *__restrict__ hist,
                                                          int nb, uint32_t *__restrict__ start,
                                                          uint32_t *__restrict__ cursor,
                                                          RunState *st) {
  __shared__ uint32_t s_w[4], s_mx[4];
  const int t = threadIdx.x;
  const int per = nb / kBlock;                  // 1..16 consecutive buckets per thread
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
  __syncthreads();
  if (t == 0) {
    const uint32_t m4 = max(max(s_mx[0], s_mx[1]), max(s_mx[2], s_mx[3]));
    st->msd_ok = m4 <= (uint32_t)kMsdCap ? 1u : 0u;
  }
}

// Scatter of 4096 keys per workgroup to their buckets: a key's slot within
// its tile and bucket comes from an LDS counter, the tile's run in each
// bucket from one atomic on that bucket's cursor.
__global__ __launch_bounds__(kBlock) void msd_scatter_kernel(const uint64_t *__restrict__ in,
                                                             uint64_t *__restrict__ out,
                                                             uint32_t n, int shift, int nb,
                                                             uint32_t *__restrict__ cursor,
                                                             const RunState *st) {
  if (!st->msd_ok) return;
  __shared__ uint32_t s_cnt[kMsdMaxBuckets], s_base[kMsdMaxBuckets];
  const int t = threadIdx.x;
  for (int i = t; i < nb; i += kBlock) s_cnt[i] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * (uint32_t)kMsdScatterTile;
  const uint32_t dmask = (uint32_t)nb - 1;
  uint64_t k[kMsdScatterIpt];
  uint32_t dr[kMsdScatterIpt];
#pragma unroll
  for (int j = 0; j < kMsdScatterIpt; ++j) {
    const uint32_t i = base + j * kBlock + t;
    k[j] = in[i < n ? i : 0u];
  }
#pragma unroll
  for (int j = 0; j < kMsdScatterIpt; ++j) {
    const uint32_t i = base + j * kBlock + t;
    dr[j] = 0;
    if (i < n) {
      const uint32_t d = (uint32_t)(k[j] >> shift) & dmask;
      dr[j] = atomicAdd(&s_cnt[d], 1u) | (d << 16);
    }
  }
  __syncthreads();
  for (int d = t; d < nb; d += kBlock) {
    const uint32_t c = s_cnt[d];
    if (c) s_base[d] = atomicAdd(&cursor[d], c);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kMsdScatterIpt; ++j) {
    const uint32_t i = base + j * kBlock + t;
    if (i < n) {
      const uint32_t pos = s_base[dr[j] >> 16] + (dr[j] & 0xffffu);
      if (pos < n) out[pos] = k[j];   // always true when the histogram matches the keys
    }
  }
}

// Sorts bucket b (keys in[start[b] .. start[b+1])) on bits [0, 8 passes) in
// LDS and writes it to out at the same place.  The workgroup holds exactly
// kMsdCap keys: slots past the bucket hold ~0, which stays behind every real
// key (a larger or equal digit in every pass, and later in the stable
// order).  Key index le = 64 IPT w + 64 j + lane lives in wave w, register
// j; each pass ranks by the same ballot matching as the onesweep tile, so it
// is stable.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void msd_bucket_kernel(const uint64_t *__restrict__ in,
                                                            uint64_t *__restrict__ out,
                                                            const uint32_t *__restrict__ start,
                                                            int passes, const RunState *st) {
  if (!st->msd_ok) return;
  const uint32_t s0 = start[blockIdx.x];
  const uint32_t cnt = min(start[blockIdx.x + 1] - s0, (uint32_t)kMsdCap);
  if (cnt == 0) return;
  constexpr int IPT = kMsdIpt;
  __shared__ uint64_t s_keys[kMsdCap];
  __shared__ uint32_t s_cnt[4 * 256];
  __shared__ uint32_t s_dig[256];
  __shared__ uint32_t s_w[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint32_t wofs = (uint32_t)w * 64u * IPT;
  uint64_t k[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t le = wofs + j * 64 + lane;
    k[j] = le < cnt ? in[s0 + le] : ~0ull;
  }
  for (int p = 0; p < passes; ++p) {
    const int sh = 8 * p;
    for (int i = t; i < 4 * 256; i += kBlock) s_cnt[i] = 0;
    __syncthreads();
    uint32_t rk[IPT];   // rank within (wave, digit) | digit << 16
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t d = (uint32_t)(k[j] >> sh) & 0xffu;
      uint32_t plo = ~0u, phi = ~0u;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const int32_t sb = __builtin_amdgcn_sbfe((int32_t)d, b, 1);
        const uint64_t bb = __ballot(sb < 0);
        plo = __builtin_amdgcn_bitop3_b32(plo, (uint32_t)bb, (uint32_t)sb, 0x90);
        phi = __builtin_amdgcn_bitop3_b32(phi, (uint32_t)(bb >> 32), (uint32_t)sb, 0x90);
      }
      const uint32_t c = s_cnt[w * 256 + d];
      const uint32_t below = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
      rk[j] = (c + below) | (d << 16);
      if (below == 0) s_cnt[w * 256 + d] = c + (uint32_t)(__popc(plo) + __popc(phi));
    }
    __syncthreads();
    // thread t owns digit t: wave offsets, then the digit's start in the tile
    uint32_t tot = 0;
#pragma unroll
    for (int ww = 0; ww < 4; ++ww) {
      const uint32_t c = s_cnt[ww * 256 + t];
      s_cnt[ww * 256 + t] = tot;
      tot += c;
    }
    uint32_t all;
    s_dig[t] = block_excl_sum<uint32_t>(tot, s_w, &all);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t d = rk[j] >> 16;
      s_keys[(rk[j] & 0xffffu) + s_dig[d] + s_cnt[w * 256 + d]] = k[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j) k[j] = s_keys[wofs + j * 64 + lane];
    __syncthreads();   // s_keys and s_cnt are rewritten by the next pass
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t le = wofs + j * 64 + lane;
    if (le < cnt) out[s0 + le] = k[j];
  }
}

}  // namespace

int msd_bits_for(int64_t n) {
  if (n < ((int64_t)1 << 23)) return 0;
  for (int b = 10; b <= 12; ++b)
    if ((double)n <= 0.9 * kMsdCap * (double)(1 << b)) return b;
  return 0;
}

void launch_msd_sort(uint64_t *keys, uint64_t *alt, uint64_t *out, int64_t n, int mbits, int b,
                     const uint32_t *hist, uint32_t *start, uint32_t *cursor, RunState *st,
                     hipStream_t s) {
  const int nb = 1 << b;
  const int shift = mbits - b;
  msd_plan_kernel<<<1, kBlock, 0, s>>>(hist, nb, start, cursor, st);
  const unsigned tiles = (unsigned)((n + kMsdScatterTile - 1) / kMsdScatterTile);
  msd_scatter_kernel<<<tiles, kBlock, 0, s>>>(keys, alt, (uint32_t)n, shift, nb, cursor, st);
  msd_bucket_kernel<<<(unsigned)nb, kBlock, 0, s>>>(alt, out, start, (shift + 7) / 8, st);
}

}  // namespace pgvd
