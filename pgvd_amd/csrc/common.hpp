// common.hpp -- device-side helpers shared by the pgvd HIP kernels (gfx950).
//
// Keys are native words: uint32_t (mbits <= 32), uint64_t (<= 64) or
// Key128 (<= 128).  They replace the reference's 11-byte BigUnsigned
// (C/BigUnsigned.h:8-16); every operation the path needs on them (xor,
// bit length, shifts, 8-bit digits, equality) is a few integer ops.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgvd {

struct alignas(16) Key128 {
  uint64_t lo, hi;
};

// A 128-bit value stored with 8-byte alignment (inside BRT node records,
// so that the record layout matches the C-ABI's pgvd_brt_node128).
struct Key128s {
  uint64_t lo, hi;
  __host__ __device__ Key128s() = default;
  __host__ __device__ Key128s(Key128 k) : lo(k.lo), hi(k.hi) {}
  __host__ __device__ operator Key128() const { return {lo, hi}; }
};
template <typename K> struct StoredKey { using type = K; };
template <> struct StoredKey<Key128> { using type = Key128s; };

// ---- key arithmetic --------------------------------------------------------
__host__ __device__ inline bool keq(uint32_t a, uint32_t b) { return a == b; }
__host__ __device__ inline bool keq(uint64_t a, uint64_t b) { return a == b; }
__host__ __device__ inline bool keq(Key128 a, Key128 b) { return a.lo == b.lo && a.hi == b.hi; }

__host__ __device__ inline uint32_t kxor(uint32_t a, uint32_t b) { return a ^ b; }
__host__ __device__ inline uint64_t kxor(uint64_t a, uint64_t b) { return a ^ b; }
__host__ __device__ inline Key128 kxor(Key128 a, Key128 b) { return {a.lo ^ b.lo, a.hi ^ b.hi}; }

__host__ __device__ inline int bitlen(uint32_t x) { return x ? 32 - __builtin_clz(x) : 0; }
__host__ __device__ inline int bitlen(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }
__host__ __device__ inline int bitlen(Key128 x) { return x.hi ? 128 - __builtin_clzll(x.hi) : bitlen(x.lo); }

// Longest common prefix length of two mbits-wide keys.  The reference finds
// it by a binary search over BigUnsigned right shifts (C/BuildBRT.c:54-88);
// that search returns mbits - bitlen(a ^ b) for every pair (equal keys give
// mbits), which is what this computes with one count-leading-zeros.
template <typename K>
__host__ __device__ inline int delta(K a, K b, int mbits) { return mbits - bitlen(kxor(a, b)); }

// Logical right shift; shifts at or past the word width give 0, as a
// BigUnsigned shift does (C/BigUnsigned.c:417-471).
__host__ __device__ inline uint32_t kshr(uint32_t a, int s) { return s >= 32 ? 0u : (a >> s); }
__host__ __device__ inline uint64_t kshr(uint64_t a, int s) { return s >= 64 ? 0ull : (a >> s); }
__host__ __device__ inline Key128 kshr(Key128 a, int s) {
  if (s >= 128) return {0, 0};
  if (s >= 64) return {a.hi >> (s - 64), 0};
  if (s == 0) return a;
  return {(a.lo >> s) | (a.hi << (64 - s)), a.hi >> s};
}

__host__ __device__ inline uint32_t klow32(uint32_t a) { return a; }
__host__ __device__ inline uint32_t klow32(uint64_t a) { return (uint32_t)a; }
__host__ __device__ inline uint32_t klow32(Key128 a) { return (uint32_t)a.lo; }

// 8-bit digit starting at bit `shift` (a multiple of 8, so it never straddles
// a 64-bit word).
__device__ inline uint32_t kdigit(uint32_t a, int shift) { return (a >> shift) & 0xffu; }
__device__ inline uint32_t kdigit(uint64_t a, int shift) { return (uint32_t)(a >> shift) & 0xffu; }
__device__ inline uint32_t kdigit(Key128 a, int shift) {
  return shift >= 64 ? (uint32_t)(a.hi >> (shift - 64)) & 0xffu : (uint32_t)(a.lo >> shift) & 0xffu;
}

// `mask`-wide digit (mask = 2^w - 1, w <= 8) starting at bit `shift`, which
// may straddle the two words of a 128-bit key.
__device__ inline uint32_t kdigit_w(uint32_t a, int shift, uint32_t mask) {
  return (a >> shift) & mask;
}
__device__ inline uint32_t kdigit_w(uint64_t a, int shift, uint32_t mask) {
  return (uint32_t)(a >> shift) & mask;
}
__device__ inline uint32_t kdigit_w(Key128 a, int shift, uint32_t mask) {
  if (shift >= 64) return (uint32_t)(a.hi >> (shift - 64)) & mask;
  const uint64_t w = shift ? (a.lo >> shift) | (a.hi << (64 - shift)) : a.lo;
  return (uint32_t)w & mask;
}

// Key with bit b (0 <= b < width) flipped.
__host__ __device__ inline uint32_t kflip(uint32_t a, int b) { return a ^ (1u << b); }
__host__ __device__ inline uint64_t kflip(uint64_t a, int b) { return a ^ (1ull << b); }
__host__ __device__ inline Key128 kflip(Key128 a, int b) {
  return b >= 64 ? Key128{a.lo, a.hi ^ (1ull << (b - 64))} : Key128{a.lo ^ (1ull << b), a.hi};
}

__host__ __device__ inline bool kzero(uint32_t a) { return a == 0; }
__host__ __device__ inline bool kzero(uint64_t a) { return a == 0; }
__host__ __device__ inline bool kzero(Key128 a) { return (a.lo | a.hi) == 0; }

template <typename K> __host__ __device__ inline K kzero_value();
template <> __host__ __device__ inline uint32_t kzero_value<uint32_t>() { return 0u; }
template <> __host__ __device__ inline uint64_t kzero_value<uint64_t>() { return 0ull; }
template <> __host__ __device__ inline Key128 kzero_value<Key128>() { return {0, 0}; }

// Truncate a 128-bit value to the key type.
template <typename K> __device__ inline K from128(uint64_t lo, uint64_t hi);
template <> __device__ inline uint32_t from128<uint32_t>(uint64_t lo, uint64_t) { return (uint32_t)lo; }
template <> __device__ inline uint64_t from128<uint64_t>(uint64_t lo, uint64_t) { return lo; }
template <> __device__ inline Key128 from128<Key128>(uint64_t lo, uint64_t hi) { return {lo, hi}; }

// ---- Morton bit spreading --------------------------------------------------
// Bit i of x goes to bit 2i (2D) or 3i (3D).  Standard magic-mask spreads.
__device__ inline uint64_t spread2(uint32_t v) {
  uint64_t x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}
__device__ inline uint64_t spread3(uint32_t v) {   // low 21 bits
  uint64_t x = v & 0x1FFFFFu;
  x = (x | (x << 32)) & 0x001F00000000FFFFull;
  x = (x | (x << 16)) & 0x001F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

// OR a 64-bit value shifted left by s (0 <= s < 128) into (lo, hi).
__device__ inline void or_shifted(uint64_t v, int s, uint64_t &lo, uint64_t &hi) {
  if (s == 0) { lo |= v; return; }
  if (s < 64) { lo |= v << s; hi |= v >> (64 - s); return; }
  hi |= v << (s - 64);
}

// Morton key of DIM coordinates: key bit i*DIM+k = bit i of coordinate k for
// i < bits (C/z_order.h:14-47).  Bits of a coordinate at or above `bits` are
// ignored, as xyz2z ignores them.
// WIDE = false: the key has at most 64 bits (bits <= 32 in 2D, <= 21 in 3D),
// so the coordinates' upper parts are zero after the mask and are skipped.
template <int DIM, bool WIDE = true>
__device__ inline void morton128(const int64_t *c, int bits, uint64_t &lo, uint64_t &hi) {
  lo = 0; hi = 0;
  const uint64_t cmask = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    const uint64_t v = (uint64_t)c[k] & cmask;
    if (DIM == 2) {
      or_shifted(spread2((uint32_t)v), k, lo, hi);
      if (WIDE) or_shifted(spread2((uint32_t)(v >> 32)), 64 + k, lo, hi);
    } else {
      or_shifted(spread3((uint32_t)(v & 0x1FFFFF)), k, lo, hi);
      if (WIDE) or_shifted(spread3((uint32_t)((v >> 21) & 0x1FFFFF)), 63 + k, lo, hi);
      // bits 42..63 of a coordinate would land at key bits >= 126 + k; the
      // widest supported key is 3*42 = 126 bits, so they are always masked.
    }
  }
}

// ---- scoped atomics (device scope; coherent across XCDs) --------------------
__device__ inline uint32_t ld_acq_u32(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_rel_u32(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t ld_acq_u64(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_rel_u64(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A flag that any number of workgroups may raise during a kernel and that
// only later kernels (or the host) read: a store kept in the XCD's L2 and
// written back when the kernel ends.  Every writer stores the same value, so
// the order of the write-backs does not matter.  (A device-scope store
// writes through the L2; thousands of workgroups doing that on one address
// serialise at the memory and stall every CU's memory pipeline behind it.)
__device__ inline void raise_flag(uint32_t *p) {
  __hip_atomic_store(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ---- wave64 helpers ---------------------------------------------------------
__device__ inline uint64_t lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <typename T>
__device__ inline T wave_incl_scan_add(T v) {
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// 32-bit inclusive scan over the wave in DPP moves: shifts by 1, 2, 4 and 8
// within each row of 16 lanes (a lane with no source adds 0), then lane 15
// of each row added to the next row (rows 1 and 3) and lane 31 to rows 2
// and 3 -- six adds, no LDS.
__device__ inline uint32_t wave_incl_scan_add(uint32_t v) {
  int x = (int)v;
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);   // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);   // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);   // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);   // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);   // row_bcast:15
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);   // row_bcast:31
  return (uint32_t)x;
}

template <typename T>
__device__ inline T wave_reduce_add(T v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// 32-bit sum over the wave: the DPP scan's last lane, read into a scalar.
// Call with the whole wave converged.
__device__ inline uint32_t wave_reduce_add(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan_add(v), 63);
}

// Appends `idx` to list[] for every lane with `take` set, one atomic per wave
// on *count.  Call with the whole wave converged.
template <typename T>
__device__ inline void wave_append(bool take, T idx, uint32_t *count, T *list) {
  const uint64_t mb = __ballot(take);
  if (!mb) return;
  const int leader = __ffsll((unsigned long long)mb) - 1;
  uint32_t base = 0;
  if ((int)__lane_id() == leader) base = atomicAdd(count, (uint32_t)__popcll(mb));
  base = __shfl(base, leader, 64);
  if (take) list[base + (uint32_t)__popcll(mb & lanemask_lt())] = idx;
}

// Miss masks: a converged wave records which of its 64 lanes take a node as
// one word, words[word] = ballot(take), stored by lane 0.  No atomics, so
// the word must belong to this wave alone (word = first index >> 6 for 64
// consecutive, 64-aligned indices).
__device__ inline void wave_mask(bool take, uint64_t *words, uint32_t word) {
  const uint64_t b = __ballot(take);
  if (__lane_id() == 0) words[word] = b;
}

// Instance owning key slot `slot`: the largest s with kofs[s] <= slot
// (kofs strictly increasing, kofs[B] = total slots).
__device__ inline uint32_t find_seg(const uint32_t *kofs, uint32_t B, uint32_t slot) {
  uint32_t lo = 0, hi = B;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (kofs[mid] <= slot) lo = mid; else hi = mid;
  }
  return lo;
}

// Workgroup b of a grid of nb runs on XCD b % 8 (round-robin dispatch).  This
// maps b to a segment index so that each XCD walks one contiguous run of
// segments: neighbouring segments, whose halos overlap, then share an L2.
__device__ inline uint32_t xcd_segment(uint32_t b, uint32_t nb) {
  const uint32_t q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Octant of `key` at octree level lvl: the DIM key bits right below a prefix
// of lvl digits (0 if the level is below the key's last digit).
template <typename K, int DIM>
__device__ inline int digit_at(K key, int lvl, int mbits) {
  const int sh = mbits - (lvl + 1) * DIM;
  return sh < 0 ? 0 : (int)(klow32(kshr(key, sh)) & ((1u << DIM) - 1));
}

// Encodes a float so that unsigned order equals float order (non-NaN).
__device__ inline uint32_t fenc(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float fdec(uint32_t e) {
  return __uint_as_float((e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e);
}

}  // namespace pgvd
