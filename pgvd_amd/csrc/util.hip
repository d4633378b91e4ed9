// util.hip -- small utility kernels: word fill (used instead of
// hipMemsetAsync so that captured graphs replay correctly), a plain copy
// kernel (also the calibration kernel for HBM byte counters), setting the
// run sizes from a prefix-sum array, and a position-dependent digest of a
// device range (result comparison without a download).
#include "kernels.hpp"

namespace pgvd {

namespace {

__global__ __launch_bounds__(kBlock) void fill_u32_kernel(uint32_t *p, size_t n, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    p[i] = v;
}

__global__ __launch_bounds__(kBlock) void copy_kernel(uint4 *__restrict__ dst,
                                                      const uint4 *__restrict__ src, size_t n16,
                                                      uint8_t *__restrict__ dtail,
                                                      const uint8_t *__restrict__ stail,
                                                      size_t ntail) {
  // four 16-byte words per thread and round, all loads before the stores
  constexpr int U = 4;
  const size_t tid = (size_t)blockIdx.x * kBlock + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * kBlock * U;
  for (size_t base = (size_t)blockIdx.x * kBlock * U + threadIdx.x; base < n16; base += stride) {
    uint4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = base + (size_t)u * kBlock;
      if (i < n16) v[u] = src[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = base + (size_t)u * kBlock;
      if (i < n16) dst[i] = v[u];
    }
  }
  if (tid < ntail) dtail[tid] = stail[tid];
}

__global__ void set_sizes_kernel(RunState *st, const uint32_t *ps, int64_t m) {
  st->m = (uint32_t)m;
  st->octree_size = ps[m - 1];
}

// The outcome of a build as four words: a pgvd_status code (the device's
// error bits in the order the host reports them; an octree above the
// output capacity is PGVD_ERR_CAPACITY), m, the octree size, malformed.
__global__ void build_state_kernel(const RunState *st, int32_t *out) {
  const uint32_t b = st->status;
  int32_t code = 0;                                        // PGVD_OK
  if (b & ST_NEG_COORD) code = 3;                          // PGVD_ERR_BAD_COORD
  else if (b & ST_DUP_KEY) code = 9;                       // PGVD_ERR_DUP_KEY
  else if (b & (ST_STACK | ST_WALK)) code = 10;            // PGVD_ERR_INTERNAL
  else if (b & ST_TOO_LARGE) code = 8;                     // PGVD_ERR_TOO_LARGE
  else if (b & ST_CAPACITY) code = 7;                      // PGVD_ERR_CAPACITY
  out[0] = code;
  out[1] = (int32_t)st->m;
  out[2] = (int32_t)st->octree_size;
  out[3] = (int32_t)st->malformed;
}

__device__ inline uint64_t mix64(uint64_t z) {   // splitmix64 finaliser
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// out += sum over 32-bit words w_i of mix64(w_i | i << 32), mod 2^64.
__global__ __launch_bounds__(kBlock) void digest_kernel(const uint32_t *__restrict__ p, size_t n,
                                                        unsigned long long *out) {
  uint64_t acc = 0;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    acc += mix64((uint64_t)p[i] | ((uint64_t)i << 32));
  acc = wave_reduce_add(acc);
  if ((threadIdx.x & 63) == 0 && acc) atomicAdd(out, (unsigned long long)acc);
}

__global__ void zero_u64_kernel(unsigned long long *p) { *p = 0ull; }

// Same digest over count[0] * per_item 32-bit words, the count read on the
// device (a size only the device knows yet); n_max bounds the grid.
__global__ __launch_bounds__(kBlock) void digest_count_kernel(const uint32_t *__restrict__ p,
                                                              const uint32_t *count,
                                                              uint32_t per_item, size_t n_max,
                                                              unsigned long long *out) {
  size_t n = (size_t)(*count) * per_item;
  if (n > n_max) n = n_max;
  uint64_t acc = 0;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    acc += mix64((uint64_t)p[i] | ((uint64_t)i << 32));
  acc = wave_reduce_add(acc);
  if ((threadIdx.x & 63) == 0 && acc) atomicAdd(out, (unsigned long long)acc);
}

inline unsigned grid_for(size_t n) {
  size_t g = (n + kBlock * 4 - 1) / (kBlock * 4);
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (unsigned)g;
}

}  // namespace

void launch_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s) {
  fill_u32_kernel<<<grid_for(n), kBlock, 0, s>>>(p, n, v);
}

void launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
  // dst and src are device allocations (256-byte aligned bases); the copy is
  // done in 16-byte words with a byte tail.
  const size_t n16 = bytes / 16, tail = bytes - n16 * 16;
  copy_kernel<<<grid_for(n16 ? n16 : 1), kBlock, 0, s>>>(
      static_cast<uint4 *>(dst), static_cast<const uint4 *>(src), n16,
      static_cast<uint8_t *>(dst) + n16 * 16, static_cast<const uint8_t *>(src) + n16 * 16, tail);
}

void launch_digest(const void *p, size_t words, uint64_t *out, bool zero_first, hipStream_t s) {
  unsigned long long *o = reinterpret_cast<unsigned long long *>(out);
  if (zero_first) zero_u64_kernel<<<1, 1, 0, s>>>(o);
  if (words) digest_kernel<<<grid_for(words), kBlock, 0, s>>>(static_cast<const uint32_t *>(p), words, o);
}

void launch_digest_count(const void *p, const uint32_t *count, uint32_t per_item, size_t n_max,
                         uint64_t *out, hipStream_t s) {
  unsigned long long *o = reinterpret_cast<unsigned long long *>(out);
  zero_u64_kernel<<<1, 1, 0, s>>>(o);
  digest_count_kernel<<<grid_for(n_max ? n_max : 1), kBlock, 0, s>>>(
      static_cast<const uint32_t *>(p), count, per_item, n_max, o);
}

void launch_build_state(const RunState *st, int32_t *out4, hipStream_t s) {
  build_state_kernel<<<1, 1, 0, s>>>(st, out4);
}

void launch_set_sizes(RunState *st, const uint32_t *ps, int64_t m, hipStream_t s) {
  set_sizes_kernel<<<1, 1, 0, s>>>(st, ps, m);
}

}  // namespace pgvd
