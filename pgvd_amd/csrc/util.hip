// util.hip -- small utility kernels: word fill (used instead of
// hipMemsetAsync so that captured graphs replay correctly), a plain copy
// kernel (also the calibration kernel for HBM byte counters), and setting
// the run sizes from a prefix-sum array.
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
  const size_t tid = (size_t)blockIdx.x * kBlock + threadIdx.x;
  for (size_t i = tid; i < n16; i += (size_t)gridDim.x * kBlock) dst[i] = src[i];
  if (tid < ntail) dtail[tid] = stail[tid];
}

__global__ void set_sizes_kernel(RunState *st, const uint32_t *ps, int64_t m) {
  st->m = (uint32_t)m;
  st->octree_size = ps[m - 1];
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

void launch_set_sizes(RunState *st, const uint32_t *ps, int64_t m, hipStream_t s) {
  set_sizes_kernel<<<1, 1, 0, s>>>(st, ps, m);
}

}  // namespace pgvd
