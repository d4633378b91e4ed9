// stages.hip -- the reference's elementary stage kernels, kept so that every
// entry point of namespace Kernels (opencl/Kernels.h:26-50) has a HIP
// counterpart.  The pipeline itself does not use them: its radix sort,
// unique and octree fill replace these bit-serial steps.
//
//   bit predicate     BitPredicate (C/ParallelAlgorithms.c:13-18)
//   unique predicate  UniquePredicate (:22-31)
//   single compact    BUSingleCompact (:73-81)
//   double compact    BUCompact (:54-71), the split scatter of one radix bit
//   init octree       brt2octree_init (C/BuildOctree.c:134-139)
#include "kernels.hpp"

namespace pgvd {

namespace {

inline unsigned grid_of(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  return (unsigned)(g < 1 ? 1 : g);
}

template <typename K>
__device__ inline uint32_t kbit(const K &k, int index);
template <> __device__ inline uint32_t kbit<uint32_t>(const uint32_t &k, int i) { return i < 32 ? (k >> i) & 1u : 0u; }
template <> __device__ inline uint32_t kbit<uint64_t>(const uint64_t &k, int i) { return i < 64 ? (uint32_t)(k >> i) & 1u : 0u; }
template <> __device__ inline uint32_t kbit<Key128>(const Key128 &k, int i) {
  return i < 64 ? (uint32_t)(k.lo >> i) & 1u : (i < 128 ? (uint32_t)(k.hi >> (i - 64)) & 1u : 0u);
}

template <typename K>
__global__ __launch_bounds__(kBlock) void bit_predicate_kernel(const K *__restrict__ in,
                                                               uint32_t *__restrict__ pred,
                                                               int index, uint32_t compared,
                                                               int64_t n) {
  const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (g < n) pred[g] = kbit(in[g], index) == compared;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void unique_predicate_kernel(const K *__restrict__ in,
                                                                  uint32_t *__restrict__ pred,
                                                                  int64_t n) {
  const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (g >= n) return;
  pred[g] = (g == 0) ? 1u : (uint32_t)!keq(in[g], in[g - 1]);
}

// result[address[g] - 1] = in[g] where pred[g] == 1 (address: inclusive scan)
template <typename K>
__global__ __launch_bounds__(kBlock) void single_compact_kernel(const K *__restrict__ in,
                                                                K *__restrict__ out,
                                                                const uint32_t *__restrict__ pred,
                                                                const uint32_t *__restrict__ addr,
                                                                int64_t n) {
  const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (g < n && pred[g] == 1) out[addr[g] - 1] = in[g];
}

// The split scatter of BUCompact: keys whose predicate is 1 go to
// addr[g] - 1, the others after all of those, in order.
template <typename K>
__global__ __launch_bounds__(kBlock) void double_compact_kernel(const K *__restrict__ in,
                                                                K *__restrict__ out,
                                                                const uint32_t *__restrict__ pred,
                                                                const uint32_t *__restrict__ addr,
                                                                int64_t n) {
  const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (g >= n) return;
  const int64_t a = addr[g];
  const int64_t total = (int64_t)addr[n - 1];   // = addr[n-2] + pred[n-1]
  const int64_t d = pred[g] ? a - 1 : g - a + total;
  out[d] = in[g];
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void init_octree_kernel(int32_t *octree, int64_t nodes) {
  const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (g >= nodes) return;
  constexpr int NC = 1 << DIM;
  int32_t *o = octree + g * (NC + 1);
#pragma unroll
  for (int c = 0; c < NC; ++c) o[c] = -1;
  o[NC] = (1 << NC) - 1;
}

}  // namespace

template <typename K>
void launch_bit_predicate(const K *in, uint32_t *pred, int index, uint32_t compared, int64_t n,
                          hipStream_t s) {
  bit_predicate_kernel<K><<<grid_of(n), kBlock, 0, s>>>(in, pred, index, compared, n);
}
template <typename K>
void launch_unique_predicate(const K *in, uint32_t *pred, int64_t n, hipStream_t s) {
  unique_predicate_kernel<K><<<grid_of(n), kBlock, 0, s>>>(in, pred, n);
}
template <typename K>
void launch_single_compact(const K *in, K *out, const uint32_t *pred, const uint32_t *addr,
                           int64_t n, hipStream_t s) {
  single_compact_kernel<K><<<grid_of(n), kBlock, 0, s>>>(in, out, pred, addr, n);
}
template <typename K>
void launch_double_compact(const K *in, K *out, const uint32_t *pred, const uint32_t *addr,
                           int64_t n, hipStream_t s) {
  double_compact_kernel<K><<<grid_of(n), kBlock, 0, s>>>(in, out, pred, addr, n);
}
void launch_init_octree(int32_t *octree, int64_t nodes, int dim, hipStream_t s) {
  if (dim == 2) init_octree_kernel<2><<<grid_of(nodes), kBlock, 0, s>>>(octree, nodes);
  else init_octree_kernel<3><<<grid_of(nodes), kBlock, 0, s>>>(octree, nodes);
}

#define PGVD_INST(K)                                                                          \
  template void launch_bit_predicate<K>(const K *, uint32_t *, int, uint32_t, int64_t,         \
                                        hipStream_t);                                          \
  template void launch_unique_predicate<K>(const K *, uint32_t *, int64_t, hipStream_t);       \
  template void launch_single_compact<K>(const K *, K *, const uint32_t *, const uint32_t *,   \
                                         int64_t, hipStream_t);                                \
  template void launch_double_compact<K>(const K *, K *, const uint32_t *, const uint32_t *,   \
                                         int64_t, hipStream_t);
PGVD_INST(uint32_t)
PGVD_INST(uint64_t)
PGVD_INST(Key128)
#undef PGVD_INST

}  // namespace pgvd
