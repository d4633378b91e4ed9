// morton.hip -- run-state prep, bounding box, FP32 quantize + Morton encode
// with fused radix digit histograms.
//
// Reference behaviour followed:
//   bbox        BoundingBox.h:166-175 / :80-87, Karras.cpp:75-83
//   quantize    Karras.cpp:42-67 (FP32 expression :53-55), :84-88 (dwidth 0)
//   Morton      C/z_order.h:14-47 (PointsToMortonKernel, kernels.cl:2-22)
//   padding     opencl/Kernels.cpp:62-75 (zero keys past n)
#include "kernels.hpp"

namespace pgvd {

namespace {

constexpr int kMaxGrid = 2048;

inline int grid_for(int64_t n, int per_block) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > kMaxGrid) g = kMaxGrid;
  return (int)g;
}

// A build's start: st reset, the zero-initialised workspace cleared.
__device__ inline void prep_run(RunState *st, uint4 *zero, size_t nzero16) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid == 0) {
    for (int k = 0; k < 3; ++k) { st->bb_min[k] = 0xFFFFFFFFu; st->bb_max[k] = 0u; }
    st->status = 0; st->malformed = 0; st->m = 0; st->octree_size = 0; st->msd_ok = 0;
    for (int k = 0; k < kNumCounters; ++k) st->counters[k] = 0;
    st->msd_dups = 0;
    st->bk_segs = 0; st->bk_keys = 0; st->bk_next_segs = 0; st->bk_next_keys = 0;
    st->bk_items = 0; st->bk_tiles = 0; st->bk_lsd = 0;
  }
  const uint4 z = make_uint4(0, 0, 0, 0);
  for (size_t i = tid; i < nzero16; i += (size_t)gridDim.x * blockDim.x) zero[i] = z;
}

__global__ __launch_bounds__(kBlock) void prep_kernel(RunState *st, uint4 *zero, size_t nzero16) {
  prep_run(st, zero, nzero16);
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void bbox_kernel(const float *__restrict__ pts, int64_t n,
                                                       RunState *st) {
  float mn[DIM], mx[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) { mn[k] = __int_as_float(0x7f800000); mx[k] = -mn[k]; }
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const float v = pts[i * DIM + k];
      mn[k] = fminf(mn[k], v);
      mx[k] = fmaxf(mx[k], v);
    }
  }
  __shared__ float smn[4][DIM], smx[4][DIM];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    for (int d = 32; d >= 1; d >>= 1) {
      mn[k] = fminf(mn[k], __shfl_xor(mn[k], d, 64));
      mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], d, 64));
    }
    if (lane == 0) { smn[w][k] = mn[k]; smx[w][k] = mx[k]; }
  }
  __syncthreads();
  if (threadIdx.x < DIM) {
    const int k = threadIdx.x;
    float a = smn[0][k], b = smx[0][k];
    for (int j = 1; j < 4; ++j) { a = fminf(a, smn[j][k]); b = fmaxf(b, smx[j][k]); }
    atomicMin(&st->bb_min[k], fenc(a));
    atomicMax(&st->bb_max[k], fenc(b));
  }
}

// One point -> quantized coordinates (Karras.cpp:50-65).  Each FP32 step is
// an explicitly rounded op so no contraction or reassociation can change it.
// SMALL: values in [0, 2^32) take the one-instruction u32 conversion, which
// truncates the same way; others the general 64-bit one.
template <int DIM, bool SMALL = false>
__device__ inline bool quantize_point(const float *p, const float *mn, float dwidth, float effw,
                                      int64_t *c) {
  bool neg = false;
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    if (dwidth == 0.0f) { c[k] = 0; continue; }
    const float t = __fdiv_rn(__fsub_rn(p[k], mn[k]), dwidth);
    const float f = __fmul_rn(effw, t);
    const double d = __dadd_rn((double)f, 0.5);
    const int64_t v = (SMALL && d >= 0.0 && d < 4294967296.0) ? (int64_t)__double2uint_rz(d)
                                                               : __double2ll_rz(d);
    neg |= (v < 0);
    c[k] = v;
  }
  return neg;
}

template <typename K, int DIM>
__global__ __launch_bounds__(kBlock) void encode_kernel(
    const float *__restrict__ pts, int64_t n, int bits, float effw, int custom, float bmin0,
    float bmin1, float bmin2, float bmax0, float bmax1, float bmax2, K *__restrict__ keys,
    int64_t *__restrict__ qout, uint32_t *__restrict__ hist, int passes, RunState *st) {
  __shared__ uint32_t sh[kMaxPasses * 256];
  for (int i = threadIdx.x; i < passes * 256; i += kBlock) sh[i] = 0;
  float mn[3], mx[3];
  if (custom) {
    mn[0] = bmin0; mn[1] = bmin1; mn[2] = bmin2;
    mx[0] = bmax0; mx[1] = bmax1; mx[2] = bmax2;
  } else {
#pragma unroll
    for (int k = 0; k < DIM; ++k) { mn[k] = fdec(st->bb_min[k]); mx[k] = fdec(st->bb_max[k]); }
  }
  // BoundingBox::max_size (BoundingBox.h:80-87)
  float dwidth = __fsub_rn(mx[0], mn[0]);
#pragma unroll
  for (int k = 1; k < DIM; ++k) {
    const float s = __fsub_rn(mx[k], mn[k]);
    dwidth = (s > dwidth) ? s : dwidth;
  }
  __syncthreads();
  bool neg = false;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    float p[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) p[k] = pts[i * DIM + k];
    int64_t c[DIM];
    neg |= quantize_point<DIM>(p, mn, dwidth, effw, c);
    if (qout) {
#pragma unroll
      for (int k = 0; k < DIM; ++k) qout[i * DIM + k] = c[k];
    }
    uint64_t lo, hi;
    morton128<DIM, (sizeof(K) > 8)>(c, bits, lo, hi);
    const K key = from128<K>(lo, hi);
    if (keys) keys[i] = key;
    for (int r = 0; r < passes; ++r) atomicAdd(&sh[r * 256 + kdigit(key, 8 * r)], 1u);
  }
  if (neg) atomicOr(&st->status, (uint32_t)ST_NEG_COORD);
  if (hist) {
    __syncthreads();
    for (int i = threadIdx.x; i < passes * 256; i += kBlock)
      if (sh[i]) atomicAdd(&hist[i], sh[i]);
  }
}

// ---- vectorised forms: a chunk is 4 points = DIM float4 loads ------------
template <int DIM>
__device__ inline void load_chunk(const float *__restrict__ pts, int64_t ch, float (&f)[4 * DIM]) {
  const float4 *p4 = reinterpret_cast<const float4 *>(pts) + ch * DIM;
#pragma unroll
  for (int v = 0; v < DIM; ++v) {
    const float4 a = p4[v];
    f[4 * v + 0] = a.x; f[4 * v + 1] = a.y; f[4 * v + 2] = a.z; f[4 * v + 3] = a.w;
  }
}

// Bounding box over the points as a flat float array: thread t reads float4
// number f = B + 256 u + t (B a multiple of 3 * 256 * kBboxU4), so each load
// instruction of a wave covers 1 KiB contiguously.  In 3D float f * 4 + j is
// component (f + j) % 3 and f % 3 = (u + t) % 3: lanes are accumulated per
// u % 3 and mapped to components once at the end; in 2D float f * 4 + j is
// component j % 2.  Floats past the last whole float4 are read one by one.
#ifndef PGVD_BBOX_U4
#define PGVD_BBOX_U4 12
#endif
#ifndef PGVD_BBOX_ROUNDS
#define PGVD_BBOX_ROUNDS 2
#endif
// Non-temporal point loads (1).  Measured at 2^24 3D points: the pass takes
// 34 us per launch instead of 54, and encode, which reads the points
// again, 90 instead of 77 (it finds fewer of them cached); the build is
// 7 us shorter.
#ifndef PGVD_BBOX_NT
#define PGVD_BBOX_NT 1
#endif
constexpr int kBboxU4 = PGVD_BBOX_U4;   // float4 loads in flight per thread (multiple of 3)

// With prep_st, also the build's start (prep_run): one launch fewer.
template <int DIM>
__global__ __launch_bounds__(kBlock) void bbox_vec_kernel(const float *__restrict__ pts, int64_t n,
                                                          float *__restrict__ partials,
                                                          RunState *prep_st, uint4 *zero,
                                                          size_t nzero16) {
  if (prep_st) prep_run(prep_st, zero, nzero16);
  static_assert(kBboxU4 % 3 == 0, "per-round span is a multiple of 3 float4s");
  constexpr int NA = DIM == 3 ? 3 : 1;
  const float kNan = __int_as_float(0x7fc00000);   // fminf / fmaxf skip it
  const float kInf = __int_as_float(0x7f800000);
  float amn[NA][4], amx[NA][4];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int j = 0; j < 4; ++j) { amn[a][j] = kInf; amx[a][j] = -kInf; }
  const int64_t nf4 = n * DIM / 4;
  const float4 *p4 = reinterpret_cast<const float4 *>(pts);
  const int64_t span = (int64_t)kBlock * kBboxU4;
  for (int64_t B = (int64_t)blockIdx.x * span; B < nf4; B += (int64_t)gridDim.x * span) {
    float4 v[kBboxU4];
#pragma unroll
    for (int u = 0; u < kBboxU4; ++u) {
      const int64_t f = B + (int64_t)u * kBlock + threadIdx.x;
#if PGVD_BBOX_NT
      {
        typedef float f4v __attribute__((ext_vector_type(4)));
        const f4v x = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(p4) + (f < nf4 ? f : 0));
        v[u] = make_float4(x.x, x.y, x.z, x.w);
      }
#else
      v[u] = p4[f < nf4 ? f : 0];
#endif
      if (f >= nf4) v[u] = make_float4(kNan, kNan, kNan, kNan);
    }
#pragma unroll
    for (int u = 0; u < kBboxU4; ++u) {
      const int a = u % NA;
      amn[a][0] = fminf(amn[a][0], v[u].x); amx[a][0] = fmaxf(amx[a][0], v[u].x);
      amn[a][1] = fminf(amn[a][1], v[u].y); amx[a][1] = fmaxf(amx[a][1], v[u].y);
      amn[a][2] = fminf(amn[a][2], v[u].z); amx[a][2] = fmaxf(amx[a][2], v[u].z);
      amn[a][3] = fminf(amn[a][3], v[u].w); amx[a][3] = fmaxf(amx[a][3], v[u].w);
    }
  }
  float mn[DIM], mx[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) { mn[k] = kInf; mx[k] = -kInf; }
  const int t3 = (int)(threadIdx.x % 3u);
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int comp = DIM == 3 ? (a + j + t3) % 3 : (j & 1);
#pragma unroll
      for (int k = 0; k < DIM; ++k)
        if (comp == k) { mn[k] = fminf(mn[k], amn[a][j]); mx[k] = fmaxf(mx[k], amx[a][j]); }
    }
  // floats after the last whole float4
  for (int64_t e = nf4 * 4 + (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n * DIM;
       e += (int64_t)gridDim.x * kBlock) {
    const int k = (int)(e % DIM);
#pragma unroll
    for (int kk = 0; kk < DIM; ++kk)
      if (kk == k) { mn[kk] = fminf(mn[kk], pts[e]); mx[kk] = fmaxf(mx[kk], pts[e]); }
  }
  __shared__ float smn[4][DIM], smx[4][DIM];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    for (int d = 32; d >= 1; d >>= 1) {
      mn[k] = fminf(mn[k], __shfl_xor(mn[k], d, 64));
      mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], d, 64));
    }
    if (lane == 0) { smn[w][k] = mn[k]; smx[w][k] = mx[k]; }
  }
  __syncthreads();
  // one record per workgroup (min[0..3), max[4..7)); no atomics, which would
  // all meet on one cache line
  if (threadIdx.x < DIM) {
    const int k = threadIdx.x;
    float a = smn[0][k], b = smx[0][k];
    for (int j = 1; j < 4; ++j) { a = fminf(a, smn[j][k]); b = fmaxf(b, smx[j][k]); }
    partials[blockIdx.x * 8 + k] = a;
    partials[blockIdx.x * 8 + 4 + k] = b;
  }
}

// Folds the per-workgroup records of bbox_vec_kernel into st's box.
template <int DIM>
__global__ __launch_bounds__(kBlock) void bbox_finish_kernel(const float *__restrict__ partials,
                                                             int nparts, RunState *st) {
  const float kInf = __int_as_float(0x7f800000);
  float mn[DIM], mx[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) { mn[k] = kInf; mx[k] = -kInf; }
  for (int b = threadIdx.x; b < nparts; b += kBlock)
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      mn[k] = fminf(mn[k], partials[b * 8 + k]);
      mx[k] = fmaxf(mx[k], partials[b * 8 + 4 + k]);
    }
  __shared__ float smn[4][DIM], smx[4][DIM];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    for (int d = 32; d >= 1; d >>= 1) {
      mn[k] = fminf(mn[k], __shfl_xor(mn[k], d, 64));
      mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], d, 64));
    }
    if (lane == 0) { smn[w][k] = mn[k]; smx[w][k] = mx[k]; }
  }
  __syncthreads();
  if (threadIdx.x < DIM) {
    const int k = threadIdx.x;
    float a = smn[0][k], b = smx[0][k];
    for (int j = 1; j < 4; ++j) { a = fminf(a, smn[j][k]); b = fmaxf(b, smx[j][k]); }
    st->bb_min[k] = min(st->bb_min[k], fenc(a));
    st->bb_max[k] = max(st->bb_max[k], fenc(b));
  }
}

template <typename K>
__device__ inline void store4(K *__restrict__ keys, int64_t ch, const K (&k)[4]) {
  static_assert((4 * sizeof(K)) % 16 == 0, "4 keys are whole 16-byte words");
  uint4 *dst = reinterpret_cast<uint4 *>(keys + ch * 4);
  const uint4 *src = reinterpret_cast<const uint4 *>(k);
#pragma unroll
  for (int v = 0; v < (int)(4 * sizeof(K) / 16); ++v) dst[v] = src[v];
}

template <typename K, int DIM, int HBINS>
__global__ __launch_bounds__(kBlock) void encode_vec_kernel(
    const float *__restrict__ pts, int64_t n, int bits, float effw, int custom, float bmin0,
    float bmin1, float bmin2, float bmax0, float bmax1, float bmax2, K *__restrict__ keys,
    uint32_t *__restrict__ hist, int passes, RunState *st, uint32_t *__restrict__ msd_hist,
    int msd_shift, int msd_nb) {
  // One LDS histogram serves either the LSD digit counts (passes > 0) or the
  // bucket counts (msd_hist, built with passes = 0): 16 KB instead of 32 KB
  // keeps eight workgroups per CU.  HBINS > 4096: the bucket counts of the
  // largest builds (up to 2^14 buckets, 64 KB).
  static_assert(kMaxPasses * 256 == (1 << kEncodeMaxBucketBits),
                "the two histograms share their LDS");
  static_assert(HBINS >= kMaxPasses * 256, "the LDS holds the digit histograms too");
  __shared__ uint32_t sh[HBINS];
  uint32_t *s_msd = sh;
  if (msd_hist) passes = 0;
  for (int i = threadIdx.x; i < passes * 256; i += kBlock) sh[i] = 0;
  // 3D keys of at most 64 bits: Morton spread by table, byte by byte
  // (s_spread3[v] = bit i of v at bit 3i), instead of 64-bit mask steps.
  constexpr bool kTable = DIM == 3 && sizeof(K) == 8;
  __shared__ uint32_t s_spread3[kTable ? 256 : 1];
  if (kTable) s_spread3[threadIdx.x] = (uint32_t)spread3(threadIdx.x);
  static_assert(!kTable || kBlock == 256, "one table entry per thread");
  const uint32_t cmask32 = bits >= 32 ? ~0u : (1u << bits) - 1u;
  if (msd_hist)
    for (int i = threadIdx.x; i < msd_nb; i += kBlock) s_msd[i] = 0;
  const uint32_t msd_mask = (uint32_t)msd_nb - 1;
  float mn[3], mx[3];
  if (custom) {
    mn[0] = bmin0; mn[1] = bmin1; mn[2] = bmin2;
    mx[0] = bmax0; mx[1] = bmax1; mx[2] = bmax2;
  } else {
#pragma unroll
    for (int k = 0; k < DIM; ++k) { mn[k] = fdec(st->bb_min[k]); mx[k] = fdec(st->bb_max[k]); }
  }
  float dwidth = __fsub_rn(mx[0], mn[0]);
#pragma unroll
  for (int k = 1; k < DIM; ++k) {
    const float s = __fsub_rn(mx[k], mn[k]);
    dwidth = (s > dwidth) ? s : dwidth;
  }
  __syncthreads();
  bool neg = false;
  const int64_t nch = n / 4;
  const int64_t tid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  auto encode_one = [&](const float *p) -> K {
    int64_t c[DIM];
    neg |= quantize_point<DIM>(p, mn, dwidth, effw, c);
    uint64_t lo, hi = 0;
    if constexpr (kTable) {
      // bits <= 21: the three coordinates' bytes j give key bits 24j ..
      const uint32_t x = (uint32_t)c[0] & cmask32, y = (uint32_t)c[1] & cmask32,
                     z = (uint32_t)c[2] & cmask32;
      uint64_t w[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int sh = 8 * j;
        w[j] = s_spread3[(x >> sh) & 255u] | s_spread3[(y >> sh) & 255u] << 1 |
               s_spread3[(z >> sh) & 255u] << 2;
      }
      lo = w[0] | w[1] << 24 | w[2] << 48;
    } else {
      morton128<DIM, (sizeof(K) > 8)>(c, bits, lo, hi);
    }
    const K key = from128<K>(lo, hi);
    for (int r = 0; r < passes; ++r) atomicAdd(&sh[r * 256 + kdigit(key, 8 * r)], 1u);
    if (msd_hist) atomicAdd(&s_msd[klow32(kshr(key, msd_shift)) & msd_mask], 1u);
    return key;
  };
  for (int64_t ch = tid; ch < nch; ch += stride) {
    float f[4 * DIM];
    load_chunk<DIM>(pts, ch, f);
    K kk[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) kk[j] = encode_one(f + j * DIM);
    store4(keys, ch, kk);
  }
  for (int64_t i = nch * 4 + tid; i < n; i += stride) keys[i] = encode_one(pts + i * DIM);
  if (neg) atomicOr(&st->status, (uint32_t)ST_NEG_COORD);
  __syncthreads();
  for (int i = threadIdx.x; i < passes * 256; i += kBlock)
    if (sh[i]) atomicAdd(&hist[i], sh[i]);
  if (msd_hist)
    for (int i = threadIdx.x; i < msd_nb; i += kBlock)
      if (s_msd[i]) atomicAdd(&msd_hist[i], s_msd[i]);
}

template <typename K, int DIM, typename C>
__global__ __launch_bounds__(kBlock) void points_to_morton_kernel(const C *__restrict__ pts,
                                                                  int64_t n, int64_t padded_n,
                                                                  int bits, K *__restrict__ keys) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < padded_n;
       i += (int64_t)gridDim.x * kBlock) {
    K key = kzero_value<K>();
    if (i < n) {
      int64_t c[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) c[k] = (int64_t)pts[i * DIM + k];
      uint64_t lo, hi;
      morton128<DIM, (sizeof(K) > 8)>(c, bits, lo, hi);
      key = from128<K>(lo, hi);
    }
    keys[i] = key;
  }
}

template <typename K, int DIM>
__global__ __launch_bounds__(kBlock) void encode_int_kernel(const int32_t *__restrict__ pts,
                                                            int64_t n, int bits,
                                                            K *__restrict__ keys,
                                                            uint32_t *__restrict__ hist,
                                                            int passes) {
  __shared__ uint32_t sh[kMaxPasses * 256];
  for (int i = threadIdx.x; i < passes * 256; i += kBlock) sh[i] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    int64_t c[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) c[k] = (int64_t)pts[i * DIM + k];
    uint64_t lo, hi;
    morton128<DIM, (sizeof(K) > 8)>(c, bits, lo, hi);
    const K key = from128<K>(lo, hi);
    keys[i] = key;
    for (int r = 0; r < passes; ++r) atomicAdd(&sh[r * 256 + kdigit(key, 8 * r)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < passes * 256; i += kBlock)
    if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

template <typename K>
__global__ __launch_bounds__(kBlock) void histogram_kernel(const K *__restrict__ keys, int64_t n,
                                                           uint32_t *__restrict__ hist,
                                                           int passes, const RunState *skip,
                                                           uint4 *zero, size_t nzero16) {
  if (skip && skip->msd_ok) return;
  // words the passes need zeroed that the build's start left alone
  const uint4 z = make_uint4(0, 0, 0, 0);
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < nzero16;
       i += (size_t)gridDim.x * kBlock)
    zero[i] = z;
  __shared__ uint32_t sh[kMaxPasses * 256];
  for (int i = threadIdx.x; i < passes * 256; i += kBlock) sh[i] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const K key = keys[i];
    for (int r = 0; r < passes; ++r) atomicAdd(&sh[r * 256 + kdigit(key, 8 * r)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < passes * 256; i += kBlock)
    if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

// ---- batched form: one workgroup per chunk of kEncodeChunk points of one
// instance ----
__global__ __launch_bounds__(kBlock) void seg_init_kernel(SegState *ss, int64_t B) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= B) return;
  SegState z;
  for (int k = 0; k < 3; ++k) { z.bb_min[k] = 0xFFFFFFFFu; z.bb_max[k] = 0u; }
  z.m = 0; z.octree_size = 0; z.oct_off = 0;
  for (int k = 0; k < 3; ++k) z.pad_[k] = 0;
  ss[i] = z;
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void bbox_seg_kernel(const float *__restrict__ pts,
                                                          const SegInfo *__restrict__ seg,
                                                          TileTable ct, SegState *ss) {
  const uint32_t chunk = blockIdx.x;
  const uint32_t s = ct.seg[chunk];
  const SegInfo si = seg[s];
  const uint32_t start = (chunk - ct.first[s]) * (uint32_t)kEncodeChunk;
  const uint32_t cnt = min((uint32_t)kEncodeChunk, si.n - start);
  const float *p = pts + (si.pofs + start) * DIM;
  float mn[DIM], mx[DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k) { mn[k] = __int_as_float(0x7f800000); mx[k] = -mn[k]; }
  for (uint32_t i = threadIdx.x; i < cnt; i += kBlock) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const float v = p[(size_t)i * DIM + k];
      mn[k] = fminf(mn[k], v);
      mx[k] = fmaxf(mx[k], v);
    }
  }
  __shared__ float smn[4][DIM], smx[4][DIM];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    for (int d = 32; d >= 1; d >>= 1) {
      mn[k] = fminf(mn[k], __shfl_xor(mn[k], d, 64));
      mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], d, 64));
    }
    if (lane == 0) { smn[w][k] = mn[k]; smx[w][k] = mx[k]; }
  }
  __syncthreads();
  if (threadIdx.x < DIM) {
    const int k = threadIdx.x;
    float a = smn[0][k], b = smx[0][k];
    for (int j = 1; j < 4; ++j) { a = fminf(a, smn[j][k]); b = fmaxf(b, smx[j][k]); }
    atomicMin(&ss[s].bb_min[k], fenc(a));
    atomicMax(&ss[s].bb_max[k], fenc(b));
  }
}

template <typename K, int DIM, bool INT>
__global__ __launch_bounds__(kBlock) void encode_seg_kernel(
    const void *__restrict__ pts, const SegInfo *__restrict__ seg, TileTable ct,
    const SegState *__restrict__ ss, int bits, float effw, K *__restrict__ keys,
    uint32_t *__restrict__ bhist, int passes, int dbits, RunState *st) {
  __shared__ uint32_t sh[kMaxPasses * 256];
  const uint32_t dmask = (1u << dbits) - 1u;
  for (int i = threadIdx.x; i < passes * 256; i += kBlock) sh[i] = 0;
  const uint32_t chunk = blockIdx.x;
  const uint32_t s = ct.seg[chunk];
  const SegInfo si = seg[s];
  const uint32_t start = (chunk - ct.first[s]) * (uint32_t)kEncodeChunk;
  const uint32_t cnt = min((uint32_t)kEncodeChunk, si.n - start);
  float mn[3] = {0.f, 0.f, 0.f};
  float dwidth = 0.f;
  if (!INT) {
    float mx[3];
#pragma unroll
    for (int k = 0; k < DIM; ++k) { mn[k] = fdec(ss[s].bb_min[k]); mx[k] = fdec(ss[s].bb_max[k]); }
    // BoundingBox::max_size (BoundingBox.h:80-87)
    dwidth = __fsub_rn(mx[0], mn[0]);
#pragma unroll
    for (int k = 1; k < DIM; ++k) {
      const float e = __fsub_rn(mx[k], mn[k]);
      dwidth = (e > dwidth) ? e : dwidth;
    }
  }
  __syncthreads();
  bool neg = false;
  K *out = keys + si.kofs + start;
  for (uint32_t i = threadIdx.x; i < cnt; i += kBlock) {
    const size_t pi = (size_t)(si.pofs + start + i) * DIM;
    int64_t c[DIM];
    if (INT) {
      const int32_t *q = static_cast<const int32_t *>(pts) + pi;
#pragma unroll
      for (int k = 0; k < DIM; ++k) c[k] = q[k];
    } else {
      const float *q = static_cast<const float *>(pts) + pi;
      float f[DIM];
#pragma unroll
      for (int k = 0; k < DIM; ++k) f[k] = q[k];
      neg |= quantize_point<DIM>(f, mn, dwidth, effw, c);
    }
    uint64_t lo, hi;
    morton128<DIM, (sizeof(K) > 8)>(c, bits, lo, hi);
    const K key = from128<K>(lo, hi);
    out[i] = key;
    for (int r = 0; r < passes; ++r) atomicAdd(&sh[r * 256 + kdigit_w(key, dbits * r, dmask)], 1u);
  }
  if (neg) atomicOr(&st->status, (uint32_t)ST_NEG_COORD);
  __syncthreads();
  uint32_t *h = bhist + (size_t)s * passes * 256;
  for (int i = threadIdx.x; i < passes * 256; i += kBlock)
    if (sh[i]) atomicAdd(&h[i], sh[i]);
}

}  // namespace

void launch_prep(RunState *st, uint32_t *zero_words, size_t nzero, hipStream_t s) {
  // zero_words is 16-byte aligned and nzero a multiple of 4 (the workspace
  // layout guarantees both).
  const size_t n16 = nzero / 4;
  const int grid = grid_for((int64_t)(n16 ? n16 : 1), kBlock * 4);
  prep_kernel<<<grid, kBlock, 0, s>>>(st, reinterpret_cast<uint4 *>(zero_words), n16);
}

inline bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

void launch_bbox(const float *pts, int64_t n, int dim, float *partials, RunState *st,
                 hipStream_t s, uint32_t *zero_words, size_t nzero, bool prep) {
  static_assert(kMaxGrid <= kBboxParts, "one partial record per workgroup");
  if (aligned16(pts)) {
    const int grid = grid_for(n * dim / 4 + 1, kBlock * kBboxU4 * PGVD_BBOX_ROUNDS);
    RunState *pst = prep ? st : nullptr;
    uint4 *z16 = reinterpret_cast<uint4 *>(zero_words);
    const size_t n16 = prep ? nzero / 4 : 0;
    if (dim == 2) {
      bbox_vec_kernel<2><<<grid, kBlock, 0, s>>>(pts, n, partials, pst, z16, n16);
      bbox_finish_kernel<2><<<1, kBlock, 0, s>>>(partials, grid, st);
    } else {
      bbox_vec_kernel<3><<<grid, kBlock, 0, s>>>(pts, n, partials, pst, z16, n16);
      bbox_finish_kernel<3><<<1, kBlock, 0, s>>>(partials, grid, st);
    }
    return;
  }
  if (prep) launch_prep(st, zero_words, nzero, s);
  const int grid = grid_for(n, kBlock * 8);
  if (dim == 2) bbox_kernel<2><<<grid, kBlock, 0, s>>>(pts, n, st);
  else bbox_kernel<3><<<grid, kBlock, 0, s>>>(pts, n, st);
}

bool encode_vectorised(const float *pts, const void *keys) {
  return aligned16(pts) && aligned16(keys);
}

// Workgroups of the vectorised encode at most.  Each flushes its histograms
// (up to 4096 bucket counters) with global atomics once: at 2^24 3D points
// 1024 workgroups took 79 us, 2048 took 90 and 512 took 96.
#ifndef PGVD_ENCODE_GRID
#define PGVD_ENCODE_GRID 1024
#endif
template <typename K>
void launch_encode(const float *pts, int64_t n, int dim, int bits, float effw, int custom_bb,
                   const float *bbmin, const float *bbmax, K *keys, int64_t *qout,
                   uint32_t *hist, int passes, RunState *st, hipStream_t s, uint32_t *msd_hist,
                   int msd_shift, int msd_nb) {
  const int grid = grid_for(n, kBlock * 4);
  float a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
  if (custom_bb)
    for (int k = 0; k < dim; ++k) { a[k] = bbmin[k]; b[k] = bbmax[k]; }
  if (keys && hist && !qout && encode_vectorised(pts, keys)) {
    const int vgrid = std::min(grid_for(n / 4 + 1, kBlock), PGVD_ENCODE_GRID);
    if (msd_nb > kBkMaxBuckets) msd_hist = nullptr;
    const bool big = msd_hist && msd_nb > (1 << kEncodeMaxBucketBits);
    if (dim == 2 && !big)
      encode_vec_kernel<K, 2, 4096><<<vgrid, kBlock, 0, s>>>(pts, n, bits, effw, custom_bb, a[0],
                                                             a[1], a[2], b[0], b[1], b[2], keys,
                                                             hist, passes, st, msd_hist,
                                                             msd_shift, msd_nb);
    else if (dim == 2)
      encode_vec_kernel<K, 2, kBkMaxBuckets><<<vgrid, kBlock, 0, s>>>(
          pts, n, bits, effw, custom_bb, a[0], a[1], a[2], b[0], b[1], b[2], keys, hist, passes,
          st, msd_hist, msd_shift, msd_nb);
    else if (!big)
      encode_vec_kernel<K, 3, 4096><<<vgrid, kBlock, 0, s>>>(pts, n, bits, effw, custom_bb, a[0],
                                                             a[1], a[2], b[0], b[1], b[2], keys,
                                                             hist, passes, st, msd_hist,
                                                             msd_shift, msd_nb);
    else
      encode_vec_kernel<K, 3, kBkMaxBuckets><<<vgrid, kBlock, 0, s>>>(
          pts, n, bits, effw, custom_bb, a[0], a[1], a[2], b[0], b[1], b[2], keys, hist, passes,
          st, msd_hist, msd_shift, msd_nb);
    return;
  }
  if (dim == 2)
    encode_kernel<K, 2><<<grid, kBlock, 0, s>>>(pts, n, bits, effw, custom_bb, a[0], a[1], a[2],
                                                b[0], b[1], b[2], keys, qout, hist, passes, st);
  else
    encode_kernel<K, 3><<<grid, kBlock, 0, s>>>(pts, n, bits, effw, custom_bb, a[0], a[1], a[2],
                                                b[0], b[1], b[2], keys, qout, hist, passes, st);
}

template <typename K>
void launch_points_to_morton(const void *pts, int coord_bytes, int64_t n, int64_t padded_n,
                             int dim, int bits, K *keys, hipStream_t s) {
  const int grid = grid_for(padded_n, kBlock * 4);
  if (coord_bytes == 4) {
    const int32_t *p = static_cast<const int32_t *>(pts);
    if (dim == 2) points_to_morton_kernel<K, 2, int32_t><<<grid, kBlock, 0, s>>>(p, n, padded_n, bits, keys);
    else points_to_morton_kernel<K, 3, int32_t><<<grid, kBlock, 0, s>>>(p, n, padded_n, bits, keys);
  } else {
    const int64_t *p = static_cast<const int64_t *>(pts);
    if (dim == 2) points_to_morton_kernel<K, 2, int64_t><<<grid, kBlock, 0, s>>>(p, n, padded_n, bits, keys);
    else points_to_morton_kernel<K, 3, int64_t><<<grid, kBlock, 0, s>>>(p, n, padded_n, bits, keys);
  }
}

template <typename K>
void launch_histogram(const K *keys, int64_t n, uint32_t *hist, int passes, hipStream_t s,
                      const RunState *skip_if_msd, uint32_t *zero_words, size_t nzero) {
  const int grid = grid_for(n, kBlock * 8);
  histogram_kernel<K><<<grid, kBlock, 0, s>>>(keys, n, hist, passes, skip_if_msd,
                                              reinterpret_cast<uint4 *>(zero_words), nzero / 4);
}

template <typename K>
void launch_encode_int(const int32_t *pts, int64_t n, int dim, int bits, K *keys, uint32_t *hist,
                       int passes, hipStream_t s) {
  const int grid = grid_for(n, kBlock * 4);
  if (dim == 2) encode_int_kernel<K, 2><<<grid, kBlock, 0, s>>>(pts, n, bits, keys, hist, passes);
  else encode_int_kernel<K, 3><<<grid, kBlock, 0, s>>>(pts, n, bits, keys, hist, passes);
}

void launch_seg_init(SegState *ss, int64_t B, hipStream_t s) {
  const unsigned grid = (unsigned)((B + kBlock - 1) / kBlock);
  seg_init_kernel<<<grid ? grid : 1, kBlock, 0, s>>>(ss, B);
}

void launch_bbox_seg(const float *pts, int dim, const BatchDev &b, hipStream_t s) {
  if (b.n_chunk <= 0) return;
  if (dim == 2) bbox_seg_kernel<2><<<(unsigned)b.n_chunk, kBlock, 0, s>>>(pts, b.seg, b.chunk, b.ss);
  else bbox_seg_kernel<3><<<(unsigned)b.n_chunk, kBlock, 0, s>>>(pts, b.seg, b.chunk, b.ss);
}

template <typename K>
void launch_encode_seg(const void *pts, int quantized, int dim, int bits, float effw,
                       const BatchDev &b, K *keys, uint32_t *bhist, int passes, RunState *st,
                       hipStream_t s) {
  if (b.n_chunk <= 0) return;
  const unsigned g = (unsigned)b.n_chunk;
  const int db = b.digit_bits;
  if (quantized) {
    if (dim == 2) encode_seg_kernel<K, 2, true><<<g, kBlock, 0, s>>>(pts, b.seg, b.chunk, b.ss, bits, effw, keys, bhist, passes, db, st);
    else encode_seg_kernel<K, 3, true><<<g, kBlock, 0, s>>>(pts, b.seg, b.chunk, b.ss, bits, effw, keys, bhist, passes, db, st);
  } else {
    if (dim == 2) encode_seg_kernel<K, 2, false><<<g, kBlock, 0, s>>>(pts, b.seg, b.chunk, b.ss, bits, effw, keys, bhist, passes, db, st);
    else encode_seg_kernel<K, 3, false><<<g, kBlock, 0, s>>>(pts, b.seg, b.chunk, b.ss, bits, effw, keys, bhist, passes, db, st);
  }
}

#define PGVD_INST(K)                                                                          \
  template void launch_encode<K>(const float *, int64_t, int, int, float, int, const float *,  \
                                 const float *, K *, int64_t *, uint32_t *, int, RunState *,   \
                                 hipStream_t, uint32_t *, int, int);                           \
  template void launch_points_to_morton<K>(const void *, int, int64_t, int64_t, int, int, K *, \
                                           hipStream_t);                                       \
  template void launch_histogram<K>(const K *, int64_t, uint32_t *, int, hipStream_t,          \
                                    const RunState *, uint32_t *, size_t);                     \
  template void launch_encode_int<K>(const int32_t *, int64_t, int, int, K *, uint32_t *, int,  \
                                     hipStream_t);                                             \
  template void launch_encode_seg<K>(const void *, int, int, int, float, const BatchDev &, K *, \
                                     uint32_t *, int, RunState *, hipStream_t);
PGVD_INST(uint32_t)
PGVD_INST(uint64_t)
PGVD_INST(Key128)
#undef PGVD_INST

}  // namespace pgvd
