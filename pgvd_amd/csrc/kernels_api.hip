// kernels_api.hip -- C++ host layer (include/pgvd_kernels.hpp): the
// reference's namespace Kernels stage API (opencl/Kernels.h:26-50) and Karras
// facade (Karras.h:16-36) over the C ABI.  Host code only.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <istream>
#include <iterator>
#include <ostream>

#include "../../include/pgvd_kernels.hpp"

namespace pgvd {

namespace {

int key_bytes_for(int mbits) { return mbits <= 32 ? 4 : (mbits <= 64 ? 8 : 16); }
size_t brt_node_bytes(int kb) { return kb == 4 ? 12 : (kb == 8 ? 16 : 24); }

std::string err() { return pgvd_last_error(); }

}  // namespace

// ---------------------------------------------------------------- Resln ---
Resln make_resln(int64_t width, int dim) {
  if (width <= 0) throw std::logic_error("No support for width of 0");
  Resln r;
  r.width = width;
  r.bits = 0;
  int64_t w = width;
  while (!(w & 1)) {
    ++r.bits;
    w >>= 1;
  }
  r.mbits = r.bits * dim;
  return r;
}

// ------------------------------------------------------- Buffer/Context ---
Buffer::Mem::~Mem() {
  if (ptr && ctx) pgvd_free(ctx.get(), ptr);
}

Buffer Buffer::parent_links() const {
  Buffer b;
  b.mem_ = aux_;
  return b;
}

Context::Context(int device) {
  pgvd_ctx *c = nullptr;
  const int rc = pgvd_ctx_create(device, &c);
  if (rc != PGVD_OK) throw std::runtime_error("pgvd_ctx_create: " + err());
  ctx_ = std::shared_ptr<pgvd_ctx>(c, [](pgvd_ctx *p) { pgvd_ctx_destroy(p); });
}

Context::~Context() = default;

Context &Context::Default() {
  static Context ctx(0);
  return ctx;
}

Buffer Context::alloc(size_t bytes) {
  Buffer b;
  auto m = std::make_shared<Buffer::Mem>();
  m->ctx = ctx_;
  if (pgvd_alloc(ctx_.get(), bytes ? bytes : 1, &m->ptr) != PGVD_OK)
    throw std::runtime_error("pgvd_alloc: " + err());
  m->bytes = bytes;
  b.mem_ = m;
  return b;
}

Buffer Context::alloc_aux(Buffer &owner, size_t bytes) {
  if (!owner.aux_ || owner.aux_->bytes < bytes) {
    Buffer a = alloc(bytes);
    owner.aux_ = a.mem_;
  }
  return owner.parent_links();
}

int Context::get(Buffer &buffer, const std::string &key, size_t bytes) {
  for (auto &kv : pool_) {
    if (kv.first != key) continue;
    if (kv.second->bytes < bytes) {
      Buffer fresh = alloc(bytes);
      kv.second = fresh.mem_;
    }
    buffer.mem_ = kv.second;
    return PGVD_OK;
  }
  Buffer fresh = alloc(bytes);
  pool_.emplace_back(key, fresh.mem_);
  buffer.mem_ = fresh.mem_;
  return PGVD_OK;
}

int Context::upload(Buffer &dst, const void *src, size_t bytes) {
  return pgvd_memcpy_htod(ctx_.get(), dst.ptr(), src, bytes, nullptr);
}

int Context::download(void *dst, const Buffer &src, size_t bytes) {
  return pgvd_memcpy_dtoh(ctx_.get(), dst, src.ptr(), bytes, nullptr);
}

int Context::synchronize() { return pgvd_synchronize(ctx_.get(), nullptr); }

namespace Kernels {

namespace {
bool g_benchmarking = false;
std::string g_bench_name;
std::chrono::steady_clock::time_point g_bench_t0;
Context &ctx() { return Context::Default(); }
}  // namespace

// opencl/Kernels.cpp:5-18: wall time (the reference used clock(), CPU time)
// with the queue drained on both sides.
void setBenchmarking(bool on) { g_benchmarking = on; }
void startBenchmark(std::string benchmarkName) {
  if (!g_benchmarking) return;
  ctx().synchronize();
  g_bench_name = benchmarkName;
  g_bench_t0 = std::chrono::steady_clock::now();
}
void stopBenchmark() {
  if (!g_benchmarking) return;
  ctx().synchronize();
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - g_bench_t0).count();
  std::printf("%s: %.3f ms\n", g_bench_name.c_str(), ms);
}

int nextPow2(int num) { return pgvd_next_pow2(num); }

template <typename P>
static int upload_points(const std::vector<P> &points, Buffer &pointsBuffer, int dim) {
  startBenchmark("Uploading points");
  const int round = nextPow2((int)points.size());
  int rc = ctx().get(pointsBuffer, "pointsBuffer", sizeof(P) * (size_t)round);
  if (rc == PGVD_OK) rc = ctx().upload(pointsBuffer, points.data(), sizeof(P) * points.size());
  pointsBuffer.dim = dim;
  pointsBuffer.coord_bytes = 4;
  stopBenchmark();
  return rc;
}
int UploadPoints(const std::vector<int2> &points, Buffer &pointsBuffer) {
  return upload_points(points, pointsBuffer, 2);
}
int UploadPoints(const std::vector<int3> &points, Buffer &pointsBuffer) {
  return upload_points(points, pointsBuffer, 3);
}

int PointsToMorton_p(Buffer &points, Buffer &zpoints, int size, int bits) {
  if (points.dim != 2 && points.dim != 3) return PGVD_ERR_INVALID;
  const int kb = key_bytes_for(bits * points.dim);
  const int global = nextPow2(size);
  int rc = ctx().get(zpoints, "zpoints", (size_t)global * kb);
  if (rc != PGVD_OK) return rc;
  startBenchmark("PointsToMorton_p");
  rc = pgvd_points_to_morton(ctx().raw(), points.ptr(), points.coord_bytes ? points.coord_bytes : 4,
                             size, points.dim, bits, zpoints.ptr(), kb, nullptr);
  stopBenchmark();
  zpoints.key_bytes = kb;
  zpoints.dim = points.dim;
  return rc;
}

int BitPredicate(Buffer &input, Buffer &predicate, unsigned int &index, unsigned char compared,
                 int globalSize) {
  int rc = ctx().get(predicate, "predicate", sizeof(uint32_t) * (size_t)globalSize);
  if (rc != PGVD_OK) return rc;
  return pgvd_bit_predicate(ctx().raw(), input.ptr(), input.key_bytes,
                            static_cast<uint32_t *>(predicate.ptr()), (int)index, compared,
                            globalSize, nullptr);
}

int UniquePredicate(Buffer &input, Buffer &predicate, int globalSize) {
  if (predicate.bytes() < sizeof(uint32_t) * (size_t)globalSize) {
    const int rc = ctx().get(predicate, "predicate", sizeof(uint32_t) * (size_t)globalSize);
    if (rc != PGVD_OK) return rc;
  }
  return pgvd_unique_predicate(ctx().raw(), input.ptr(), input.key_bytes,
                               static_cast<uint32_t *>(predicate.ptr()), globalSize, nullptr);
}

int StreamScan_p(Buffer &input, Buffer &result, int globalSize) {
  if (result.bytes() < sizeof(uint32_t) * (size_t)globalSize) {
    const int rc = ctx().get(result, "scanResult", sizeof(uint32_t) * (size_t)globalSize);
    if (rc != PGVD_OK) return rc;
  }
  return pgvd_stream_scan(ctx().raw(), static_cast<const uint32_t *>(input.ptr()),
                          static_cast<uint32_t *>(result.ptr()), globalSize, nullptr);
}

static int compact(Buffer &input, Buffer &result, Buffer &predicate, Buffer &address,
                   int globalSize, bool dbl) {
  const int kb = input.key_bytes;
  if (result.bytes() < (size_t)kb * globalSize) {
    const int rc = ctx().get(result, dbl ? "doubleCompact" : "singleCompact", (size_t)kb * globalSize);
    if (rc != PGVD_OK) return rc;
  }
  result.key_bytes = kb;
  result.dim = input.dim;
  auto *p = static_cast<const uint32_t *>(predicate.ptr());
  auto *a = static_cast<const uint32_t *>(address.ptr());
  return dbl ? pgvd_double_compact(ctx().raw(), input.ptr(), result.ptr(), kb, p, a, globalSize, nullptr)
             : pgvd_single_compact(ctx().raw(), input.ptr(), result.ptr(), kb, p, a, globalSize, nullptr);
}
int SingleCompact(Buffer &input, Buffer &result, Buffer &predicate, Buffer &address,
                  int globalSize) {
  return compact(input, result, predicate, address, globalSize, false);
}
int DoubleCompact(Buffer &input, Buffer &result, Buffer &predicate, Buffer &address,
                  int globalSize) {
  return compact(input, result, predicate, address, globalSize, true);
}

// opencl/Kernels.cpp:198-217: unique over the padded array; rebinds input.
int UniqueSorted(Buffer &input, int &size) {
  startBenchmark("UniqueSorted");
  const int global = nextPow2(size);
  Buffer result;
  int rc = ctx().get(result, "result", (size_t)input.key_bytes * global);
  if (rc != PGVD_OK) return rc;
  int64_t m = 0;
  rc = pgvd_unique_sorted(ctx().raw(), input.ptr(), result.ptr(), global, input.key_bytes, &m,
                          nullptr);
  result.key_bytes = input.key_bytes;
  result.dim = input.dim;
  input = result;
  size = (int)m;
  stopBenchmark();
  return rc;
}

// opencl/Kernels.cpp:219-247: sorts the nextPow2(size) keys of the buffer.
int RadixSortBigUnsigned(Buffer &input, int size, int mbits) {
  const int kb = input.key_bytes;
  if (kb != 4 && kb != 8 && kb != 16) return PGVD_ERR_INVALID;
  const int bits = mbits < kb * 8 ? mbits : kb * 8;   // bits past the key width are zero
  startBenchmark("RadixSortBigUnsigned");
  const int rc = pgvd_radix_sort(ctx().raw(), input.ptr(), nullptr, nextPow2(size), kb, bits, nullptr);
  stopBenchmark();
  return rc;
}

int BuildBinaryRadixTree_p(Buffer &zpoints, Buffer &internalBRTNodes, int size, int mbits) {
  startBenchmark("BuildBinaryRadixTree_p");
  const int kb = zpoints.key_bytes;
  const int global = nextPow2(size);
  int rc = ctx().get(internalBRTNodes, "internalBRTNodes", brt_node_bytes(kb) * (size_t)global);
  if (rc != PGVD_OK) return rc;
  Buffer parent = ctx().alloc_aux(internalBRTNodes, sizeof(int32_t) * (size_t)global);
  rc = pgvd_build_brt(ctx().raw(), zpoints.ptr(), size, kb, mbits, internalBRTNodes.ptr(),
                      static_cast<int32_t *>(parent.ptr()), nullptr);
  internalBRTNodes.key_bytes = kb;
  internalBRTNodes.dim = zpoints.dim;
  stopBenchmark();
  return rc;
}

int ComputeLocalSplits_p(Buffer &internalBRTNodes, Buffer &localSplits, int size) {
  startBenchmark("ComputeLocalSplits_p");
  int rc = ctx().get(localSplits, "localSplits", sizeof(uint32_t) * (size_t)nextPow2(size));
  if (rc != PGVD_OK) return rc;
  rc = pgvd_local_splits(ctx().raw(), internalBRTNodes.ptr(), size, internalBRTNodes.key_bytes,
                         internalBRTNodes.dim, static_cast<uint32_t *>(localSplits.ptr()), nullptr);
  stopBenchmark();
  return rc;
}

int InitOctree(Buffer &internalBRTNodes, Buffer &octree, Buffer &, Buffer &, int, int octreeSize) {
  startBenchmark("InitOctree");
  const int dim = internalBRTNodes.dim;
  const size_t stride = sizeof(int32_t) * ((1u << dim) + 1);
  if (octree.bytes() < stride * (size_t)octreeSize) {
    const int rc = ctx().get(octree, "octree", stride * (size_t)nextPow2(octreeSize));
    if (rc != PGVD_OK) return rc;
  }
  const int rc = pgvd_init_octree(ctx().raw(), octree.ptr(), octreeSize, dim, nullptr);
  stopBenchmark();
  return rc;
}

template <int DIM>
static int brt_to_octree(Buffer &I, std::vector<OctNodeT<DIM>> &octree_vec, int size) {
  if (I.dim != DIM) return PGVD_ERR_INVALID;
  startBenchmark("BinaryRadixToOctree_p");
  int64_t cap = 2 * (int64_t)size + 16, osz = 0;
  Buffer oct;
  int rc;
  for (;;) {
    rc = ctx().get(oct, "octree", sizeof(OctNodeT<DIM>) * (size_t)cap);
    if (rc != PGVD_OK) return rc;
    rc = pgvd_brt_to_octree(ctx().raw(), I.ptr(), static_cast<const int32_t *>(I.parent_links().ptr()),
                            size, I.key_bytes, DIM, oct.ptr(), cap, &osz, nullptr);
    if (rc != PGVD_ERR_CAPACITY) break;
    cap = osz + osz / 8 + 16;
  }
  if (rc != PGVD_OK) return rc;
  octree_vec.resize((size_t)osz);
  rc = ctx().download(octree_vec.data(), oct, sizeof(OctNodeT<DIM>) * (size_t)osz);
  stopBenchmark();
  return rc;
}
int BinaryRadixToOctree_p(Buffer &I, std::vector<OctNode2> &octree_vec, int size) {
  return brt_to_octree<2>(I, octree_vec, size);
}
int BinaryRadixToOctree_p(Buffer &I, std::vector<OctNode3> &octree_vec, int size) {
  return brt_to_octree<3>(I, octree_vec, size);
}

// opencl/Kernels.cpp:413-429, all on the device.
template <typename P, int DIM>
static int build_octree(const std::vector<P> &points, std::vector<OctNodeT<DIM>> &octree,
                        int bits, int mbits) {
  if (points.empty()) throw std::logic_error("Zero points not supported");
  if (mbits != bits * DIM) return PGVD_ERR_INVALID;
  Buffer pts;
  int rc = upload_points(points, pts, DIM);
  if (rc != PGVD_OK) return rc;
  pgvd_build_opts opts;
  std::memset(&opts, 0, sizeof(opts));
  opts.quantized_input = 1;
  rc = pgvd_build_octree(ctx().raw(), pts.ptr(), (int64_t)points.size(), DIM, bits, nullptr, &opts,
                         nullptr);
  if (rc != PGVD_OK) return rc;
  pgvd_octree_result res;
  rc = pgvd_build_octree_finish(ctx().raw(), &res);
  if (rc != PGVD_OK) return rc;
  octree.resize((size_t)res.octree_size);
  return pgvd_memcpy_dtoh(ctx().raw(), octree.data(), res.octree,
                          sizeof(OctNodeT<DIM>) * (size_t)res.octree_size, nullptr);
}
int BuildOctree_p(const std::vector<int2> &points, std::vector<OctNode2> &octree, int bits,
                  int mbits) {
  return build_octree<int2, 2>(points, octree, bits, mbits);
}
int BuildOctree_p(const std::vector<int3> &points, std::vector<OctNode3> &octree, int bits,
                  int mbits) {
  return build_octree<int3, 3>(points, octree, bits, mbits);
}

int ReadKeys(const Buffer &keys, int count, std::vector<uint64_t> &lo, std::vector<uint64_t> &hi) {
  const int kb = keys.key_bytes;
  std::vector<unsigned char> raw((size_t)kb * count);
  const int rc = ctx().download(raw.data(), keys, raw.size());
  lo.assign(count, 0);
  hi.assign(count, 0);
  for (int i = 0; i < count; ++i) {
    if (kb == 4) { uint32_t v; std::memcpy(&v, &raw[(size_t)4 * i], 4); lo[i] = v; }
    else if (kb == 8) std::memcpy(&lo[i], &raw[(size_t)8 * i], 8);
    else { std::memcpy(&lo[i], &raw[(size_t)16 * i], 8); std::memcpy(&hi[i], &raw[(size_t)16 * i + 8], 8); }
  }
  return rc;
}

int ReadBigUnsigned(const Buffer &keys, int count, std::vector<BigUnsigned> &out) {
  const int kb = keys.key_bytes;
  std::vector<unsigned char> raw((size_t)kb * (size_t)(count > 0 ? count : 0));
  int rc = ctx().download(raw.data(), keys, raw.size());
  out.assign((size_t)(count > 0 ? count : 0), BigUnsigned{});
  if (rc == PGVD_OK && count > 0) rc = pgvd_keys_to_big_unsigned(raw.data(), kb, count, out.data());
  return rc;
}

int WriteBigUnsigned(const std::vector<BigUnsigned> &in, int key_bytes, Buffer &keys) {
  std::vector<unsigned char> raw((size_t)key_bytes * in.size());
  int rc = pgvd_big_unsigned_to_keys(in.data(), (int64_t)in.size(), key_bytes, raw.data());
  if (rc != PGVD_OK) return rc;
  if (keys.bytes() < raw.size()) keys = ctx().alloc(raw.size());
  keys.key_bytes = key_bytes;
  return ctx().upload(keys, raw.data(), raw.size());
}

int ReadBrtNodes(const Buffer &I, int size, std::vector<RefBrtNode> &out) {
  const int kb = I.key_bytes;
  const size_t nb = brt_node_bytes(kb);
  const int n = size > 1 ? size - 1 : 0;
  std::vector<unsigned char> raw(nb * (size_t)n);
  std::vector<int32_t> parent((size_t)n);
  int rc = ctx().download(raw.data(), I, raw.size());
  if (rc == PGVD_OK) rc = ctx().download(parent.data(), I.parent_links(), sizeof(int32_t) * (size_t)n);
  out.assign((size_t)n, RefBrtNode{});
  if (rc == PGVD_OK && n > 0) rc = pgvd_brt_to_ref_nodes(raw.data(), parent.data(), n, kb, out.data());
  return rc;
}

int ReadBrt(const Buffer &I, int size, std::vector<BrtNode> &nodes) {
  const int kb = I.key_bytes;
  const size_t nb = brt_node_bytes(kb);
  const int n = size > 1 ? size - 1 : 0;
  std::vector<unsigned char> raw(nb * (size_t)n);
  std::vector<int32_t> parent((size_t)n);
  int rc = ctx().download(raw.data(), I, raw.size());
  if (rc == PGVD_OK) rc = ctx().download(parent.data(), I.parent_links(), sizeof(int32_t) * (size_t)n);
  nodes.resize((size_t)n);
  for (int i = 0; i < n; ++i) {
    const unsigned char *r = &raw[nb * i];
    BrtNode &b = nodes[i];
    b.lcp_lo = b.lcp_hi = 0;
    size_t off;
    if (kb == 4) { uint32_t v; std::memcpy(&v, r, 4); b.lcp_lo = v; off = 4; }
    else if (kb == 8) { std::memcpy(&b.lcp_lo, r, 8); off = 8; }
    else { std::memcpy(&b.lcp_lo, r, 8); std::memcpy(&b.lcp_hi, r + 8, 8); off = 16; }
    uint16_t len;
    std::memcpy(&b.left, r + off, 4);
    std::memcpy(&len, r + off + 4, 2);
    b.left_leaf = r[off + 6] != 0;
    b.right_leaf = r[off + 7] != 0;
    b.lcp_length = len;
    b.parent = parent[i];
  }
  return rc;
}

}  // namespace Kernels

namespace Karras {

template <typename F, typename I, int DIM>
static std::vector<I> quantize(const std::vector<F> &points, const Resln &resln, const F *bbMin,
                               const F *bbMax, bool clamped) {
  std::vector<I> out;
  if (points.empty()) return out;
  if (resln.bits > 31) throw std::logic_error("int grid points need bits <= 31");
  Context &c = Context::Default();
  Buffer pts = c.alloc(sizeof(F) * points.size());
  c.upload(pts, points.data(), sizeof(F) * points.size());
  Buffer q = c.alloc(sizeof(int64_t) * DIM * points.size());
  float bb[6];
  const float *custom = nullptr;
  if (bbMin && bbMax) {
    std::memcpy(bb, bbMin, sizeof(F));
    std::memcpy(bb + DIM, bbMax, sizeof(F));
    custom = bb;
  }
  int64_t nout = 0;
  const int rc = pgvd_quantize(c.raw(), static_cast<const float *>(pts.ptr()), (int64_t)points.size(),
                               DIM, resln.bits, custom, clamped ? 1 : 0,
                               static_cast<int64_t *>(q.ptr()), &nout, nullptr);
  if (rc == PGVD_ERR_BAD_COORD) throw std::logic_error("bad coordinate");
  if (rc != PGVD_OK) throw std::runtime_error("pgvd_quantize: " + std::string(pgvd_last_error()));
  std::vector<int64_t> q64((size_t)nout * DIM);
  c.download(q64.data(), q, sizeof(int64_t) * q64.size());
  out.resize((size_t)nout);
  for (int64_t i = 0; i < nout; ++i) {
    int32_t v[DIM];
    for (int k = 0; k < DIM; ++k) v[k] = (int32_t)q64[(size_t)i * DIM + k];
    std::memcpy(&out[(size_t)i], v, sizeof(I));
  }
  return out;
}

std::vector<int2> Quantize(const std::vector<float2> &points, const Resln &resln,
                           const float2 *bbMin, const float2 *bbMax, bool clamped) {
  return quantize<float2, int2, 2>(points, resln, bbMin, bbMax, clamped);
}
std::vector<int3> Quantize(const std::vector<float3> &points, const Resln &resln,
                           const float3 *bbMin, const float3 *bbMax, bool clamped) {
  return quantize<float3, int3, 3>(points, resln, bbMin, bbMax, clamped);
}

std::vector<OctNode2> BuildOctreeInParallel(const std::vector<int2> &points, const Resln &r, bool) {
  std::vector<OctNode2> octree;
  const int rc = Kernels::BuildOctree_p(points, octree, r.bits, r.mbits);
  if (rc != PGVD_OK) throw std::runtime_error("BuildOctree_p: " + std::string(pgvd_last_error()));
  return octree;
}
std::vector<OctNode3> BuildOctreeInParallel(const std::vector<int3> &points, const Resln &r, bool) {
  std::vector<OctNode3> octree;
  const int rc = Kernels::BuildOctree_p(points, octree, r.bits, r.mbits);
  if (rc != PGVD_OK) throw std::runtime_error("BuildOctree_p: " + std::string(pgvd_last_error()));
  return octree;
}

std::vector<OctNode3> BuildOctreeFromPoints(const std::vector<float3> &points, const Resln &r) {
  if (points.empty()) throw std::logic_error("Zero points not supported");
  Context &c = Context::Default();
  Buffer pts = c.alloc(sizeof(float3) * points.size());
  c.upload(pts, points.data(), sizeof(float3) * points.size());
  int rc = pgvd_build_octree(c.raw(), pts.ptr(), (int64_t)points.size(), 3, r.bits, nullptr,
                             nullptr, nullptr);
  pgvd_octree_result res;
  if (rc == PGVD_OK) rc = pgvd_build_octree_finish(c.raw(), &res);
  if (rc == PGVD_ERR_BAD_COORD) throw std::logic_error("bad coordinate");
  if (rc != PGVD_OK) throw std::runtime_error("pgvd_build_octree: " + std::string(pgvd_last_error()));
  std::vector<OctNode3> octree((size_t)res.octree_size);
  pgvd_memcpy_dtoh(c.raw(), octree.data(), res.octree, sizeof(OctNode3) * octree.size(), nullptr);
  return octree;
}

}  // namespace Karras

// ------------------------------------------------------------ OctreeUtils ---
namespace OctreeUtils {

namespace {

void check(int rc, const char *what) {
  if (rc != PGVD_OK) throw std::runtime_error(std::string(what) + ": " + pgvd_last_error());
}

template <int DIM, typename P>
std::vector<OctCellT<DIM>> find_leaves(const std::vector<P> &points,
                                       const std::vector<OctNodeT<DIM>> &octree,
                                       const Resln &resln) {
  std::vector<OctCellT<DIM>> cells(points.size());
  if (points.empty()) return cells;
  if (octree.empty()) throw std::logic_error("FindLeaf on an empty octree");
  Context &c = Context::Default();
  const size_t q = points.size();
  Buffer dp = c.alloc(sizeof(P) * q), doct = c.alloc(sizeof(OctNodeT<DIM>) * octree.size());
  Buffer out = c.alloc(5 * 4 * q);
  c.upload(dp, points.data(), sizeof(P) * q);
  c.upload(doct, octree.data(), sizeof(OctNodeT<DIM>) * octree.size());
  int32_t *o = static_cast<int32_t *>(out.ptr());
  check(pgvd_find_leaf(c.raw(), static_cast<const int32_t *>(dp.ptr()), (int64_t)q, DIM,
                       resln.bits, static_cast<const int32_t *>(doct.ptr()),
                       (int64_t)octree.size(), o, o + q, o + 2 * q, o + 3 * q,
                       reinterpret_cast<uint32_t *>(o + 4 * q), nullptr),
        "pgvd_find_leaf");
  std::vector<int32_t> h(5 * q);
  c.download(h.data(), out, h.size() * 4);
  for (size_t i = 0; i < q; ++i) {
    if (h[4 * q + i] != 0) throw std::logic_error("FindLeaf: no leaf reached");
    const int32_t *pc = reinterpret_cast<const int32_t *>(&points[i]);
    const int32_t w = (int32_t)(1u << (resln.bits - h[3 * q + i]));
    // the octant offsets along the descent (OctreeUtils.cpp:41-47): the
    // coordinate bits the key holds, above the cell's width
    const int32_t gmask = (int32_t)((1u << resln.bits) - 1u);
    for (int k = 0; k < DIM; ++k) cells[i].origin[k] = (pc[k] & gmask) & ~(w - 1);
    cells[i].width = w;
    cells[i].parent_idx = h[i];
    cells[i].octant = h[q + i];
    cells[i].data = h[2 * q + i];
  }
  return cells;
}

}  // namespace

OctCell2 FindLeaf(const int2 &p, const std::vector<OctNode2> &octree, const Resln &resln) {
  return find_leaves<2>(std::vector<int2>{p}, octree, resln)[0];
}
OctCell3 FindLeaf(const int3 &p, const std::vector<OctNode3> &octree, const Resln &resln) {
  return find_leaves<3>(std::vector<int3>{p}, octree, resln)[0];
}
std::vector<OctCell2> FindLeaves(const std::vector<int2> &points,
                                 const std::vector<OctNode2> &octree, const Resln &resln) {
  return find_leaves<2>(points, octree, resln);
}
std::vector<OctCell3> FindLeaves(const std::vector<int3> &points,
                                 const std::vector<OctNode3> &octree, const Resln &resln) {
  return find_leaves<3>(points, octree, resln);
}

namespace {

// The walks and each one's status (1: past 10000 cells).
std::vector<std::vector<OctCell2>> walks(const std::vector<float2> &a,
                                         const std::vector<float2> &b,
                                         const std::vector<OctNode2> &octree, const Resln &resln,
                                         std::vector<int32_t> &st) {
  if (a.size() != b.size()) throw std::logic_error("CellWalks: a and b differ in length");
  const size_t n = a.size();
  std::vector<std::vector<OctCell2>> out(n);
  if (n == 0) return out;
  if (octree.empty()) throw std::logic_error("CellWalk on an empty octree");
  Context &c = Context::Default();
  std::vector<float> segs(4 * n);
  for (size_t i = 0; i < n; ++i) {
    segs[4 * i] = a[i].x;
    segs[4 * i + 1] = a[i].y;
    segs[4 * i + 2] = b[i].x;
    segs[4 * i + 3] = b[i].y;
  }
  Buffer ds = c.alloc(4 * segs.size()), doct = c.alloc(sizeof(OctNode2) * octree.size());
  Buffer doff = c.alloc(8 * (n + 1)), dst = c.alloc(4 * n);
  c.upload(ds, segs.data(), 4 * segs.size());
  c.upload(doct, octree.data(), sizeof(OctNode2) * octree.size());
  int64_t total = 0;
  const auto *sp = static_cast<const float *>(ds.ptr());
  const auto *op = static_cast<const int32_t *>(doct.ptr());
  auto *offp = static_cast<int64_t *>(doff.ptr());
  check(pgvd_cell_walk(c.raw(), sp, (int64_t)n, resln.bits, op, (int64_t)octree.size(), offp,
                       nullptr, 0, &total, static_cast<int32_t *>(dst.ptr()), nullptr),
        "pgvd_cell_walk");
  Buffer dc = c.alloc(20 * (size_t)std::max<int64_t>(total, 1));
  check(pgvd_cell_walk(c.raw(), sp, (int64_t)n, resln.bits, op, (int64_t)octree.size(), offp,
                       static_cast<int32_t *>(dc.ptr()), total, &total,
                       static_cast<int32_t *>(dst.ptr()), nullptr),
        "pgvd_cell_walk");
  std::vector<int64_t> off(n + 1);
  std::vector<int32_t> cells(5 * (size_t)total);
  st.assign(n, 0);
  c.download(off.data(), doff, 8 * (n + 1));
  c.download(st.data(), dst, 4 * n);
  if (total) c.download(cells.data(), dc, 4 * cells.size());
  for (size_t i = 0; i < n; ++i) {
    for (int64_t k = off[i]; k < off[i + 1]; ++k) {
      const int32_t *r = &cells[5 * k];
      OctCell2 cell;
      cell.origin[0] = r[0];
      cell.origin[1] = r[1];
      cell.width = r[2];
      cell.parent_idx = r[3];
      cell.octant = r[4];
      cell.data = octree[r[3]].children[r[4]];
      out[i].push_back(cell);
    }
  }
  return out;
}

}  // namespace

std::vector<std::vector<OctCell2>> CellWalks(const std::vector<float2> &a,
                                             const std::vector<float2> &b,
                                             const std::vector<OctNode2> &octree,
                                             const Resln &resln) {
  std::vector<int32_t> st;
  auto out = walks(a, b, octree, resln, st);
  for (int32_t x : st)
    if (x) throw std::logic_error("Infinite loop");
  return out;
}

void CellWalk(const float2 &a, const float2 &b, const std::vector<OctNode2> &octree,
              const Resln &resln, cwv v, void *data) {
  std::vector<int32_t> st;
  const auto w = walks(std::vector<float2>{a}, std::vector<float2>{b}, octree, resln, st);
  for (const OctCell2 &cell : w[0]) v(cell, a, b, octree, resln, data);
  if (st[0]) throw std::logic_error("Infinite loop");   // viewer/Octree2.cpp:286-287
}

}  // namespace OctreeUtils

// ------------------------------------------------------ GVD refinement ---
static_assert(sizeof(CellIntersections) == sizeof(pgvd_cell_intersections),
              "CellIntersections has the layout of pgvd_cell_intersections");

namespace {

void gvd_check(int rc, const char *what) {
  if (rc == PGVD_ERR_GEOMETRY || rc == PGVD_ERR_BAD_COORD)
    throw std::logic_error(std::string(what) + ": " + pgvd_last_error());
  if (rc != PGVD_OK) throw std::runtime_error(std::string(what) + ": " + pgvd_last_error());
}

}  // namespace

namespace Geom {

void FitBoxes(const FloatSegment &A, const FloatSegment &B, float min_d,
              std::vector<float2> *samples, std::vector<float2> *origins,
              std::vector<float> *lengths) {
  Context &c = Context::Default();
  const float pair[8] = {A.a.x, A.a.y, A.b.x, A.b.y, B.a.x, B.a.y, B.b.x, B.b.y};
  Buffer dp = c.alloc(sizeof pair);
  c.upload(dp, pair, sizeof pair);
  const auto *pp = static_cast<const float *>(dp.ptr());
  int64_t ns = 0, nb = 0;
  const int rc = pgvd_fit_boxes(c.raw(), pp, 1, min_d, nullptr, nullptr, nullptr, 0, nullptr,
                                nullptr, &ns, &nb, nullptr, nullptr);
  if (rc != PGVD_OK && rc != PGVD_ERR_CAPACITY && rc != PGVD_ERR_GEOMETRY)
    gvd_check(rc, "Geom::FitBoxes");
  const int64_t cap = std::max<int64_t>(std::max(ns, nb), 1);
  Buffer ds = c.alloc(8 * (size_t)cap), dorg = c.alloc(8 * (size_t)cap), dl = c.alloc(4 * (size_t)cap);
  const int rc2 = pgvd_fit_boxes(c.raw(), pp, 1, min_d, static_cast<float *>(ds.ptr()),
                                 static_cast<float *>(dorg.ptr()), static_cast<float *>(dl.ptr()),
                                 cap, nullptr, nullptr, &ns, &nb, nullptr, nullptr);
  // the reference pushes what it found before it throws
  std::vector<float2> s((size_t)ns), o((size_t)nb);
  std::vector<float> l((size_t)nb);
  if (ns) c.download(s.data(), ds, 8 * (size_t)ns);
  if (nb) {
    c.download(o.data(), dorg, 8 * (size_t)nb);
    c.download(l.data(), dl, 4 * (size_t)nb);
  }
  if (samples) samples->insert(samples->end(), s.begin(), s.end());
  if (origins) origins->insert(origins->end(), o.begin(), o.end());
  if (lengths) lengths->insert(lengths->end(), l.begin(), l.end());
  gvd_check(rc2, "Geom::FitBoxes");
}

}  // namespace Geom

Octree2::Octree2(const Resln &resln, int karras_iterations, float min_d)
    : resln_(resln), karras_iterations_(karras_iterations), min_d_(min_d) {}

void Octree2::build(const std::vector<std::vector<float2>> &polygons, const float2 *bbMin,
                    const float2 *bbMax) {
  octree_.clear();
  ci_.clear();
  qpoints_.clear();
  extra_.clear();
  iterations_ = 0;
  std::vector<float2> verts;
  std::vector<int64_t> counts;
  for (const auto &p : polygons) {
    verts.insert(verts.end(), p.begin(), p.end());
    counts.push_back((int64_t)p.size());
  }
  if (verts.empty()) return;               // Octree2.cpp:152-155
  Context &c = Context::Default();
  Buffer dv = c.alloc(sizeof(float2) * verts.size());
  c.upload(dv, verts.data(), sizeof(float2) * verts.size());
  float bb[4];
  const float *bbp = nullptr;
  if (bbMin && bbMax) {
    bb[0] = bbMin->x;
    bb[1] = bbMin->y;
    bb[2] = bbMax->x;
    bb[3] = bbMax->y;
    bbp = bb;
  }
  pgvd_gvd_result r;
  gvd_check(pgvd_gvd_refine(c.raw(), static_cast<const float *>(dv.ptr()), counts.data(),
                            (int64_t)counts.size(), resln_.bits, bbp, karras_iterations_, min_d_,
                            nullptr, &r, nullptr),
            "Octree2::build");
  iterations_ = r.iterations;
  qpoints_.resize((size_t)r.n_qpoints);
  gvd_check(pgvd_memcpy_dtoh(c.raw(), qpoints_.data(), r.qpoints, 8 * qpoints_.size(), nullptr),
        "download");
  extra_.resize((size_t)r.n_extra);
  if (r.n_extra)
    gvd_check(pgvd_memcpy_dtoh(c.raw(), extra_.data(), r.extra, 8 * extra_.size(), nullptr), "download");
  octree_.resize((size_t)r.octree_size);
  ci_.resize((size_t)r.octree_size);
  if (r.octree_size) {
    gvd_check(pgvd_memcpy_dtoh(c.raw(), octree_.data(), r.octree, sizeof(OctNode2) * octree_.size(),
                           nullptr),
          "download");
    gvd_check(pgvd_memcpy_dtoh(c.raw(), ci_.data(), r.ci, sizeof(CellIntersections) * ci_.size(),
                           nullptr),
          "download");
  }
}

// ------------------------------------------------- octree stream operators ---
namespace {

template <int DIM>
void write_octree(std::ostream &out, const std::vector<OctNodeT<DIM>> &octree) {
  Context &c = Context::Default();
  const size_t bytes = sizeof(OctNodeT<DIM>) * octree.size();
  Buffer d = c.alloc(std::max<size_t>(bytes, 4));
  if (bytes) c.upload(d, octree.data(), bytes);
  int64_t len = 0;
  const auto *p = static_cast<const int32_t *>(d.ptr());
  int rc = pgvd_octree_format_text(c.raw(), p, (int64_t)octree.size(), DIM, nullptr, 0, &len,
                                   nullptr);
  if (rc != PGVD_OK) throw std::runtime_error(std::string("format: ") + pgvd_last_error());
  Buffer t = c.alloc((size_t)len);
  rc = pgvd_octree_format_text(c.raw(), p, (int64_t)octree.size(), DIM,
                               static_cast<char *>(t.ptr()), len, &len, nullptr);
  if (rc != PGVD_OK) throw std::runtime_error(std::string("format: ") + pgvd_last_error());
  std::string text((size_t)len, ' ');
  c.download(&text[0], t, (size_t)len);
  out.write(text.data(), (std::streamsize)text.size());
}

template <int DIM>
void read_octree(std::istream &in, std::vector<OctNodeT<DIM>> &octree) {
  const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  if (text.empty()) {
    in.setstate(std::ios::failbit);
    return;
  }
  Context &c = Context::Default();
  Buffer t = c.alloc(text.size());
  c.upload(t, text.data(), text.size());
  const int64_t cap = (int64_t)(text.size() / (2 * ((1 << DIM) + 1)) + 1);
  Buffer d = c.alloc(sizeof(OctNodeT<DIM>) * (size_t)cap);
  int64_t O = 0;
  const int rc = pgvd_octree_parse_text(c.raw(), static_cast<const char *>(t.ptr()),
                                        (int64_t)text.size(), DIM, static_cast<int32_t *>(d.ptr()),
                                        cap, &O, nullptr);
  if (rc != PGVD_OK) {
    in.setstate(std::ios::failbit);
    return;
  }
  octree.resize((size_t)O);
  if (O) c.download(octree.data(), d, sizeof(OctNodeT<DIM>) * (size_t)O);
}

}  // namespace

std::ostream &operator<<(std::ostream &out, const std::vector<OctNode2> &octree) {
  write_octree<2>(out, octree);
  return out;
}
std::ostream &operator<<(std::ostream &out, const std::vector<OctNode3> &octree) {
  write_octree<3>(out, octree);
  return out;
}
std::istream &operator>>(std::istream &in, std::vector<OctNode2> &octree) {
  read_octree<2>(in, octree);
  return in;
}
std::istream &operator>>(std::istream &in, std::vector<OctNode3> &octree) {
  read_octree<3>(in, octree);
  return in;
}

}  // namespace pgvd
