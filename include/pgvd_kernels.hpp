// pgvd_kernels.hpp -- C++ host layer of libpgvd.so with the reference's stage
// API: namespace Kernels (opencl/Kernels.h:26-50) and the Karras facade
// (Karras.h:16-36, Resln.h:29-83), implemented over the C ABI in pgvd.h.
//
// Differences from the reference's declarations, each forced by the
// platform rather than chosen:
//   * cl::Buffer -> pgvd::Buffer, a shared handle to device memory (copies
//     alias, as cl::Buffer copies do; UniqueSorted and RadixSortBigUnsigned
//     rebind the caller's handle as the reference does).
//   * cl_int -> int, 0 on success, else a pgvd_status code; std::logic_error
//     where the reference throws (empty input, bad coordinate).
//   * DIM is a property of the data (2 or 3), not a compile-time macro:
//     points are pgvd::int2 / pgvd::int3, octree nodes OctNodeT<2> / <3>.
//   * Keys are 32/64/128-bit words picked from mbits instead of the 11-byte
//     BigUnsigned; a key buffer records its width.  ReadBigUnsigned,
//     WriteBigUnsigned and ReadBrtNodes convert to and from the reference's
//     BigUnsigned and BrtNode records.
// The serial `_s` functions of the reference are the test oracle
// (oracle/pgvd_oracle.h), not part of this library.
#pragma once

#include <cstdint>
#include <iosfwd>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pgvd.h"

namespace pgvd {

struct int2 { int32_t x, y; };
struct int3 { int32_t x, y, z; };
struct float2 { float x, y; };
struct float3 { float x, y, z; };

// OctNode (C/OctNode.h:16-26).  leaf bit o is set when children[o] == -1.
template <int DIM>
struct OctNodeT {
  int32_t children[1 << DIM];
  int32_t leaf;
  const int32_t &operator[](int i) const { return children[i]; }
};
using OctNode2 = OctNodeT<2>;
using OctNode3 = OctNodeT<3>;

// BrtNode (C/BrtNode.h:10-21) on the host, with a 128-bit lcp in place of
// the BigUnsigned.
struct BrtNode {
  int32_t left;
  bool left_leaf, right_leaf;
  uint64_t lcp_lo, lcp_hi;
  int32_t lcp_length;
  int32_t parent;
};

// The reference's own records, byte for byte: BigUnsigned
// (C/BigUnsigned.h:8-16, 16 bytes: len, then 11 8-bit blocks least
// significant first, no zero block at len - 1) and BrtNode
// (C/BrtNode.h:10-21, 32 bytes).  Kernels::ReadBigUnsigned / ReadBrtNodes
// read device results into them, as the reference's tests read their
// buffers (tests/KernelTests.cpp:74,130,190); CompareBU is compareBU.
using BigUnsigned = pgvd_big_unsigned;
using RefBrtNode = pgvd_ref_brt_node;
inline int CompareBU(const BigUnsigned &x, const BigUnsigned &y) {
  return pgvd_compare_big_unsigned(&x, &y);
}

// Resln (Resln.h:29-60) and make_resln (:64-83), with 64-bit width so that
// grids finer than 2^30 per axis are expressible.
struct Resln {
  int64_t width;
  int bits;
  int mbits;
};
Resln make_resln(int64_t width, int dim);

class Context;

// Device memory handle standing in for cl::Buffer.
class Buffer {
 public:
  Buffer() = default;
  void *ptr() const { return mem_ ? mem_->ptr : nullptr; }
  size_t bytes() const { return mem_ ? mem_->bytes : 0; }
  explicit operator bool() const { return ptr() != nullptr; }
  // What the buffer holds, recorded by the stage that filled it.
  int key_bytes = 0;   // key buffers: 4, 8 or 16
  int dim = 0;         // point / key / BRT buffers: 2 or 3
  int coord_bytes = 0; // point buffers: 4 (int32 coordinates)
  Buffer parent_links() const;   // BRT buffers: the int32 parent array

 private:
  friend class Context;
  struct Mem {
    std::shared_ptr<pgvd_ctx> ctx;   // keeps the context alive while memory is
    void *ptr = nullptr;
    size_t bytes = 0;
    ~Mem();
  };
  std::shared_ptr<Mem> mem_;
  std::shared_ptr<Mem> aux_;   // BRT parent links
};

// One device context (HIP stream + keyed buffer pool), the role of CLFW's
// static state (CLFW/include/clfw.hpp:21-79).
class Context {
 public:
  explicit Context(int device = 0);
  ~Context();
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  pgvd_ctx *raw() const { return ctx_.get(); }
  // CLFW::get(buffer, key, size) (CLFW/src/clfw.cpp:212-226): reuse the
  // pooled buffer under `key` when it is large enough, else reallocate.
  int get(Buffer &buffer, const std::string &key, size_t bytes);
  Buffer alloc(size_t bytes);
  Buffer alloc_aux(Buffer &owner, size_t bytes);
  int upload(Buffer &dst, const void *src, size_t bytes);
  int download(void *dst, const Buffer &src, size_t bytes);
  int synchronize();
  static Context &Default();       // device 0, created on first use

 private:
  std::shared_ptr<pgvd_ctx> ctx_;
  std::vector<std::pair<std::string, std::shared_ptr<Buffer::Mem>>> pool_;
};

namespace Kernels {

void startBenchmark(std::string benchmarkName);
void stopBenchmark();
void setBenchmarking(bool on);

int nextPow2(int num);
int UploadPoints(const std::vector<int2> &points, Buffer &pointsBuffer);
int UploadPoints(const std::vector<int3> &points, Buffer &pointsBuffer);
int PointsToMorton_p(Buffer &points, Buffer &zpoints, int size, int bits);
int BitPredicate(Buffer &input, Buffer &predicate, unsigned int &index, unsigned char compared,
                 int globalSize);
int UniquePredicate(Buffer &input, Buffer &predicate, int globalSize);
int StreamScan_p(Buffer &input, Buffer &result, int globalSize);
int SingleCompact(Buffer &input, Buffer &result, Buffer &predicate, Buffer &address,
                  int globalSize);
int DoubleCompact(Buffer &input, Buffer &result, Buffer &predicate, Buffer &address,
                  int globalSize);
int UniqueSorted(Buffer &input, int &size);
int RadixSortBigUnsigned(Buffer &input, int size, int mbits);
int BuildBinaryRadixTree_p(Buffer &zpoints, Buffer &internalBRTNodes, int size, int mbits);
int ComputeLocalSplits_p(Buffer &internalBRTNodes, Buffer &localSplits, int size);
int InitOctree(Buffer &internalBRTNodes, Buffer &octree, Buffer &localSplits,
               Buffer &scannedSplits, int size, int octreeSize);
int BinaryRadixToOctree_p(Buffer &internalBRTNodes, std::vector<OctNode2> &octree_vec, int size);
int BinaryRadixToOctree_p(Buffer &internalBRTNodes, std::vector<OctNode3> &octree_vec, int size);
int BuildOctree_p(const std::vector<int2> &points, std::vector<OctNode2> &octree, int bits,
                  int mbits);
int BuildOctree_p(const std::vector<int3> &points, std::vector<OctNode3> &octree, int bits,
                  int mbits);

// Host copies of device results (the reference reads them with
// enqueueReadBuffer in its tests).
int ReadKeys(const Buffer &keys, int count, std::vector<uint64_t> &lo, std::vector<uint64_t> &hi);
int ReadBrt(const Buffer &internalBRTNodes, int size, std::vector<BrtNode> &nodes);
// The same results in the reference's layouts: count keys as BigUnsigned
// (keys of at most 88 bits), the size - 1 BRT nodes with their parent links
// as BrtNode.  WriteBigUnsigned uploads BigUnsigned values as key_bytes
// words (a key buffer for RadixSortBigUnsigned and the other stages).
int ReadBigUnsigned(const Buffer &keys, int count, std::vector<BigUnsigned> &out);
int WriteBigUnsigned(const std::vector<BigUnsigned> &in, int key_bytes, Buffer &keys);
int ReadBrtNodes(const Buffer &internalBRTNodes, int size, std::vector<RefBrtNode> &nodes);

}  // namespace Kernels

namespace Karras {

// Karras::Quantize (Karras.cpp:69-99) on the device: float points to grid
// points.  Throws std::logic_error where the reference throws.
std::vector<int2> Quantize(const std::vector<float2> &points, const Resln &resln,
                           const float2 *bbMin = nullptr, const float2 *bbMax = nullptr,
                           bool clamped = false);
std::vector<int3> Quantize(const std::vector<float3> &points, const Resln &resln,
                           const float3 *bbMin = nullptr, const float3 *bbMax = nullptr,
                           bool clamped = false);
// Karras::BuildOctreeInParallel (Karras.cpp:116-120).
std::vector<OctNode2> BuildOctreeInParallel(const std::vector<int2> &points, const Resln &r,
                                            bool verbose = false);
std::vector<OctNode3> BuildOctreeInParallel(const std::vector<int3> &points, const Resln &r,
                                            bool verbose = false);
// Float points straight to the octree on the device (bbox, quantize and the
// whole path in one call, nothing but the octree crossing to the host).
std::vector<OctNode3> BuildOctreeFromPoints(const std::vector<float3> &points, const Resln &r);

}  // namespace Karras

// OctCell (OctCell.h:9-60) without its node pointers: the leaf cell a point
// falls in, with the node holding it and the leaf's data.
template <int DIM>
struct OctCellT {
  int32_t origin[DIM];
  int32_t width;
  int32_t parent_idx;
  int32_t octant;
  int32_t data;   // children[octant] of the parent (-1: an empty cell)
};
using OctCell2 = OctCellT<2>;
using OctCell3 = OctCellT<3>;

namespace OctreeUtils {

// OctreeUtils::FindLeaf (OctreeUtils.cpp:18-71) with the octant masked to
// DIM bits at every level (the reference leaves that mask commented out).
// The single-point form uploads the octree on each call; FindLeaves locates
// a batch in one launch.  Throws std::logic_error where no leaf is reached.
OctCell2 FindLeaf(const int2 &p, const std::vector<OctNode2> &octree, const Resln &resln);
OctCell3 FindLeaf(const int3 &p, const std::vector<OctNode3> &octree, const Resln &resln);
std::vector<OctCell2> FindLeaves(const std::vector<int2> &points,
                                 const std::vector<OctNode2> &octree, const Resln &resln);
std::vector<OctCell3> FindLeaves(const std::vector<int3> &points,
                                 const std::vector<OctNode3> &octree, const Resln &resln);

// CellWalk (viewer/Octree2.cpp:259-313): calls v for every cell the segment
// ab visits, in order.  Throws std::logic_error("Infinite loop") after
// 10000 cells, as the reference does (once the first 10000 are visited).
typedef void (*cwv)(const OctCell2 &cell, const float2 &a, const float2 &b,
                    const std::vector<OctNode2> &octree, const Resln &resln, void *data);
void CellWalk(const float2 &a, const float2 &b, const std::vector<OctNode2> &octree,
              const Resln &resln, cwv v, void *data);
// Many walks in two launches: cells[i] are segment i's cells in order;
// throws std::logic_error("Infinite loop") when any walk passed 10000 cells.
std::vector<std::vector<OctCell2>> CellWalks(const std::vector<float2> &a,
                                             const std::vector<float2> &b,
                                             const std::vector<OctNode2> &octree,
                                             const Resln &resln);

}  // namespace OctreeUtils

// ---- GVD refinement (2D) -----------------------------------------------------

// FloatSegment (opencl/FloatSegment.h): two float points.
struct FloatSegment {
  float2 a, b;
};

// CellIntersections (CellIntersections.h:9-55): per octant the first two
// labels whose segments reached the cell and the longest part of each;
// the record pgvd_cell_intersections lays out (160 bytes).
struct CellIntersections {
  static const int NUM_LABELS = 2;
  static const int NUM_OCTANTS = 4;
  int32_t l[NUM_LABELS * NUM_OCTANTS];
  FloatSegment segs[NUM_LABELS * NUM_OCTANTS];
  bool is_multi(int octant) const { return l[octant * NUM_LABELS + 1] > -1; }
  int num_labels(int octant) const {
    for (int i = 0; i < NUM_LABELS; ++i)
      if (l[octant * NUM_LABELS + i] < 0) return i;
    return NUM_LABELS;
  }
  FloatSegment seg(int i, int octant) const { return segs[octant * NUM_LABELS + i]; }
  int label(int i, int octant) const { return l[octant * NUM_LABELS + i]; }
};

namespace Geom {

// Geom::FitBoxes (opencl/Geom.cpp:667-716) on the device: appends the
// samples, box origins and widths.  Throws std::logic_error where the
// reference throws (and when its loop passes 4096 rounds).
void FitBoxes(const FloatSegment &A, const FloatSegment &B, float min_d,
              std::vector<float2> *samples, std::vector<float2> *origins,
              std::vector<float> *lengths);

}  // namespace Geom

// Octree2's polyline build (viewer/Octree2.cpp:127-190) with its
// FindMultiCells (:430-506) in place: the octree of the polylines'
// vertices, refined by the FitBoxes samples of the cells two polylines
// share, for up to karras_iterations builds; all on the device.
class Octree2 {
 public:
  explicit Octree2(const Resln &resln, int karras_iterations = 1, float min_d = 1.0f);
  // Octree2::build(const Polylines&, const BoundingBox<float2>*): each
  // polygon is a polyline (closed loops repeat their first vertex).
  void build(const std::vector<std::vector<float2>> &polygons, const float2 *bbMin = nullptr,
             const float2 *bbMax = nullptr);
  const std::vector<OctNode2> &octree() const { return octree_; }
  // cell_intersections of the last FindMultiCells, one per octree node
  const std::vector<CellIntersections> &cell_intersections() const { return ci_; }
  // the points of the last build (the quantized vertices, then extra points)
  const std::vector<int2> &qpoints() const { return qpoints_; }
  // extra_qpoints of the last FindMultiCells (not added to a build)
  const std::vector<int2> &extra_qpoints() const { return extra_; }
  int iterations() const { return iterations_; }

 private:
  Resln resln_;
  int karras_iterations_;
  float min_d_;
  int iterations_ = 0;
  std::vector<OctNode2> octree_;
  std::vector<CellIntersections> ci_;
  std::vector<int2> qpoints_, extra_;
};

// operator<< / operator>> for the octree (C/OctNode.h:58-98, commented out
// in the reference), formatted and parsed on the device.  >> reads the rest
// of the stream and sets failbit on malformed text.
std::ostream &operator<<(std::ostream &out, const std::vector<OctNode2> &octree);
std::ostream &operator<<(std::ostream &out, const std::vector<OctNode3> &octree);
std::istream &operator>>(std::istream &in, std::vector<OctNode2> &octree);
std::istream &operator>>(std::istream &in, std::vector<OctNode3> &octree);

}  // namespace pgvd
