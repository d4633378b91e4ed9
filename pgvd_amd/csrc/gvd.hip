// gvd.hip -- the GVD refinement step (2D) on the device: FindMultiCells
// (viewer/Octree2.cpp:430-506) and Geom::FitBoxes (opencl/Geom.cpp:467-716).
//
// Multi-cells.  FindMultiCells walks every polyline segment through the
// octree (the batched cell walk, walk.hip) and, for every visited cell,
// MCCallback (viewer/Octree2.cpp:366-428) keeps the segment's part inside
// the cell under (parent node, octant) in a CellIntersections record
// (CellIntersections.h:9-55): per octant the first two labels to arrive,
// and for each the longest part (a strictly longer part replaces; the
// first of equal ones stays).  The reference folds the visits one by one;
// here every visit is a record at its position g in walk order (polygon by
// polygon, segment by segment, cell by cell) and the fold is three
// order-independent reductions per (node, octant):
//   first  = min g of the visits that keep a part      -> label 0
//   second = min g of those whose label differs         -> label 1
//   best   = max (length^2 bits << 32 | ~g) per label   -> the kept part
// which is what the sequential fold keeps.
//
// FitBoxes.  One lane per pair of segments runs the reference's loop with
// the same float and double operations (no contraction, correctly rounded
// division and square root): a counting pass, a scan of the counts, and a
// writing pass.  The reference's loop has no bound; here a pair stops with
// an error after kFitMaxRounds rounds.
#include "kernels.hpp"
#include "tilescan.hpp"
#include "walk_geom.hpp"

namespace pgvd {

namespace {

struct v2 {
  float x, y;
};
struct seg2 {
  v2 a, b;
};

__device__ inline v2 mk(float x, float y) { return v2{x, y}; }
__device__ inline v2 add(v2 u, v2 v) { return mk(u.x + v.x, u.y + v.y); }
__device__ inline v2 sub(v2 u, v2 v) { return mk(u.x - v.x, u.y - v.y); }
__device__ inline v2 muls(v2 u, float s) { return mk(u.x * s, u.y * s); }
__device__ inline v2 divs(v2 u, float s) { return mk(u.x / s, u.y / s); }
__device__ inline float dotv(v2 u, v2 v) { return dot2(u.x, u.y, v.x, v.y); }
__device__ inline float lenv(v2 v) { return sqrtf(dotv(v, v)); }
__device__ inline v2 unitv(v2 v) { return divs(v, lenv(v)); }
__device__ inline bool eqv(v2 u, v2 v) { return u.x == v.x && u.y == v.y; }
__device__ inline float distv(v2 a, v2 b) { return lenv(sub(a, b)); }     // opencl/Geom.h:16-18
__device__ inline seg2 sg(v2 a, v2 b) { return seg2{a, b}; }
__device__ inline v2 vecs(seg2 s) { return sub(s.b, s.a); }
__device__ inline seg2 rev(seg2 s) { return sg(s.b, s.a); }
__device__ inline int sgnf(float f) { return f < 0 ? -1 : (f > 0 ? 1 : 0); }  // Geom.h:10-12

// Geom::closest(p, s) (opencl/Geom.cpp:15-29); its epsilon is a float.
__device__ v2 closest_pt(v2 p, seg2 s) {
  if (eqv(s.a, s.b)) return s.a;
  const float eps = (float)1e-6;
  v2 v = vecs(s);
  const float len = lenv(v);
  v = divs(v, len);
  const v2 u = sub(p, s.a);
  const float t = dotv(u, v);
  v2 q = add(s.a, muls(v, t));
  if (t < 0 || distv(q, s.a) < eps) q = s.a;
  else if (t > len || distv(q, s.b) < eps) q = s.b;
  return q;
}

// Geom::x_given_y, y_given_x (opencl/Geom.cpp:31-53).
__device__ inline bool x_given_y(seg2 s, float y, float &x) {
  const v2 v = vecs(s);
  const float t = (y - s.a.y) / v.y;
  x = s.a.x + t * v.x;
  return t <= 1;
}
__device__ inline bool y_given_x(seg2 s, float x, float &y) {
  const v2 v = vecs(s);
  const float t = (x - s.a.x) / v.x;
  y = s.a.y + t * v.y;
  return t <= 1;
}

// Geom::clip_to_box (opencl/Geom.cpp:55-96).
__device__ bool clip_to_box(seg2 s, float d, seg2 &out) {
  const v2 c = s.a;
  const float d2 = d / 2;
  const v2 sv = vecs(s);
  bool valid;
  float x, y;
  if (fabsf(sv.x) > fabsf(sv.y)) {
    x = c.x + d2;
    if (sv.x < 0) x = c.x - d2;
    valid = y_given_x(s, x, y);
  } else {
    y = c.y + d2;
    if (sv.y < 0) y = c.y - d2;
    valid = x_given_y(s, y, x);
  }
  if (valid) out = sg(mk(x, y), s.b);
  return valid;
}

// Geom::multi_intersection (opencl/Geom.cpp:109-151): the slope is a float
// quotient widened to double and the collinearity test compares in double.
__device__ bool multi_isect(v2 p1, v2 p2, v2 p3, v2 p4) {
  const float x1 = p1.x, x2 = p2.x, x3 = p3.x, x4 = p4.x;
  const float y1 = p1.y, y2 = p2.y, y3 = p3.y, y4 = p4.y;
  const float d = (x1 - x2) * (y3 - y4) - (y1 - y2) * (x3 - x4);
  if (d == 0) {
    if (x2 - x1 != 0) {
      const double m = (double)((y2 - y1) / (x2 - x1));
      if ((double)(y1 - y3) == m * (double)(x1 - x3)) {
        const float len1 = fabsf(x2 - x1), len2 = fabsf(x4 - x3);
        const float len = fmaxf(x1, fmaxf(x2, fmaxf(x3, x4))) - fminf(x1, fminf(x2, fminf(x3, x4)));
        if (len < len1 + len2) return true;
      }
    } else if (x1 == x3) {
      const float len1 = fabsf(y2 - y1), len2 = fabsf(y4 - y3);
      const float len = fmaxf(y1, fmaxf(y2, fmaxf(y3, y4))) - fminf(y1, fminf(y2, fminf(y3, y4)));
      if (len < len1 + len2) return true;
    }
  }
  return false;
}

// Geom::line_intersection (opencl/Geom.cpp:165-206), in double.
__device__ v2 line_isect(v2 p1, v2 p2, v2 p3, v2 p4, bool &segs) {
  const double x1 = p1.x, x2 = p2.x, x3 = p3.x, x4 = p4.x;
  const double y1 = p1.y, y2 = p2.y, y3 = p3.y, y4 = p4.y;
  const double d = (x1 - x2) * (y3 - y4) - (y1 - y2) * (x3 - x4);
  if (d == 0) {
    segs = false;
    return mk(0, 0);
  }
  const double pre = (x1 * y2 - y1 * x2), post = (x3 * y4 - y3 * x4);
  const double x = (pre * (x3 - x4) - (x1 - x2) * post) / d;
  const double y = (pre * (y3 - y4) - (y1 - y2) * post) / d;
  segs = !(x < fmin(x1, x2) || x > fmax(x1, x2) || x < fmin(x3, x4) || x > fmax(x3, x4)) &&
         !(y < fmin(y1, y2) || y > fmax(y1, y2) || y < fmin(y3, y4) || y > fmax(y3, y4));
  return mk((float)x, (float)y);
}

// Geom::closest(a, b, ...) for two segments (opencl/Geom.cpp:259-346).
__device__ void closest_segs(seg2 a, seg2 b, v2 &ca, v2 &cb, bool &ca_end, bool &cb_end) {
  bool si;
  const v2 l = line_isect(a.a, a.b, b.a, b.b, si);
  if (si) {
    ca = cb = l;
    ca_end = cb_end = false;
    return;
  }
  v2 c = closest_pt(a.a, b);
  ca = a.a;
  cb = c;
  ca_end = true;
  cb_end = eqv(c, b.a) || eqv(c, b.b);
  float dist = lenv(sub(a.a, c));
  v2 c_ = closest_pt(a.b, b);
  float dist_ = lenv(sub(a.b, c_));
  if (dist_ < dist) {
    c = c_;
    ca = a.b;
    cb = c;
    ca_end = true;
    cb_end = eqv(c, b.a) || eqv(c, b.b);
    dist = dist_;
  }
  c_ = closest_pt(b.a, a);
  dist_ = lenv(sub(b.a, c_));
  if (dist_ < dist) {
    c = c_;
    ca = c;
    cb = b.a;
    ca_end = eqv(c, a.a) || eqv(c, a.b);
    cb_end = true;
    dist = dist_;
  }
  c_ = closest_pt(b.b, a);
  dist_ = lenv(sub(b.b, c_));
  if (dist_ < dist) {
    c = c_;
    ca = c;
    cb = b.b;
    ca_end = eqv(c, a.a) || eqv(c, a.b);
    cb_end = true;
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const v2 a0 = i ? a.b : a.a, a1 = i ? a.a : a.b;
      const v2 b0 = j ? b.b : b.a, b1 = j ? b.a : b.b;
      if (a0.x == b0.x) {
        const int day = sgnf(a1.y - a0.y), dby = sgnf(b1.y - b0.y), daby = sgnf(b0.y - a0.y);
        if (day != dby && day != daby) {
          ca = a0;
          cb = b0;
          ca_end = cb_end = true;
        }
      } else if (a0.y == b0.y) {
        const int dax = sgnf(a1.x - a0.x), dbx = sgnf(b1.x - b0.x), dabx = sgnf(b0.x - a0.x);
        if (dax != dbx && dax != dabx) {
          ca = a0;
          cb = b0;
          ca_end = cb_end = true;
        }
      }
    }
}

constexpr int kFitMaxRounds = 4096;
constexpr int kFitOffsetFactor = 4;

// Where FitBoxes puts what it finds: counts always, values while WRITE.
template <bool WRITE>
struct FitSink {
  float *samples, *origins, *lengths;
  int32_t *extra;
  int64_t ns = 0, nb = 0;  // running positions (start at the item's offsets)
  __device__ void sample(v2 p) {
    if (WRITE) {
      samples[2 * ns] = p.x;
      samples[2 * ns + 1] = p.y;
      if (extra) {
        extra[2 * ns] = (int32_t)p.x;      // convert_intn (truncation)
        extra[2 * ns + 1] = (int32_t)p.y;
      }
    }
    ++ns;
  }
  __device__ void box(v2 o, float d) {
    if (WRITE) {
      origins[2 * nb] = o.x;
      origins[2 * nb + 1] = o.y;
      lengths[nb] = d;
    }
    ++nb;
  }
};

// Geom::FitBoxesNoIntersection (opencl/Geom.cpp:471-664).  Returns 0, or 1,
// 2, 3 for the reference's logic_errors, 4 past kFitMaxRounds rounds.
template <class Sink>
__device__ int fit_no_isect(seg2 A, seg2 B, float min_d, Sink &o) {
  bool done = false;
  int rounds = 0;
  while (!done) {
    if (++rounds > kFitMaxRounds) return 4;
    v2 ca, cb, c_cb;
    bool ca_end, cb_end, c_cb_end;
    closest_segs(A, B, ca, cb, ca_end, cb_end);
    if (eqv(ca, cb)) return 1;
    if (!ca_end && !cb_end) return 2;
    if (cb_end) {
      const v2 t = ca;
      ca = cb;
      cb = t;
      const bool te = ca_end;
      ca_end = cb_end;
      cb_end = te;
      const seg2 ts = A;
      A = B;
      B = ts;
    }
    if (!ca_end) return 3;
    if (!eqv(A.a, ca)) A = rev(A);
    if (cb_end && !eqv(B.a, cb)) B = rev(B);

    float d;
    v2 org;
    v2 off = mk(0, 0);
    if (distv(ca, cb) < min_d) {
      org = sub(divs(add(ca, cb), 2), mk(min_d / 2, min_d / 2));
      d = min_d;
      c_cb = cb;
      c_cb_end = cb_end;
    } else if (cb_end && (A.a.x == B.a.x || A.a.y == B.a.y)) {
      d = distv(A.a, B.a);
      org = mk(fminf(A.a.x, B.a.x), fminf(A.a.y, B.a.y));
      if (A.a.x == B.a.x) {
        off = mk(d / kFitOffsetFactor, 0);
        if (vecs(A).x < 0) {
          org = sub(org, mk(d, 0));
          off = muls(off, -1);
        }
      } else if (A.a.y == B.a.y) {
        off = mk(0, d / 16);
        if (vecs(A).y < 0) {
          org = sub(org, mk(0, d));
          off = muls(off, -1);
        }
      }
      c_cb = cb;
      c_cb_end = true;
    } else {
      const v2 v = sub(cb, ca);
      d = lenv(v);
      const float cx = (v.x >= 0) ? 2 * d : -2 * d;
      const float cy = (v.y >= 0) ? 2 * d : -2 * d;
      const seg2 C = sg(ca, add(ca, mk(cx, cy)));
      v2 c_cc;
      bool c_cc_end;
      closest_segs(B, C, c_cb, c_cc, c_cb_end, c_cc_end);
      float d_ = distv(ca, c_cb) / (float)1.4142135623730951;   // (float)sqrt(2)
      if (!eqv(c_cb, c_cc)) d_ = fmaxf(fabsf(c_cb.x - ca.x), fabsf(c_cb.y - ca.y));
      if (d_ < d) d = d_;
      const float bxd = (v.x >= 0) ? 0 : -d;
      const float byd = (v.y >= 0) ? 0 : -d;
      org = add(ca, mk(bxd, byd));
    }
    o.sample(add(add(ca, divs(sub(cb, ca), kFitOffsetFactor)), off));
    o.sample(add(add(cb, divs(sub(ca, cb), kFitOffsetFactor)), off));
    o.box(org, d);

    if (c_cb_end) {
      if (!eqv(B.a, c_cb)) B = rev(B);
    } else {
      const seg2 B0 = sg(c_cb, B.a), B1 = sg(c_cb, B.b);
      B = dotv(unitv(vecs(A)), unitv(vecs(B0))) > 0 ? B0 : B1;
    }

    if (dotv(unitv(vecs(A)), unitv(vecs(B))) <= 0) {
      done = true;
    } else {
      const v2 bis = unitv(add(unitv(vecs(A)), unitv(vecs(B))));
      bool horizontal;
      if (A.a.x == B.a.x) horizontal = true;
      else if (A.a.y == B.a.y) horizontal = false;
      else horizontal = fabsf(bis.x) > fabsf(bis.y);
      float Ax, Ay, Bx, By;
      bool Ai, Bi;
      if (horizontal) {
        Ax = Bx = (bis.x < 0) ? org.x : org.x + d;
        Ai = y_given_x(A, Ax, Ay);
        Bi = y_given_x(B, Bx, By);
      } else {
        Ay = By = (bis.y < 0) ? org.y : org.y + d;
        Ai = x_given_y(A, Ay, Ax);
        Bi = x_given_y(B, By, Bx);
      }
      done = !Ai || !Bi;
      if (eqv(A.a, mk(Ax, Ay)) && eqv(B.a, mk(Bx, By))) done = true;
      A = sg(mk(Ax, Ay), A.b);
      B = sg(mk(Bx, By), B.b);
    }
  }
  return 0;
}

// Geom::FitBoxes (opencl/Geom.cpp:667-716).  Where the segments meet, the
// reference reads the unit vectors of the first clipped pair even when a
// clip failed (uninitialised segments); a failed clip here leaves a zero
// segment, whose unit vector is NaN, so the pairs are not swapped.
template <class Sink>
__device__ int fit_boxes(seg2 A, seg2 B, float min_d, Sink &o) {
  v2 ca, cb;
  bool ca_end, cb_end;
  closest_segs(A, B, ca, cb, ca_end, cb_end);
  if (!eqv(ca, cb)) return fit_no_isect(A, B, min_d, o);
  o.sample(ca);
  o.box(sub(ca, mk(min_d / 2, min_d / 2)), min_d);
  if (A.a.x > A.b.x) A = rev(A);
  if (B.a.x > B.b.x) B = rev(B);
  const seg2 zero = sg(mk(0, 0), mk(0, 0));
  seg2 A0 = zero, A1 = zero, B0 = zero, B1 = zero;
  bool A0v = clip_to_box(sg(ca, A.a), min_d, A0);
  bool A1v = clip_to_box(sg(ca, A.b), min_d, A1);
  const bool B0v = clip_to_box(sg(cb, B.a), min_d, B0);
  const bool B1v = clip_to_box(sg(cb, B.b), min_d, B1);
  if (dotv(unitv(vecs(A0)), unitv(vecs(B0))) < 0) {
    const seg2 t = A0;
    A0 = A1;
    A1 = t;
    const bool tv = A0v;
    A0v = A1v;
    A1v = tv;
  }
  int rc = 0;
  if (A0v && B0v) rc = fit_no_isect(A0, B0, min_d, o);
  if (!rc && A1v && B1v) rc = fit_no_isect(A1, B1, min_d, o);
  return rc;
}

__device__ inline seg2 load_seg(const float *p) { return sg(mk(p[0], p[1]), mk(p[2], p[3])); }

// The pair of item i: from a pair list (8 floats each), or from the
// CellIntersections records (slot i = node * 4 + octant), where the slot
// is an item when it holds two labels whose parts do not overlap on one
// line (viewer/Octree2.cpp:466-479).
template <bool FROM_CI>
__device__ inline bool fit_item(const float *__restrict__ src, int64_t i, seg2 &A, seg2 &B) {
  if (!FROM_CI) {
    A = load_seg(src + 8 * i);
    B = load_seg(src + 8 * i + 4);
    return true;
  }
  const int64_t node = i >> 2;
  const int oc = (int)(i & 3);
  const int32_t *lab = reinterpret_cast<const int32_t *>(src) + 40 * node;
  if (lab[2 * oc + 1] < 0) return false;
  const float *s = src + 40 * node + 8 + 8 * oc;
  A = load_seg(s);
  B = load_seg(s + 4);
  return !multi_isect(A.a, A.b, B.a, B.b);
}

// Counting pass: per item the number of samples and boxes; the first
// failing item and its error go to *err as (item << 8 | code) (min).
template <bool FROM_CI>
__global__ __launch_bounds__(kBlock) void fit_count_kernel(const float *__restrict__ src,
                                                           int64_t n, float min_d,
                                                           uint32_t *__restrict__ ns,
                                                           uint32_t *__restrict__ nb,
                                                           int32_t *__restrict__ status,
                                                           unsigned long long *__restrict__ err) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    seg2 A, B;
    FitSink<false> o{nullptr, nullptr, nullptr, nullptr};
    int rc = 0;
    if (fit_item<FROM_CI>(src, i, A, B)) rc = fit_boxes(A, B, min_d, o);
    ns[i] = (uint32_t)o.ns;
    nb[i] = (uint32_t)o.nb;
    if (status) status[i] = rc;
    if (rc) atomicMin(err, ((unsigned long long)i << 8) | (unsigned long long)rc);
  }
}

template <bool FROM_CI>
__global__ __launch_bounds__(kBlock) void fit_write_kernel(
    const float *__restrict__ src, int64_t n, float min_d, const int64_t *__restrict__ soff,
    const int64_t *__restrict__ boff, const uint32_t *__restrict__ ns, float *__restrict__ samples,
    float *__restrict__ origins, float *__restrict__ lengths, int32_t *__restrict__ extra) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    if (ns[i] == 0) continue;
    seg2 A, B;
    if (!fit_item<FROM_CI>(src, i, A, B)) continue;
    FitSink<true> o{samples, origins, lengths, extra};
    o.ns = soff[i];
    o.nb = boff[i];
    (void)fit_boxes(A, B, min_d, o);
  }
}

// ---- multi-cells ----------------------------------------------------------

// Octree2::obj2Oct (viewer/Octree2.cpp:215-225) of every vertex, then the
// segments of every polygon: vertex i of polygon j (vertices [voff[j],
// voff[j+1])) starts segment soff[j] + (i - voff[j]) when i + 1 < voff[j+1],
// labelled j.
__global__ __launch_bounds__(kBlock) void poly_segments_kernel(
    const float *__restrict__ verts, int64_t nv, const int64_t *__restrict__ voff,
    const int64_t *__restrict__ soff, int64_t P, float mnx, float mny, float bbw, float width,
    float clampv, float *__restrict__ segs, int32_t *__restrict__ labels,
    uint32_t *__restrict__ bad) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nv;
       i += (int64_t)gridDim.x * kBlock) {
    int64_t lo = 0, hi = P - 1;       // last j with voff[j] <= i
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (voff[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    const int64_t j = lo;
    float c[2];
    const float mn[2] = {mnx, mny};
    bool neg = false;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      float q = ((verts[2 * i + k] - mn[k]) / bbw) * width;
      if (q >= width) q = clampv;
      neg |= q < 0;
      c[k] = q;
    }
    if (neg) atomicOr(bad, 1u);
    const int64_t k = i - voff[j];
    if (i + 1 < voff[j + 1]) {          // vertex i starts a segment
      const int64_t sidx = soff[j] + k;
      segs[4 * sidx] = c[0];
      segs[4 * sidx + 1] = c[1];
      labels[sidx] = (int32_t)j;
    }
    if (k > 0) {                        // and ends the one before it
      const int64_t sidx = soff[j] + k - 1;
      segs[4 * sidx + 2] = c[0];
      segs[4 * sidx + 3] = c[1];
    }
  }
}

// MCCallback for every visited cell: the part kept (float4) and the label,
// or -1 where nothing is kept.  One lane per segment, over its cells.
__global__ __launch_bounds__(kBlock) void mc_visit_kernel(
    const float *__restrict__ segs, const int32_t *__restrict__ labels, int64_t n,
    const int64_t *__restrict__ offsets, const int32_t *__restrict__ cells,
    float4 *__restrict__ parts, int32_t *__restrict__ vlabel) {
  for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * kBlock) {
    const float ax = segs[4 * s], ay = segs[4 * s + 1], bx = segs[4 * s + 2], by = segs[4 * s + 3];
    const int32_t ai0 = (int32_t)ax, ai1 = (int32_t)ay, bi0 = (int32_t)bx, bi1 = (int32_t)by;
    const int32_t lab = labels[s];
    for (int64_t g = offsets[s]; g < offsets[s + 1]; ++g) {
      const int32_t *c = cells + 5 * g;
      const int32_t ox = c[0], oy = c[1], w = c[2];
      float ts[4], px[4], py[4];
      const uint32_t keep = walk_isects(ax, ay, bx, by, ox, oy, w, ts, px, py);
      const int nk = __popc(keep);
      float4 part = make_float4(0, 0, 0, 0);
      int32_t l = -1;
      if (nk >= 1) {
        const int f0 = __ffs(keep) - 1;
        float x0 = 0, y0 = 0, x1 = 0, y1 = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (i == f0) { x0 = px[i]; y0 = py[i]; }
        if (nk == 2) {
          const int f1 = 31 - __clz(keep);
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (i == f1) { x1 = px[i]; y1 = py[i]; }
          part = make_float4(x0, y0, x1, y1);
          l = lab;
        } else if (nk == 1) {
          // the endpoint whose truncated coordinates lie in the cell's
          // half-open box (BoundingBox.h:97-103)
          const bool ain = ai0 >= ox && ai0 < ox + w && ai1 >= oy && ai1 < oy + w;
          const bool bin = bi0 >= ox && bi0 < ox + w && bi1 >= oy && bi1 < oy + w;
          if (ain) {
            part = make_float4(x0, y0, ax, ay);
            l = lab;
          } else if (bin) {
            part = make_float4(x0, y0, bx, by);
            l = lab;
          }
        }
      }
      parts[g] = part;
      vlabel[g] = l;
    }
  }
}

__device__ inline int64_t slot_of(const int32_t *c) { return (int64_t)c[3] * 4 + c[4]; }

// first[slot] = min g of the visits that keep a part.
__global__ __launch_bounds__(kBlock) void mc_first_kernel(const int32_t *__restrict__ cells,
                                                          const int32_t *__restrict__ vlabel,
                                                          int64_t total,
                                                          uint32_t *__restrict__ first) {
  for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * kBlock)
    if (vlabel[g] >= 0) atomicMin(first + slot_of(cells + 5 * g), (uint32_t)g);
}

// second[slot] = min g of the kept visits whose label is not first's.
__global__ __launch_bounds__(kBlock) void mc_second_kernel(const int32_t *__restrict__ cells,
                                                           const int32_t *__restrict__ vlabel,
                                                           int64_t total,
                                                           const uint32_t *__restrict__ first,
                                                           uint32_t *__restrict__ second) {
  for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * kBlock) {
    const int32_t l = vlabel[g];
    if (l < 0) continue;
    const int64_t slot = slot_of(cells + 5 * g);
    if (l != vlabel[first[slot]]) atomicMin(second + slot, (uint32_t)g);
  }
}

// best[2 slot + j] = max over the visits of label j of
// (bits of length^2 << 32 | ~g): the longest part, the first of equals.
__global__ __launch_bounds__(kBlock) void mc_best_kernel(const int32_t *__restrict__ cells,
                                                         const int32_t *__restrict__ vlabel,
                                                         const float4 *__restrict__ parts,
                                                         int64_t total,
                                                         const uint32_t *__restrict__ first,
                                                         const uint32_t *__restrict__ second,
                                                         unsigned long long *__restrict__ best) {
  for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * kBlock) {
    const int32_t l = vlabel[g];
    if (l < 0) continue;
    const int64_t slot = slot_of(cells + 5 * g);
    int j;
    if (l == vlabel[first[slot]]) j = 0;
    else if (second[slot] != 0xFFFFFFFFu && l == vlabel[second[slot]]) j = 1;
    else continue;
    const float4 p = parts[g];
    const float len2 = dot2(p.z - p.x, p.w - p.y, p.z - p.x, p.w - p.y);   // FloatSegment::length2
    const unsigned long long key =
        ((unsigned long long)__float_as_uint(len2) << 32) | (unsigned long long)(~(uint32_t)g);
    atomicMax(best + 2 * slot + j, key);
  }
}

// CellIntersections records: node i's labels (8 int32) then its parts (8
// float4), 40 words per node.
__global__ __launch_bounds__(kBlock) void mc_write_kernel(int64_t slots,
                                                          const int32_t *__restrict__ vlabel,
                                                          const float4 *__restrict__ parts,
                                                          const uint32_t *__restrict__ first,
                                                          const uint32_t *__restrict__ second,
                                                          const unsigned long long *__restrict__ best,
                                                          int32_t *__restrict__ ci) {
  for (int64_t slot = (int64_t)blockIdx.x * kBlock + threadIdx.x; slot < slots;
       slot += (int64_t)gridDim.x * kBlock) {
    const int64_t node = slot >> 2;
    const int oc = (int)(slot & 3);
    int32_t *rec = ci + 40 * node;
    float4 *seg = reinterpret_cast<float4 *>(rec + 8);
    const uint32_t f[2] = {first[slot], second[slot]};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      int32_t l = -1;
      float4 p = make_float4(0, 0, 0, 0);
      if (f[j] != 0xFFFFFFFFu && (j == 0 || f[0] != 0xFFFFFFFFu)) {
        l = vlabel[f[j]];
        const uint32_t g = ~(uint32_t)(best[2 * slot + j] & 0xFFFFFFFFull);
        p = parts[g];
      }
      rec[2 * oc + j] = l;
      seg[2 * oc + j] = p;
    }
  }
}

__global__ void fill_u32_kernel(uint32_t *p, int64_t n, uint32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    p[i] = v;
}

__global__ void fill_u64_kernel(unsigned long long *p, int64_t n, unsigned long long v) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    p[i] = v;
}

inline unsigned grid_for(int64_t n, int64_t cap = 16384) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace

void launch_poly_segments(const float *verts, int64_t nv, const int64_t *voff, const int64_t *soff,
                          int64_t P, const float bbmin[2], float bbw, int bits, float *segs,
                          int32_t *labels, uint32_t *bad, hipStream_t s) {
  if (nv == 0) return;
  const double width = (double)(int64_t(1) << bits);
  const float clampv = (float)(width - 0.0001);   // Octree2.cpp:220 (int - double)
  poly_segments_kernel<<<grid_for(nv), kBlock, 0, s>>>(verts, nv, voff, soff, P, bbmin[0],
                                                       bbmin[1], bbw, (float)width, clampv,
                                                       segs, labels, bad);
}

void launch_multi_cells(const float *segs, const int32_t *labels, int64_t n,
                        const int64_t *offsets, const int32_t *cells, int64_t total, int64_t O,
                        float4 *parts, int32_t *vlabel, uint32_t *first, uint32_t *second,
                        unsigned long long *best, int32_t *ci, hipStream_t s) {
  const int64_t slots = 4 * O;
  fill_u32_kernel<<<grid_for(slots), kBlock, 0, s>>>(first, slots, 0xFFFFFFFFu);
  fill_u32_kernel<<<grid_for(slots), kBlock, 0, s>>>(second, slots, 0xFFFFFFFFu);
  fill_u64_kernel<<<grid_for(2 * slots), kBlock, 0, s>>>(best, 2 * slots, 0ull);
  if (n > 0 && total > 0) {
    mc_visit_kernel<<<grid_for(n), kBlock, 0, s>>>(segs, labels, n, offsets, cells, parts, vlabel);
    mc_first_kernel<<<grid_for(total), kBlock, 0, s>>>(cells, vlabel, total, first);
    mc_second_kernel<<<grid_for(total), kBlock, 0, s>>>(cells, vlabel, total, first, second);
    mc_best_kernel<<<grid_for(total), kBlock, 0, s>>>(cells, vlabel, parts, total, first, second,
                                                       best);
  }
  mc_write_kernel<<<grid_for(slots), kBlock, 0, s>>>(slots, vlabel, parts, first, second, best, ci);
}

void launch_fit_count(const float *src, bool from_ci, int64_t items, float min_d, uint32_t *ns,
                      uint32_t *nb, int32_t *status, unsigned long long *err, hipStream_t s) {
  fill_u64_kernel<<<1, kBlock, 0, s>>>(err, 1, ~0ull);
  if (items == 0) return;
  if (from_ci)
    fit_count_kernel<true><<<grid_for(items), kBlock, 0, s>>>(src, items, min_d, ns, nb, status, err);
  else
    fit_count_kernel<false><<<grid_for(items), kBlock, 0, s>>>(src, items, min_d, ns, nb, status,
                                                                err);
}

void launch_fit_write(const float *src, bool from_ci, int64_t items, float min_d,
                      const int64_t *soff, const int64_t *boff, const uint32_t *ns, float *samples,
                      float *origins, float *lengths, int32_t *extra, hipStream_t s) {
  if (items == 0) return;
  if (from_ci)
    fit_write_kernel<true><<<grid_for(items), kBlock, 0, s>>>(src, items, min_d, soff, boff, ns,
                                                              samples, origins, lengths, extra);
  else
    fit_write_kernel<false><<<grid_for(items), kBlock, 0, s>>>(src, items, min_d, soff, boff, ns,
                                                               samples, origins, lengths, extra);
}

}  // namespace pgvd

namespace pgvd {
namespace {
__global__ void narrow_i64_kernel(const int64_t *__restrict__ in, int64_t n, int32_t *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    out[i] = (int32_t)in[i];
}
}  // namespace

void launch_narrow_i64(const int64_t *in, int64_t n, int32_t *out, hipStream_t s) {
  if (n > 0) narrow_i64_kernel<<<grid_for(n), kBlock, 0, s>>>(in, n, out);
}
}  // namespace pgvd

namespace pgvd {
namespace {
__global__ void any_nonzero_kernel(const int32_t *__restrict__ v, int64_t n, uint32_t *flag) {
  bool any = false;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    any |= v[i] != 0;
  if (__any(any) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}
}  // namespace

void launch_any_nonzero(const int32_t *v, int64_t n, uint32_t *flag, hipStream_t s) {
  fill_u32_kernel<<<1, kBlock, 0, s>>>(flag, 1, 0u);
  if (n > 0) any_nonzero_kernel<<<grid_for(n, 2048), kBlock, 0, s>>>(v, n, flag);
}
}  // namespace pgvd
