// ref_geom.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Links the reference's own 2D geometry (opencl/Geom.cpp: FitBoxes,
// multi_intersection; OctreeUtils.cpp: FindIntersections) and its
// CellIntersections class (CellIntersections.h), compiled unmodified from
// the reference tree, behind a C interface the tests call through ctypes.
// It is the checker of the oracle's restatement of the GVD refinement step;
// the product library never loads it.
//
// MCCallback (viewer/Octree2.cpp:366-428) lives in a file that needs the
// viewer's OpenGL headers, so its dozen lines are restated here on the
// reference's own types and functions.
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "OctreeUtils.h"
#include "opencl/Geom.h"
#include "CellIntersections.h"

namespace {

floatn f2(const float *p) { return make_floatn(p[0], p[1]); }

struct MCState {
  std::vector<CellIntersections> labels;
};

}  // namespace

extern "C" {

int refg_multi_intersection(const float *p1, const float *p2, const float *p3, const float *p4) {
  return Geom::multi_intersection(FloatSegment(f2(p1), f2(p2)), FloatSegment(f2(p3), f2(p4))) ? 1 : 0;
}

// Geom::FitBoxes(A, B, min_d, ...): samples, box origins and widths up to
// cap; returns the number of samples, *boxes_out the number of boxes and
// *status 0, or the logic_error it threw: 1 "doesn't support intersections",
// 2 "no end but not intersecting", 3 "must include one endpoint", 9 other.
int64_t refg_fit_boxes(const float *a, const float *b, float min_d, float *samples,
                       float *origins, float *lengths, int64_t cap, int64_t *boxes_out,
                       int *status) {
  std::vector<floatn> s, o;
  std::vector<float> l;
  *status = 0;
  try {
    Geom::FitBoxes(FloatSegment(f2(a), f2(a + 2)), FloatSegment(f2(b), f2(b + 2)), min_d, &s, &o,
                   &l);
  } catch (const std::logic_error &e) {
    const std::string w = e.what();
    if (w.find("doesn't support intersections") != std::string::npos) *status = 1;
    else if (w.find("no end but not intersecting") != std::string::npos) *status = 2;
    else if (w.find("must include one endpoint") != std::string::npos) *status = 3;
    else *status = 9;
  }
  for (size_t i = 0; i < s.size() && (int64_t)i < cap; ++i) {
    samples[2 * i] = s[i].s[0];
    samples[2 * i + 1] = s[i].s[1];
  }
  for (size_t i = 0; i < o.size() && (int64_t)i < cap; ++i) {
    origins[2 * i] = o[i].s[0];
    origins[2 * i + 1] = o[i].s[1];
    lengths[i] = l[i];
  }
  *boxes_out = (int64_t)o.size();
  return (int64_t)s.size();
}

// OctreeUtils::FindIntersections(a, b, origin, width): up to 4 (t, x, y).
int refg_find_intersections(const float *a, const float *b, const int32_t *origin, int width,
                            float *out) {
  Resln r;
  std::memset(&r, 0, sizeof r);
  intn o;
  o.s[0] = origin[0];
  o.s[1] = origin[1];
  const std::vector<OctreeUtils::CellIntersection> v =
      OctreeUtils::FindIntersections(f2(a), f2(b), o, width, r);
  for (size_t i = 0; i < v.size() && i < 4; ++i) {
    out[3 * i] = v[i].t;
    out[3 * i + 1] = v[i].p.s[0];
    out[3 * i + 2] = v[i].p.s[1];
  }
  return (int)v.size();
}

void *refg_mc_new(int64_t nodes) {
  MCState *m = new MCState;
  m->labels.resize((size_t)nodes, CellIntersections());
  return m;
}

void refg_mc_free(void *h) { delete static_cast<MCState *>(h); }

// MCCallback for one visited cell (origin x, origin y, width, parent node,
// octant) of the walk of segment a-b with label `label`.  Returns the
// number of intersections FindIntersections gave.
int refg_mc_visit(void *h, const int32_t *cell, const float *a, const float *b, int label) {
  MCState *m = static_cast<MCState *>(h);
  intn origin;
  origin.s[0] = cell[0];
  origin.s[1] = cell[1];
  const OctCell c(origin, cell[2], cell[3], nullptr, cell[4], nullptr, 0);
  Resln r;
  std::memset(&r, 0, sizeof r);
  const floatn fa = f2(a), fb = f2(b);
  const std::vector<OctreeUtils::CellIntersection> is =
      OctreeUtils::FindIntersections(fa, fb, c, r);
  if (is.size() == 2) {
    m->labels[c.get_parent_idx()].set(c.get_octant(), label, FloatSegment(is[0].p, is[1].p));
  } else if (is.size() == 1) {
    const BoundingBox<intn> bb = c.bb();
    floatn q;
    bool degenerate = false;
    if (bb.in_half_open(convert_intn(fa))) q = fa;
    else if (bb.in_half_open(convert_intn(fb))) q = fb;
    else degenerate = true;
    if (!degenerate)
      m->labels[c.get_parent_idx()].set(c.get_octant(), label, FloatSegment(is[0].p, q));
  }
  return (int)is.size();
}

// The CellIntersections of every node: labels (8 per node, -1 empty) and
// the segment of every set label (4 floats; zero where empty).
void refg_mc_read(void *h, int32_t *labels, float *segs) {
  MCState *m = static_cast<MCState *>(h);
  for (size_t i = 0; i < m->labels.size(); ++i)
    for (int o = 0; o < 4; ++o)
      for (int j = 0; j < 2; ++j) {
        const int k = (int)(8 * i) + 2 * o + j;
        const int l = m->labels[i].label(j, o);
        labels[k] = l;
        const FloatSegment s = m->labels[i].seg(j, o);
        segs[4 * k] = l < 0 ? 0.0f : s.a().s[0];
        segs[4 * k + 1] = l < 0 ? 0.0f : s.a().s[1];
        segs[4 * k + 2] = l < 0 ? 0.0f : s.b().s[0];
        segs[4 * k + 3] = l < 0 ? 0.0f : s.b().s[1];
      }
}

}  // extern "C"
