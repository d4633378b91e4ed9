// walk_geom.hpp -- the float geometry of one cell-walk step, shared by the
// cell walk (walk.hip) and the multi-cell pass of the GVD refinement
// (gvd.hip).  Internal.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pgvd {

constexpr double kWalkEps = 1e-6;      // opencl/defs.h:135 (a double)

__device__ inline float dot2(float ux, float uy, float vx, float vy) {
  float r = 0.0f;
  r += ux * vx;
  r += uy * vy;
  return r;
}

// OctreeUtils.cpp:80-172 for the cell at (ox, oy) of width w: candidates
// in the reference's order (bottom, right, top, left) in fixed slots t, x,
// y; returns the mask of those kept (the reference's list is the kept slots
// in order).
__device__ inline uint32_t walk_isects(float ax, float ay, float bx, float by, int32_t ox,
                                       int32_t oy, int32_t w, float (&ts)[4], float (&px)[4],
                                       float (&py)[4]) {
  float vx = bx - ax, vy = by - ay;
  const float len = sqrtf(dot2(vx, vy, vx, vy));
  vx = vx / len;
  vy = vy / len;
  const float mnx = (float)ox, mny = (float)oy;
  const float mxx = (float)ox + (float)w, mxy = (float)oy + (float)w;
  ts[0] = ((float)oy - ay) / vy;          // bottom
  px[0] = ax + vx * ts[0];
  py[0] = (float)oy;
  ts[1] = ((float)(ox + w) - ax) / vx;    // right
  px[1] = (float)(ox + w);
  py[1] = ay + vy * ts[1];
  ts[2] = ((float)(oy + w) - ay) / vy;    // top
  px[2] = ax + vx * ts[2];
  py[2] = (float)(oy + w);
  ts[3] = ((float)ox - ax) / vx;          // left
  px[3] = (float)ox;
  py[3] = ay + vy * ts[3];
  const float eps = (float)kWalkEps;
  uint32_t keep = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float t = ts[i];
    if ((double)t >= -kWalkEps && (double)t < (double)len + kWalkEps &&
        !(px[i] < mnx - eps || px[i] > mxx + eps) && !(py[i] < mny - eps || py[i] > mxy + eps))
      keep |= 1u << i;
  }
  const int n = __popc(keep);
  if (n == 2) {
    const int f0 = __ffs(keep) - 1, f1 = 31 - __clz(keep);
    float x0 = 0, y0 = 0, x1 = 0, y1 = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i == f0) { x0 = px[i]; y0 = py[i]; }
      if (i == f1) { x1 = px[i]; y1 = py[i]; }
    }
    const float dx = x0 - x1, dy = y0 - y1;
    if ((double)sqrtf(dot2(dx, dy, dx, dy)) < kWalkEps) keep &= ~(1u << f1);
  } else if (n > 2) {
    // the first two kept slots unless a pair's |dot| with v exceeds 0
    int bi = __ffs(keep) - 1;
    int bj = __ffs(keep & (keep - 1)) - 1;
    float best = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = i + 1; j < 4; ++j) {
        if (((keep >> i) & 1) && ((keep >> j) & 1)) {
          float ux = px[j] - px[i], uy = py[j] - py[i];
          const float l = sqrtf(dot2(ux, uy, ux, uy));
          ux = ux / l;
          uy = uy / l;
          const float cd = fabsf(dot2(ux, uy, vx, vy));
          if (cd > best) {
            best = cd;
            bi = i;
            bj = j;
          }
        }
      }
    keep = (1u << bi) | (1u << bj);
  }
  return keep;
}

}  // namespace pgvd
