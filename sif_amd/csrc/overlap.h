// The exact overlap test of sif__overlap_exact (src/finder/utils.c:850-948),
// split so that an index of any shape can return the reference's answer.
//
// The reference answers "some void in the catalogue collides": it scans the
// cells of a 32^3 void cell list between floor((c -+ s) / cell) on each axis,
// s = scale (r + max_r - f min(r, max_r)) for the catalogue's largest radius
// max_r, and applies a pair test to every void listed there -- the direct
// distance when the whole scan lies inside the box, else the periodic one.
// The answer is therefore
//
//     OR over voids j of  in_range(j, max_r) AND pair(j, interior(max_r)).
//
// A pair that collides under the direct distance also collides under the
// periodic one (each periodic axis distance is at most the direct one), and
// in_range only grows with max_r. So, for a set of voids and a reference
// max_r, overlap_state() returns
//
//     1  some void of the set collides as the reference decides at max_r
//        (and therefore at every larger max_r),
//     0  no void of the set collides under the periodic distance at all
//        (so none ever collides, whatever max_r),
//     2  otherwise: a periodic collision that the reference's scan at max_r
//        does not see; the caller decides with the reference's own scan.
//
// It scans a superset of the cells any periodically colliding void of the set
// can sit in (the set's own largest radius, widened by a margin far above
// float rounding), and decides each pair as the reference does. Everything is
// float with no contraction (the build compiles with -ffp-contract=off), the
// same operations in the same order as the reference, on host and device.
#pragma once
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define SIF_HD __host__ __device__ __forceinline__
#else
#define SIF_HD inline
#endif

namespace sif_amd {

// exodus_finder.c:215: the void cell list is 32 cells a side.
constexpr int32_t VOID_CELLS = 32;

// A grown candidate's state word from the device (grown_overlap): the
// overlap state against the catalogue as the rung began in bits 0-1, and
// this bit when the pole probe at its grown radius hits the mask.
constexpr uint32_t GROWN_POLE_BIT = 16u;

SIF_HD uint32_t void_wrap(int32_t v) {
  const int32_t w = v % VOID_CELLS;
  return (uint32_t)(w < 0 ? w + VOID_CELLS : w);
}

// The reference's cell of a void centre (cell_linked_list.c:41-59).
SIF_HD uint32_t void_axis_cell(float v, float inv_l) {
  return void_wrap((int32_t)floorf(v * inv_l));
}

struct OverlapQuery {
  float cx, cy, cz, r;
  float box, half, f, scale;
  // the reference's scan at the reference max_r
  int32_t lo[3];
  int32_t span[3];  // hi - lo; 31 or more covers every cell
  bool interior;
};

SIF_HD float collision_reach(float r, float rj, float f, float scale) {
  const float mr = r < rj ? r : rj;
  return scale * (r + rj - f * mr);
}

SIF_HD void overlap_query(OverlapQuery& q, float cx, float cy, float cz, float r,
  float max_r, float box, float f, float scale, float inv_l) {
  q.cx = cx;
  q.cy = cy;
  q.cz = cz;
  q.r = r;
  q.box = box;
  q.half = box * 0.5f;
  q.f = f;
  q.scale = scale;
  const float sr = collision_reach(r, max_r, f, scale);
  const float c[3] = {cx, cy, cz};
  for (int a = 0; a < 3; a++) {
    const int32_t lo = (int32_t)floorf((c[a] - sr) * inv_l);
    const int32_t hi = (int32_t)floorf((c[a] + sr) * inv_l);
    q.lo[a] = lo;
    q.span[a] = hi - lo;
  }
  q.interior = (cx - sr >= 0.0f) && (cx + sr < box) && (cy - sr >= 0.0f) &&
               (cy + sr < box) && (cz - sr >= 0.0f) && (cz + sr < box);
}

// Is a void stored in wrapped cell w (one axis) among the reference's cells?
SIF_HD bool in_scan(const OverlapQuery& q, int a, uint32_t w) {
  if (q.span[a] >= VOID_CELLS - 1)
    return true;
  int32_t d = (int32_t)w - (int32_t)void_wrap(q.lo[a]);
  if (d < 0)
    d += VOID_CELLS;
  return d <= q.span[a];
}

// The reference's pair test (utils.c:916-945), direct or periodic.
SIF_HD bool pair_collides(const OverlapQuery& q, bool direct, float vx, float vy,
  float vz, float vr) {
  const float tot = collision_reach(q.r, vr, q.f, q.scale);
  const float t2 = tot * tot;
  if (direct) {
    const float dx = q.cx - vx, dy = q.cy - vy, dz = q.cz - vz;
    return dx * dx + dy * dy + dz * dz < t2;
  }
  float dx = fabsf(q.cx - vx);
  dx = dx > q.half ? q.box - dx : dx;
  const float dx2 = dx * dx;
  if (dx2 >= t2)
    return false;
  float dy = fabsf(q.cy - vy);
  dy = dy > q.half ? q.box - dy : dy;
  const float dy2 = dy * dy;
  if (dx2 + dy2 >= t2)
    return false;
  float dz = fabsf(q.cz - vz);
  dz = dz > q.half ? q.box - dz : dz;
  return dx2 + dy2 + dz * dz < t2;
}

// The cells (per axis, before wrapping) that hold every void of a set whose
// largest radius is set_max_r and that could collide periodically.
SIF_HD void superset_scan(const OverlapQuery& q, float set_max_r, float inv_l,
  int32_t lo[3], int32_t hi[3]) {
  const float reach = collision_reach(q.r, set_max_r, q.f, q.scale);
  const float s = reach + reach * 1e-5f + q.box * 1e-4f;
  const float c[3] = {q.cx, q.cy, q.cz};
  for (int a = 0; a < 3; a++) {
    lo[a] = (int32_t)floorf((c[a] - s) * inv_l);
    hi[a] = (int32_t)floorf((c[a] + s) * inv_l);
    if (hi[a] - lo[a] >= VOID_CELLS) {  // the whole axis, each cell once
      lo[a] = 0;
      hi[a] = VOID_CELLS - 1;
    }
  }
}

// One void found in wrapped cell (wx, wy, wz): its contribution to the state.
SIF_HD int pair_state(const OverlapQuery& q, uint32_t wx, uint32_t wy, uint32_t wz,
  float vx, float vy, float vz, float vr) {
  if (!pair_collides(q, false, vx, vy, vz, vr))
    return 0;
  if (in_scan(q, 0, wx) && in_scan(q, 1, wy) && in_scan(q, 2, wz) &&
      pair_collides(q, q.interior, vx, vy, vz, vr))
    return 1;
  return 2;
}

}  // namespace sif_amd
