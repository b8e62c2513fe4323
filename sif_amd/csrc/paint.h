// The node test of sif__mark_sphere (src/finder/utils.c:965-1033): which
// nodes of the n^3 grid a void paints into the mask, for host and device.
//
// The reference paints by walking the cube of nodes (o - c .. o + c) around
// the centre, o = (int)(centre / l), c = (int)(r / l) + 1, and setting node g
// (wrapped) when its folded axis distances pass dx^2 <= r^2, then
// dx^2 + dy^2 <= r^2, then dx^2 + dy^2 + dz^2 <= r^2. paint_hits() answers
// the same question for one wrapped node by trying every unwrapped image of it
// inside that cube, with the same float operations in the same order (the
// build compiles with -ffp-contract=off); the device paint kernel walks the
// cube as the reference does. Both give the reference's bits.
#pragma once
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define SIF_HD __host__ __device__ __forceinline__
#else
#define SIF_HD inline
#endif

namespace sif_amd {

struct PaintSphere {
  float cx, cy, cz;
  float r2, l, L, half;
  int32_t c, ox, oy, oz;
  int32_t n;
};

SIF_HD void paint_sphere(PaintSphere& P, float cx, float cy, float cz, float r, float l,
  uint32_t n) {
  P.cx = cx;
  P.cy = cy;
  P.cz = cz;
  P.r2 = r * r;
  const float inv_l = 1.0f / l;
  P.l = l;
  P.L = (float)n * l;
  P.half = P.L * 0.5f;
  P.c = (int32_t)(r * inv_l) + 1;
  P.ox = (int32_t)(cx * inv_l);
  P.oy = (int32_t)(cy * inv_l);
  P.oz = (int32_t)(cz * inv_l);
  P.n = (int32_t)n;
}

// |g l - centre|, folded at half the box
SIF_HD float paint_dist(const PaintSphere& P, int32_t g, float centre) {
  const float d = fabsf((float)g * P.l - centre);
  return d > P.half ? P.L - d : d;
}

SIF_HD uint32_t paint_wrap(int32_t g, int32_t n) {
  if (g >= -n && g < 2 * n)  // within a box of the grid: no division
    return (uint32_t)(g < 0 ? g + n : (g >= n ? g - n : g));
  const int32_t w = g % n;
  return (uint32_t)(w < 0 ? w + n : w);
}

// the first image i + k n at or above lo
SIF_HD int32_t paint_first_image(uint32_t i, int32_t lo, int32_t n) {
  const int32_t d = lo - (int32_t)i;
  if (d > -n && d <= n)  // within a box of i: no division
    return (int32_t)i + (d > 0 ? n : 0);
  const int32_t k = d >= 0 ? (d + n - 1) / n : -((-d) / n);
  return (int32_t)i + k * n;
}

// Does painting this sphere set wrapped node (ix, iy, iz)?
SIF_HD bool paint_hits(const PaintSphere& P, uint32_t ix, uint32_t iy, uint32_t iz) {
  for (int32_t gx = paint_first_image(ix, P.ox - P.c, P.n); gx <= P.ox + P.c; gx += P.n) {
    const float ddx = paint_dist(P, gx, P.cx);
    const float dx2 = ddx * ddx;
    if (dx2 > P.r2)
      continue;
    for (int32_t gy = paint_first_image(iy, P.oy - P.c, P.n); gy <= P.oy + P.c; gy += P.n) {
      const float ddy = paint_dist(P, gy, P.cy);
      const float dxy2 = dx2 + ddy * ddy;
      if (dxy2 > P.r2)
        continue;
      for (int32_t gz = paint_first_image(iz, P.oz - P.c, P.n); gz <= P.oz + P.c; gz += P.n) {
        const float ddz = paint_dist(P, gz, P.cz);
        if (dxy2 + ddz * ddz <= P.r2)
          return true;
      }
    }
  }
  return false;
}

}  // namespace sif_amd
