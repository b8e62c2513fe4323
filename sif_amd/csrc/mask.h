// The host mask of accepted-void nodes, n^3 bits in flat (x, y, z) order: the
// pole probe of sif__overlap_quick (src/finder/utils.c:819-848) and the
// painting of sif__mark_sphere (utils.c:965-1033).
//
// The reference paints by testing every node of the sphere's bounding cube.
// Here each (x, y) row is painted as one run of z nodes: within the cube the
// z distance |z l - cz| falls to its smallest value and rises again, so the
// nodes that pass the reference's own test form one run. Its two ends are
// found by stepping with that same test from the continuous estimate, and its
// bits are set a word at a time. This holds while no z distance of the cube
// reaches half the box (the periodic fold would bend it); a sphere that wide
// is painted node by node, as the reference does. Same float operations, same
// bits.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace sif_amd {

inline uint32_t wrap_n(int32_t v, uint32_t n) {
  const int32_t w = v % (int32_t)n;
  return (uint32_t)(w < 0 ? w + (int32_t)n : w);
}

struct Mask {
  std::vector<uint64_t> words;
  uint32_t n = 0;
  bool dirty = true;
  void init(uint32_t n_) {
    n = n_;
    words.assign(((uint64_t)n * n * n + 63) / 64, 0ull);
    dirty = true;
  }
  bool get(uint64_t i) const { return (words[i >> 6] >> (i & 63)) & 1ull; }
  void set(uint64_t i) { words[i >> 6] |= 1ull << (i & 63); }
  // bits a..b inclusive
  void set_range(uint64_t a, uint64_t b) {
    const uint64_t wa = a >> 6, wb = b >> 6;
    if (wa == wb) {
      words[wa] |= (~0ull >> (63 - (b - a))) << (a & 63);
      return;
    }
    words[wa] |= ~0ull << (a & 63);
    for (uint64_t w = wa + 1; w < wb; w++)
      words[w] = ~0ull;
    words[wb] |= ~0ull >> (63 - (b & 63));
  }
};

// utils.c:819-848
inline bool overlap_quick(const Mask& m, uint32_t ix, uint32_t iy, uint32_t iz, uint32_t r) {
  const uint64_t n = m.n;
  if (m.get(((uint64_t)wrap_n((int32_t)(ix + r), m.n) * n + iy) * n + iz))
    return true;
  if (m.get(((uint64_t)wrap_n((int32_t)ix - (int32_t)r, m.n) * n + iy) * n + iz))
    return true;
  if (m.get(((uint64_t)ix * n + wrap_n((int32_t)(iy + r), m.n)) * n + iz))
    return true;
  if (m.get(((uint64_t)ix * n + wrap_n((int32_t)iy - (int32_t)r, m.n)) * n + iz))
    return true;
  if (m.get(((uint64_t)ix * n + iy) * n + wrap_n((int32_t)(iz + r), m.n)))
    return true;
  if (m.get(((uint64_t)ix * n + iy) * n + wrap_n((int32_t)iz - (int32_t)r, m.n)))
    return true;
  return false;
}

// utils.c:965-1033, setting form: every node within r of the centre.
inline void mark_sphere(Mask& m, float cx, float cy, float cz, float r, float l) {
  const float r2 = r * r;
  const float inv_l = 1.0f / l;
  const float L = (float)m.n * l;
  const float half = L * 0.5f;
  const int32_t c = (int32_t)(r * inv_l) + 1;
  const int32_t ox = (int32_t)(cx * inv_l), oy = (int32_t)(cy * inv_l),
                oz = (int32_t)(cz * inv_l);
  const uint64_t n = m.n;
  // every z distance of the cube stays below half the box: rows are runs
  const bool runs = (float)(c + 2) * l < half;
  for (int32_t dx = -c; dx <= c; dx++) {
    const int32_t gx = ox + dx;
    float ddx = fabsf((float)gx * l - cx);
    ddx = ddx > half ? L - ddx : ddx;
    const float dx2 = ddx * ddx;
    if (dx2 > r2)
      continue;
    const uint32_t wx = wrap_n(gx, m.n);
    for (int32_t dy = -c; dy <= c; dy++) {
      const int32_t gy = oy + dy;
      float ddy = fabsf((float)gy * l - cy);
      ddy = ddy > half ? L - ddy : ddy;
      const float dxy2 = dx2 + ddy * ddy;
      if (dxy2 > r2)
        continue;
      const uint64_t row = ((uint64_t)wx * n + wrap_n(gy, m.n)) * n;
      if (!runs) {
        for (int32_t dz = -c; dz <= c; dz++) {
          const int32_t gz = oz + dz;
          float ddz = fabsf((float)gz * l - cz);
          ddz = ddz > half ? L - ddz : ddz;
          if (dxy2 + ddz * ddz <= r2)
            m.set(row + wrap_n(gz, m.n));
        }
        continue;
      }
      // the reference's test at node gz of this row (no fold: runs)
      auto inside = [&](int32_t gz) {
        const float ddz = fabsf((float)gz * l - cz);
        return dxy2 + ddz * ddz <= r2;
      };
      const int32_t zlo = oz - c, zhi = oz + c;
      // g0: the first node of the cube at or past the centre
      int32_t g0 = oz - 1 < zlo ? zlo : oz - 1;
      while (g0 > zlo && (float)g0 * l - cz >= 0.0f)
        g0--;
      while (g0 <= zhi && (float)g0 * l - cz < 0.0f)
        g0++;
      // each end of the run starts from the continuous estimate cz +- w and
      // steps with the exact test; the estimate only saves steps
      const float w = sqrtf(r2 - dxy2);
      int32_t a = 0, b = -1;  // the run, a > b when empty
      if (g0 <= zhi && inside(g0)) {  // rising side: last node inside
        int32_t g = (int32_t)floorf((cz + w) * inv_l);
        g = g < g0 ? g0 : g > zhi ? zhi : g;
        if (inside(g)) {
          while (g < zhi && inside(g + 1))
            g++;
        } else {
          while (!inside(g))  // ends at g0 at the latest
            g--;
        }
        a = g0;
        b = g;
      }
      if (g0 - 1 >= zlo && inside(g0 - 1)) {  // falling side: first node inside
        int32_t g = (int32_t)ceilf((cz - w) * inv_l);
        g = g > g0 - 1 ? g0 - 1 : g < zlo ? zlo : g;
        if (inside(g)) {
          while (g > zlo && inside(g - 1))
            g--;
        } else {
          while (!inside(g))  // ends at g0 - 1 at the latest
            g++;
        }
        if (b < a)
          b = g0 - 1;
        a = g;
      }
      if (b < a)
        continue;
      const uint64_t wa = wrap_n(a, m.n), len = (uint64_t)(b - a);
      if (wa + len < n) {
        m.set_range(row + wa, row + wa + len);
      } else {
        m.set_range(row + wa, row + n - 1);
        m.set_range(row, row + (wa + len - n));
      }
    }
  }
  m.dirty = true;
}

}  // namespace sif_amd
