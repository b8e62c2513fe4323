// The voids accepted since a rung began, on the host, in a 32^3 cell list of
// their own. During a rung the device holds everything older: the catalogue
// as the rung began (the overlap index) and the mask it painted. What the
// commit must still ask of the new voids -- do they collide with a sphere
// (the overlap state of overlap.h), do they paint a candidate's node or one
// of its six pole nodes (the mask and the pole probe of
// sif__overlap_quick, src/finder/utils.c:819-848) -- probe() answers in one
// pass over the cells around the candidate.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

#include "overlap.h"
#include "paint.h"

namespace sif_amd {

struct RungProbe {
  int state = 0;      // overlap state (overlap.h) of the new voids
  bool node = false;  // a new void paints the candidate's node
  bool pole = false;  // a new void paints one of its six pole nodes
};

struct RungVoids {
  std::vector<float> cx, cy, cz, r;
  std::vector<int32_t> head, next;
  std::vector<uint32_t> touched;
  std::vector<uint64_t> occupied;  // one bit per cell: a probe skips empty cells
  float box = 0.0f, inv_l = 0.0f, max_r = 0.0f;
  void init(float box_) {
    box = box_;
    inv_l = (float)VOID_CELLS / box;
    head.assign(VOID_CELLS * VOID_CELLS * VOID_CELLS, -1);
    occupied.assign(VOID_CELLS * VOID_CELLS * VOID_CELLS / 64, 0ull);
  }
  void clear() {
    for (uint32_t c : touched) {
      head[c] = -1;
      occupied[c >> 6] = 0ull;
    }
    touched.clear();
    cx.clear();
    cy.clear();
    cz.clear();
    r.clear();
    next.clear();
    max_r = 0.0f;
  }
  void add(float x, float y, float z, float rr) {
    const int32_t k = (int32_t)cx.size();
    cx.push_back(x);
    cy.push_back(y);
    cz.push_back(z);
    r.push_back(rr);
    const uint32_t c =
      (void_axis_cell(x, inv_l) * VOID_CELLS + void_axis_cell(y, inv_l)) * VOID_CELLS +
      void_axis_cell(z, inv_l);
    if (head[c] == -1) {
      touched.push_back(c);
      occupied[c >> 6] |= 1ull << (c & 63);
    }
    next.push_back(head[c]);
    head[c] = k;
    if (rr > max_r)
      max_r = rr;
  }

  // The candidate sits on node (ix, iy, iz) of an n^3 grid of spacing l, at
  // (x, y, z). q: the overlap query, or null for none. want_node: test the
  // node's own mask bit. pole > 0: test the six nodes pole away along the
  // axes. `scale` is the painting scale of a void's radius.
  RungProbe probe(const OverlapQuery* q, uint32_t ix, uint32_t iy, uint32_t iz, float x,
    float y, float z, bool want_node, int32_t pole, float scale, float l, uint32_t n) const {
    RungProbe out;
    if (cx.empty())
      return out;
    const bool want_pole = pole > 0;
    // reach: the overlap superset, and painting (a void paints only nodes
    // within scale r + 2 l of its centre on each axis) around the node and
    // its poles
    int32_t lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
    if (q)
      superset_scan(*q, max_r, inv_l, lo, hi);
    if (want_node || want_pole) {
      const float paint_reach = scale * max_r + 2.0f * l + (want_pole ? (float)pole * l : 0.0f);
      const float s = paint_reach + paint_reach * 1e-5f + box * 1e-4f;
      const float c3[3] = {x, y, z};
      for (int a = 0; a < 3; a++) {
        int32_t plo = (int32_t)floorf((c3[a] - s) * inv_l);
        int32_t phi = (int32_t)floorf((c3[a] + s) * inv_l);
        if (hi[a] < lo[a]) {
          lo[a] = plo;
          hi[a] = phi;
        } else {
          lo[a] = plo < lo[a] ? plo : lo[a];
          hi[a] = phi > hi[a] ? phi : hi[a];
        }
        if (hi[a] - lo[a] >= VOID_CELLS) {
          lo[a] = 0;
          hi[a] = VOID_CELLS - 1;
        }
      }
    }
    const uint32_t px[2] = {paint_wrap((int32_t)ix + pole, (int32_t)n),
      paint_wrap((int32_t)ix - pole, (int32_t)n)};
    const uint32_t py[2] = {paint_wrap((int32_t)iy + pole, (int32_t)n),
      paint_wrap((int32_t)iy - pole, (int32_t)n)};
    const uint32_t pz[2] = {paint_wrap((int32_t)iz + pole, (int32_t)n),
      paint_wrap((int32_t)iz - pole, (int32_t)n)};
    for (int32_t a = lo[0]; a <= hi[0]; a++) {
      const uint32_t wa = void_wrap(a);
      for (int32_t b = lo[1]; b <= hi[1]; b++) {
        const uint32_t wb = void_wrap(b);
        for (int32_t c = lo[2]; c <= hi[2]; c++) {
          const uint32_t wc = void_wrap(c);
          const uint32_t cell = (wa * VOID_CELLS + wb) * VOID_CELLS + wc;
          if (!((occupied[cell >> 6] >> (cell & 63)) & 1ull))
            continue;
          for (int32_t i = head[cell]; i != -1; i = next[i]) {
            if (q && out.state != 1) {
              const int s = pair_state(*q, wa, wb, wc, cx[i], cy[i], cz[i], r[i]);
              out.state = s == 1 ? 1 : (out.state | s);
            }
            if ((want_node && !out.node) || (want_pole && !out.pole)) {
              PaintSphere P;
              paint_sphere(P, cx[i], cy[i], cz[i], scale * r[i], l, n);
              if (want_node && !out.node)
                out.node = paint_hits(P, ix, iy, iz);
              if (want_pole && !out.pole)
                out.pole = paint_hits(P, px[0], iy, iz) || paint_hits(P, px[1], iy, iz) ||
                           paint_hits(P, ix, py[0], iz) || paint_hits(P, ix, py[1], iz) ||
                           paint_hits(P, ix, iy, pz[0]) || paint_hits(P, ix, iy, pz[1]);
            }
          }
        }
      }
    }
    return out;
  }
};

}  // namespace sif_amd
