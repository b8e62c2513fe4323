// The commit of a batch of grown candidates (exodus_finder.c:2174-2329) takes
// them in order, and a candidate accepted joins the catalogue that every
// later one is tested against. Every test is an OR over the catalogue's
// voids, and a candidate that is accepted becomes a void of the radius it
// was grown to, known before the commit begins. So what an earlier candidate
// a of the batch would do to a later candidate b, were a accepted, can be
// found for every pair at once (on the host pool): the commit then only ORs
// the effects of the candidates it has accepted.
//
// pair_effects() lists, for candidate b, every earlier candidate a whose
// void would paint b's node or one of its pole nodes, at the rung's pole
// distance or at b's own, or whose overlap state (overlap.h) with b, at the
// rung radius or at b's grown radius, is not 0. The overlap states are taken
// with the catalogue's largest radius as the batch began; one that grows
// before b is tested only turns a state 2 into a 1 (overlap.h), and a 2 sends
// b to the reference's own scan, which decides it at the radius of the day.
#pragma once
#include <cstdint>

#include "rung_voids.h"

namespace sif_amd {

// What candidate b still has to learn from the batch's earlier candidates.
enum : uint8_t {
  NEED_NODE = 1,     // does one paint its node
  NEED_POLE = 2,     // ... one of its poles at the rung's pole distance
  NEED_STATE = 4,    // overlap state at the rung radius
  NEED_POLE_R = 8,   // ... one of its poles at its grown radius's distance
  NEED_STATE_R = 16  // overlap state at its grown radius
};

// One earlier candidate's effect on b.
struct PairEffect {
  uint32_t a;       // its place in the batch
  uint8_t node : 1, pole : 1, pole_r : 1;
  uint8_t state : 2, state_r : 2;
};

// A candidate's effects count up to BATCH_EFF_FULL, which stands for "that
// many or more": such a candidate, or one whose effects run past the batch's
// list, probes the commit's accepted voids instead.
constexpr uint32_t BATCH_EFF_FULL = 255;
// The batch's list holds this many effects a candidate on average (a few
// times what a batch of Zel'dovich candidates needs).
constexpr uint32_t BATCH_EFF_PER = 4;

// What batch_pairs() needs of the run and the rung.
struct BatchPairParams {
  uint32_t n;        // grid side
  float l, box, f, scale;
  float radius;      // the rung's
  float cat_max_r;   // the catalogue's largest radius as the batch is launched
  float inv_l;       // VOID_CELLS / box
  int32_t proxy;     // the rung's pole distance in nodes
};

// A candidate's void cell (overlap.h), (wa 32 + wb) 32 + wc, as the pair
// scan's region test reads it; BATCH_NO_CELL for one that did not grow, whose
// void would paint and reach nothing.
constexpr uint32_t BATCH_NO_CELL = 0xffffffffu;
SIF_HD uint32_t batch_cell_code(uint32_t node, float r, const BatchPairParams& P) {
  if (!(r >= 0.0f))
    return BATCH_NO_CELL;
  const uint32_t n = P.n, nn = n * n;
  const uint32_t jx = node / nn, jy = (node / n) % n, jz = node % n;
  const uint32_t wa = void_axis_cell((float)jx * P.l, P.inv_l),
                 wb = void_axis_cell((float)jy * P.l, P.inv_l),
                 wc = void_axis_cell((float)jz * P.l, P.inv_l);
  return (wa << 10) | (wb << 5) | wc;
}

// Device scratch and results of batch_pairs(): per candidate its count
// (cnt8, read by the host) and offset into eff (the exclusive sum of the
// counts in batch order, which the host can sum again from cnt8).
struct BatchPairScratch {
  float* set_max_r;   // 1: the batch's largest grown radius
  uint32_t* cells;    // count: batch_cell_code
  uint32_t* cnt32;    // count
  uint8_t* cnt8;      // count
  uint64_t* offset;   // count
  PairEffect* eff;    // capacity
  uint64_t capacity;
};

// For each of `count` grown candidates (device: flat nodes, grown radii, the
// grown-overlap states), its effects from the earlier ones (batch_pairs.hip).
int batch_pairs(const uint32_t* flat, const float* rad, const uint32_t* grown_state,
  uint32_t count, const BatchPairParams& P, BatchPairScratch& W);

struct PairQuery {
  uint32_t b;              // the candidate's place in the batch
  uint32_t ix, iy, iz;     // its node
  float x, y, z;           // and centre
  float r;                 // its grown radius (>= 0 when NEED_*_R is set)
  int32_t pole, pole_r;    // pole distances, rung and grown, in nodes
  uint8_t need;
};

// The batch's grown candidates as voids (`index`, in batch order, `place`[k]
// the batch place of void k): appends the effects on q of those before it to
// out[0 .. cap) and returns their number, or -1 when they do not fit.
inline int pair_effects(const RungVoids& index, const uint32_t* place, const PairQuery& c,
  float radius, float cat_max_r, float f, float scale, float l, uint32_t n, PairEffect* out,
  int cap) {
  if (!c.need || index.size() == 0)
    return 0;
  const bool want_node = c.need & NEED_NODE, want_pole = (c.need & NEED_POLE) && c.pole > 0;
  const bool want_pole_r = (c.need & NEED_POLE_R) && c.pole_r > 0;
  OverlapQuery q, qr;
  CellRegion R;
  if (c.need & NEED_STATE) {
    overlap_query(q, c.x, c.y, c.z, radius, cat_max_r, index.box, f, scale, index.inv_l);
    R.add_overlap(q, index.max_r, index.inv_l);
  }
  if (c.need & NEED_STATE_R) {
    overlap_query(qr, c.x, c.y, c.z, c.r, cat_max_r, index.box, f, scale, index.inv_l);
    R.add_overlap(qr, index.max_r, index.inv_l);
  }
  const float paint = scale * index.max_r + 2.0f * l;
  if (want_node || want_pole)
    R.add_paint(c.x, c.y, c.z, paint + (want_pole ? (float)c.pole * l : 0.0f), index.box,
      index.inv_l);
  if (want_pole_r)
    R.add_paint(c.x, c.y, c.z, paint + (float)c.pole_r * l, index.box, index.inv_l);
  PoleNodes poles, poles_r;
  bool poles_ready = false;  // wrapped when a first void needs them
  int count = 0;
  index.walk(R, [&](uint32_t k, const RungVoids::Rec& v, uint32_t wa, uint32_t wb,
                  uint32_t wc) {
    if (place[k] >= c.b)
      return true;
    PairEffect e{};
    e.a = place[k];
    if (c.need & NEED_STATE)
      e.state = (uint8_t)pair_state(q, wa, wb, wc, v.x, v.y, v.z, v.r);
    if (c.need & NEED_STATE_R)
      e.state_r = (uint8_t)pair_state(qr, wa, wb, wc, v.x, v.y, v.z, v.r);
    if (want_node || want_pole || want_pole_r) {
      if (!poles_ready) {
        if (want_pole)
          poles.set(c.ix, c.iy, c.iz, c.pole, n);
        if (want_pole_r)
          poles_r.set(c.ix, c.iy, c.iz, c.pole_r, n);
        poles_ready = true;
      }
      PaintSphere P;
      paint_sphere(P, v.x, v.y, v.z, scale * v.r, l, n);
      e.node = want_node && paint_hits(P, c.ix, c.iy, c.iz);
      e.pole = want_pole && poles.painted(P, c.ix, c.iy, c.iz);
      e.pole_r = want_pole_r && poles_r.painted(P, c.ix, c.iy, c.iz);
    }
    if (e.node | e.pole | e.pole_r | e.state | e.state_r) {
      if (count == cap) {
        count = -1;
        return false;
      }
      out[count++] = e;
    }
    return true;
  });
  return count;
}

}  // namespace sif_amd
