// The voids accepted since a rung began, on the host, in a 32^3 cell list of
// their own. During a rung the device holds everything older: the catalogue
// as the rung began (the overlap index) and the mask it painted. What the
// commit must still ask of the new voids -- do they collide with a sphere
// (the overlap state of overlap.h), do they paint a candidate's node or one
// of its six pole nodes (the mask and the pole probe of
// sif__overlap_quick, src/finder/utils.c:819-848) -- probe() answers in one
// pass over the cells around the candidate. walk() visits the voids of any
// such reach one by one, for callers that need each void's answer.
#pragma once
#include <algorithm>
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

// The six nodes `pole` away from (ix, iy, iz) along the axes, wrapped.
struct PoleNodes {
  uint32_t px[2], py[2], pz[2];
  SIF_HD void set(uint32_t ix, uint32_t iy, uint32_t iz, int32_t pole, uint32_t n) {
    px[0] = paint_wrap((int32_t)ix + pole, (int32_t)n);
    px[1] = paint_wrap((int32_t)ix - pole, (int32_t)n);
    py[0] = paint_wrap((int32_t)iy + pole, (int32_t)n);
    py[1] = paint_wrap((int32_t)iy - pole, (int32_t)n);
    pz[0] = paint_wrap((int32_t)iz + pole, (int32_t)n);
    pz[1] = paint_wrap((int32_t)iz - pole, (int32_t)n);
  }
  SIF_HD bool painted(const PaintSphere& P, uint32_t ix, uint32_t iy, uint32_t iz) const {
    return paint_hits(P, px[0], iy, iz) || paint_hits(P, px[1], iy, iz) ||
           paint_hits(P, ix, py[0], iz) || paint_hits(P, ix, py[1], iz) ||
           paint_hits(P, ix, iy, pz[0]) || paint_hits(P, ix, iy, pz[1]);
  }
};

// Cells of the 32^3 list, per axis before wrapping: a union of reaches.
struct CellRegion {
  int32_t lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
  SIF_HD bool empty() const { return hi[0] < lo[0]; }
  SIF_HD void merge(const int32_t l3[3], const int32_t h3[3]) {
    for (int a = 0; a < 3; a++) {
      if (hi[a] < lo[a]) {
        lo[a] = l3[a];
        hi[a] = h3[a];
      } else {
        lo[a] = l3[a] < lo[a] ? l3[a] : lo[a];
        hi[a] = h3[a] > hi[a] ? h3[a] : hi[a];
      }
      if (hi[a] - lo[a] >= VOID_CELLS) {  // the whole axis, each cell once
        lo[a] = 0;
        hi[a] = VOID_CELLS - 1;
      }
    }
  }
  // every void whose overlap with q may hold, for a set of largest radius
  // set_max_r (overlap.h)
  SIF_HD void add_overlap(const OverlapQuery& q, float set_max_r, float inv_l) {
    int32_t l3[3], h3[3];
    superset_scan(q, set_max_r, inv_l, l3, h3);
    merge(l3, h3);
  }
  // every void that may paint a node within `reach` of (x, y, z) on each
  // axis, widened by a margin far above float rounding
  SIF_HD void add_paint(float x, float y, float z, float reach, float box, float inv_l) {
    const float s = reach + reach * 1e-5f + box * 1e-4f;
    const float c3[3] = {x, y, z};
    int32_t l3[3], h3[3];
    for (int a = 0; a < 3; a++) {
      l3[a] = (int32_t)floorf((c3[a] - s) * inv_l);
      h3[a] = (int32_t)floorf((c3[a] + s) * inv_l);
    }
    merge(l3, h3);
  }
  // The region's wrapped cells on axis a as the bits of a word: holds()
  // tests cell w of that axis as bit w.
  SIF_HD uint32_t axis_bits(int a) const {
    if (empty())
      return 0u;
    const uint32_t span = (uint32_t)(hi[a] - lo[a]);
    if (span >= (uint32_t)(VOID_CELLS - 1))
      return 0xffffffffu;
    const uint32_t run = (2u << span) - 1u;  // span + 1 cells from lo
    const uint32_t s = void_wrap(lo[a]);
    return s ? (run << s) | (run >> (32u - s)) : run;
  }
  // Does the region hold wrapped cell (wa, wb, wc)? The same cells as a scan
  // of the region visits.
  SIF_HD bool holds(uint32_t wa, uint32_t wb, uint32_t wc) const {
    if (empty())
      return false;
    const uint32_t w[3] = {wa, wb, wc};
    for (int a = 0; a < 3; a++)
      if (((w[a] - void_wrap(lo[a])) & (uint32_t)(VOID_CELLS - 1)) > (uint32_t)(hi[a] - lo[a]))
        return false;
    return true;
  }
};

struct RungVoids {
  // one cache line's half per void: what a probe reads of it
  struct alignas(32) Rec {
    float x, y, z, r;
    int32_t next;   // the next void of its cell, or -1
    uint32_t cell;  // its cell, (wx 32 + wy) 32 + wz
  };
  std::vector<float> cx, cy, cz, r;  // the same voids as columns, for upload
  bool columns = true;                // kept only when asked for
  std::vector<Rec> rec;
  // Per cell, its first void as (generation << 32 | index), and per row of
  // 32 cells along z, the cells holding a void as (generation << 32 | bits):
  // an entry of an earlier generation stands for an empty cell or row, so
  // clear() starts a generation instead of rewriting the cells (lines that
  // other cores have just read in probes).
  std::vector<uint64_t> head;
  std::vector<uint64_t> rows;
  uint32_t gen = 1;
  float box = 0.0f, inv_l = 0.0f, max_r = 0.0f;
  void init(float box_) {
    box = box_;
    inv_l = (float)VOID_CELLS / box;
    head.assign(VOID_CELLS * VOID_CELLS * VOID_CELLS, 0ull);
    rows.assign(VOID_CELLS * VOID_CELLS, 0ull);
    gen = 1;
  }
  size_t size() const { return rec.size(); }
  // a cell's first void, or -1
  int32_t first(uint32_t cell) const {
    const uint64_t e = head[cell];
    return (uint32_t)(e >> 32) == gen ? (int32_t)(uint32_t)e : -1;
  }
  // a row's cells holding a void, as bits
  uint32_t row_cells(uint32_t row) const {
    const uint64_t e = rows[row];
    return (uint32_t)(e >> 32) == gen ? (uint32_t)e : 0u;
  }
  void clear() {
    if (++gen == 0) {  // after 2^32 - 1 generations: every entry reads empty again
      std::fill(head.begin(), head.end(), 0ull);
      std::fill(rows.begin(), rows.end(), 0ull);
      gen = 1;
    }
    cx.clear();
    cy.clear();
    cz.clear();
    r.clear();
    rec.clear();
    max_r = 0.0f;
  }
  void add(float x, float y, float z, float rr) {
    const int32_t k = (int32_t)rec.size();
    if (columns) {
      cx.push_back(x);
      cy.push_back(y);
      cz.push_back(z);
      r.push_back(rr);
    }
    const uint32_t c =
      (void_axis_cell(x, inv_l) * VOID_CELLS + void_axis_cell(y, inv_l)) * VOID_CELLS +
      void_axis_cell(z, inv_l);
    const uint64_t g = (uint64_t)gen << 32;
    rows[c >> 5] = g | row_cells(c >> 5) | (1u << (c & 31u));
    rec.push_back(Rec{x, y, z, rr, first(c), c});
    head[c] = g | (uint32_t)k;
    if (rr > max_r)
      max_r = rr;
  }

  // Appends n voids at once, get(i, &x, &y, &z, &r) giving the i-th, over
  // run(n, fn(begin, end)), a pool's parallel for. Each void's record and
  // columns go to the slot its place fixes, and each cell's list is linked by
  // an atomic exchange of the cell's head, so a cell lists its voids in no
  // fixed order (walk() visits them in no particular order anyway). Added one
  // by one, each void's writes go to lines that other cores have just read
  // in probes, a transfer between cores apiece; spread over the pool, they
  // overlap.
  template <class Get, class Run>
  void add_many(size_t n, Get get, Run run) {
    if (n == 0)
      return;
    const size_t base = rec.size();
    rec.resize(base + n);
    if (columns) {
      cx.resize(base + n);
      cy.resize(base + n);
      cz.resize(base + n);
      r.resize(base + n);
    }
    run(n, [&](uint64_t b0, uint64_t b1) {
      for (uint64_t i = b0; i < b1; i++) {
        float x, y, z, rr;
        get(i, &x, &y, &z, &rr);
        const size_t k = base + i;
        if (columns) {
          cx[k] = x;
          cy[k] = y;
          cz[k] = z;
          r[k] = rr;
        }
        const uint32_t c =
          (void_axis_cell(x, inv_l) * VOID_CELLS + void_axis_cell(y, inv_l)) * VOID_CELLS +
          void_axis_cell(z, inv_l);
        const uint64_t g = (uint64_t)gen << 32;
        const uint64_t prev = __atomic_exchange_n(&head[c], g | (uint32_t)k, __ATOMIC_RELAXED);
        rec[k] = Rec{x, y, z, rr, (uint32_t)(prev >> 32) == gen ? (int32_t)(uint32_t)prev : -1, c};
        // the row's bit, the row's word restarted if it is of an earlier
        // generation
        uint64_t old = __atomic_load_n(&rows[c >> 5], __ATOMIC_RELAXED);
        for (;;) {
          const uint64_t want =
            ((uint32_t)(old >> 32) == gen ? old : g) | (1ull << (c & 31u));
          if (want == old ||
              __atomic_compare_exchange_n(&rows[c >> 5], &old, want, true, __ATOMIC_RELAXED,
                __ATOMIC_RELAXED))
            break;
        }
      }
    });
    for (size_t k = base; k < base + n; k++)
      if (rec[k].r > max_r)
        max_r = rec[k].r;
  }

  // Calls visit(void index, record, wa, wb, wc) for every void stored in the
  // region's cells (wrapped cell coordinates wa, wb, wc), in no particular
  // order, until visit returns false. A list shorter than the region's rows
  // is walked directly, testing each void's cell against the region, which
  // visits the same voids as the scan of the cells.
  template <class Visit>
  void walk(const CellRegion& R, Visit visit) const {
    if (rec.empty() || R.empty())
      return;
    const uint32_t span[3] = {(uint32_t)(R.hi[0] - R.lo[0]), (uint32_t)(R.hi[1] - R.lo[1]),
      (uint32_t)(R.hi[2] - R.lo[2])};
    const uint32_t rows = (span[0] + 1) * (span[1] + 1);
    if (rec.size() <= rows) {
      const uint32_t w0[3] = {void_wrap(R.lo[0]), void_wrap(R.lo[1]), void_wrap(R.lo[2])};
      for (size_t i = 0; i < rec.size(); i++) {
        const Rec& v = rec[i];
        const uint32_t wa = v.cell >> 10, wb = (v.cell >> 5) & 31u, wc = v.cell & 31u;
        if (((wa - w0[0]) & 31u) > span[0] || ((wb - w0[1]) & 31u) > span[1] ||
            ((wc - w0[2]) & 31u) > span[2])
          continue;
        if (!visit((uint32_t)i, v, wa, wb, wc))
          return;
      }
      return;
    }
    // the z range as a mask over a row's 32 cells (wrapping)
    uint32_t zmask = 0xffffffffu;
    if (span[2] + 1 < (uint32_t)VOID_CELLS) {
      const uint64_t m = ((1ull << (span[2] + 1)) - 1ull) << void_wrap(R.lo[2]);
      zmask = (uint32_t)(m | (m >> VOID_CELLS));
    }
    for (int32_t a = R.lo[0]; a <= R.hi[0]; a++) {
      const uint32_t wa = void_wrap(a);
      for (int32_t b = R.lo[1]; b <= R.hi[1]; b++) {
        const uint32_t wb = void_wrap(b);
        // a row of 32 cells is half an occupancy word: only its set cells
        const uint32_t row = wa * VOID_CELLS + wb;
        uint32_t cells = row_cells(row) & zmask;
        while (cells) {
          const uint32_t wc = (uint32_t)__builtin_ctz(cells);
          cells &= cells - 1u;
          const uint32_t cell = row * VOID_CELLS + wc;
          for (int32_t i = first(cell); i != -1; i = rec[i].next)
            if (!visit((uint32_t)i, rec[i], wa, wb, wc))
              return;
        }
      }
    }
  }

  // The candidate sits on node (ix, iy, iz) of an n^3 grid of spacing l, at
  // (x, y, z). q: the overlap query, or null for none. want_node: test the
  // node's own mask bit. pole > 0: test the six nodes pole away along the
  // axes. `scale` is the painting scale of a void's radius.
  //
  // The caller refuses on the first of node, pole, overlap that holds, so
  // the probe stops once its answer can no longer change that: at a node hit,
  // or at a pole hit when the node is not asked; after a pole hit it no
  // longer needs the state. Every test is an OR over voids, so the voids may
  // be visited in any order.
  RungProbe probe(const OverlapQuery* q, uint32_t ix, uint32_t iy, uint32_t iz, float x,
    float y, float z, bool want_node, int32_t pole, float scale, float l, uint32_t n) const {
    RungProbe out;
    if (rec.empty())
      return out;
    const bool want_pole = pole > 0;
    if (!q && !want_node && !want_pole)
      return out;
    // reach: the overlap superset, and painting (a void paints only nodes
    // within scale r + 2 l of its centre on each axis) around the node and
    // its poles
    CellRegion R;
    if (q)
      R.add_overlap(*q, max_r, inv_l);
    if (want_node || want_pole)
      R.add_paint(x, y, z, scale * max_r + 2.0f * l + (want_pole ? (float)pole * l : 0.0f),
        box, inv_l);
    PoleNodes poles;
    bool poles_ready = false;  // wrapped when a first void needs them
    walk(R, [&](uint32_t, const Rec& v, uint32_t wa, uint32_t wb, uint32_t wc) {
      if (q && out.state != 1 && !(want_pole && out.pole)) {
        const int s = pair_state(*q, wa, wb, wc, v.x, v.y, v.z, v.r);
        out.state = s == 1 ? 1 : (out.state | s);
      }
      if ((want_node && !out.node) || (want_pole && !out.pole)) {
        PaintSphere P;
        paint_sphere(P, v.x, v.y, v.z, scale * v.r, l, n);
        if (want_node && !out.node)
          out.node = paint_hits(P, ix, iy, iz);
        if (want_pole && !out.pole) {
          if (!poles_ready) {
            poles.set(ix, iy, iz, pole, n);
            poles_ready = true;
          }
          out.pole = poles.painted(P, ix, iy, iz);
        }
      }
      return !(out.node || (out.pole && !want_node));
    });
    return out;
  }
};

}  // namespace sif_amd
