// TEST INFRASTRUCTURE ONLY: the painting predicate (sif_amd/csrc/paint.h) and
// the per-rung void probe (sif_amd/csrc/rung_voids.h) built for the CPU, for
// comparison with the oracle's painting and overlap scan.
#include <cstdint>
#include <cstring>

#include "paint.h"
#include "rung_voids.h"

using namespace sif_amd;

// The mask every node of which is decided by paint_hits: all nodes when
// `all_nodes`, else the nodes of each sphere's cube (the only ones it can
// paint).
extern "C" void paint_nodes(uint32_t n, float l, uint64_t count, const float* cx,
  const float* cy, const float* cz, const float* r, int all_nodes, uint64_t* words) {
  const uint64_t cells = (uint64_t)n * n * n;
  memset(words, 0, (cells + 63) / 64 * sizeof(uint64_t));
  for (uint64_t s = 0; s < count; s++) {
    PaintSphere P;
    paint_sphere(P, cx[s], cy[s], cz[s], r[s], l, n);
    auto test = [&](uint32_t ix, uint32_t iy, uint32_t iz) {
      if (paint_hits(P, ix, iy, iz)) {
        const uint64_t i = ((uint64_t)ix * n + iy) * n + iz;
        words[i >> 6] |= 1ull << (i & 63);
      }
    };
    if (all_nodes) {
      for (uint32_t ix = 0; ix < n; ix++)
        for (uint32_t iy = 0; iy < n; iy++)
          for (uint32_t iz = 0; iz < n; iz++)
            test(ix, iy, iz);
    } else {
      for (int32_t gx = P.ox - P.c; gx <= P.ox + P.c; gx++)
        for (int32_t gy = P.oy - P.c; gy <= P.oy + P.c; gy++)
          for (int32_t gz = P.oz - P.c; gz <= P.oz + P.c; gz++)
            test(paint_wrap(gx, P.n), paint_wrap(gy, P.n), paint_wrap(gz, P.n));
    }
  }
}

// RungVoids::probe for nq candidate nodes against nv voids. qr[q] < 0: no
// overlap query. want_node/pole as probe() takes them.
extern "C" void rv_probe_many(uint64_t nv, const float* vx, const float* vy, const float* vz,
  const float* vr, float box, uint32_t n, float scale, float f, float ref_max_r, uint64_t nq,
  const uint32_t* qix, const uint32_t* qiy, const uint32_t* qiz, const float* qr,
  const uint8_t* want_node, const int32_t* pole, uint8_t* out_state, uint8_t* out_node,
  uint8_t* out_pole) {
  RungVoids RV;
  RV.init(box);
  for (uint64_t v = 0; v < nv; v++)
    RV.add(vx[v], vy[v], vz[v], vr[v]);
  const float l = box / (float)n;
  const float inv_l = (float)VOID_CELLS / box;
  for (uint64_t k = 0; k < nq; k++) {
    const float x = (float)qix[k] * l, y = (float)qiy[k] * l, z = (float)qiz[k] * l;
    OverlapQuery q;
    const bool want_q = qr[k] >= 0.0f;
    if (want_q)
      overlap_query(q, x, y, z, qr[k], ref_max_r, box, f, scale, inv_l);
    const RungProbe p = RV.probe(want_q ? &q : nullptr, qix[k], qiy[k], qiz[k], x, y, z,
      want_node[k] != 0, pole[k], scale, l, n);
    out_state[k] = (uint8_t)p.state;
    out_node[k] = p.node ? 1 : 0;
    out_pole[k] = p.pole ? 1 : 0;
  }
}
