// TEST INFRASTRUCTURE ONLY: the overlap states of sif_amd/csrc/overlap.h on
// the CPU, over a cell list built the way the device index is (each void
// under the reference's cell), for comparison with the oracle's restatement
// of the reference scan.
#include <cstdint>
#include <vector>

#include "overlap.h"

using namespace sif_amd;

// n_classes 1: one list; 4: the radius classes of the device index (a
// quarter, a half, one void-cell width, and more), each scanned with its own
// largest radius.
extern "C" void ovl_states(const float* vcx, const float* vcy, const float* vcz,
  const float* vr, uint64_t nv, float ref_max_r, const float* qx, const float* qy,
  const float* qz, const float* qr, uint64_t nq, float box, float f, float scale,
  int n_classes, uint8_t* out) {
  const float inv_l = (float)VOID_CELLS / box;
  const float cell_w = box / (float)VOID_CELLS;
  const uint32_t ncell = VOID_CELLS * VOID_CELLS * VOID_CELLS;
  std::vector<std::vector<uint32_t>> cells((size_t)n_classes * ncell);
  std::vector<float> cmax(n_classes, -1.0f);
  for (uint64_t k = 0; k < nv; k++) {
    const float r = vr[k];
    const int cls = n_classes == 1 ? 0
                    : r < 0.25f * cell_w ? 0
                    : r < 0.5f * cell_w  ? 1
                    : r < cell_w         ? 2
                                         : 3;
    const uint32_t c = (void_axis_cell(vcx[k], inv_l) * VOID_CELLS +
                        void_axis_cell(vcy[k], inv_l)) * VOID_CELLS +
                       void_axis_cell(vcz[k], inv_l);
    cells[(size_t)cls * ncell + c].push_back((uint32_t)k);
    if (r > cmax[cls])
      cmax[cls] = r;
  }
  for (uint64_t q = 0; q < nq; q++) {
    OverlapQuery Q;
    overlap_query(Q, qx[q], qy[q], qz[q], qr[q], ref_max_r, box, f, scale, inv_l);
    int st = 0;
    for (int cls = 0; cls < n_classes && st != 1; cls++) {
      if (cmax[cls] < 0.0f)
        continue;
      int32_t lo[3], hi[3];
      superset_scan(Q, cmax[cls], inv_l, lo, hi);
      for (int32_t a = lo[0]; a <= hi[0] && st != 1; a++)
        for (int32_t b = lo[1]; b <= hi[1] && st != 1; b++)
          for (int32_t c = lo[2]; c <= hi[2] && st != 1; c++) {
            const uint32_t wa = void_wrap(a), wb = void_wrap(b), wc = void_wrap(c);
            for (uint32_t i : cells[(size_t)cls * ncell + (wa * VOID_CELLS + wb) * VOID_CELLS + wc]) {
              const int s = pair_state(Q, wa, wb, wc, vcx[i], vcy[i], vcz[i], vr[i]);
              if (s == 1) {
                st = 1;
                break;
              }
              st |= s;
            }
          }
    }
    out[q] = (uint8_t)st;
  }
}
