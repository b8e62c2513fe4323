// TEST INFRASTRUCTURE ONLY: the host mask of sif_amd/csrc/mask.h built for
// comparison with the oracle's node-by-node painting.
#include <cstdint>
#include <cstring>

#include "mask.h"

using namespace sif_amd;

extern "C" void mask_paint(uint32_t n, float l, uint64_t count, const float* cx,
  const float* cy, const float* cz, const float* r, uint64_t* words) {
  Mask m;
  m.init(n);
  for (uint64_t i = 0; i < count; i++)
    mark_sphere(m, cx[i], cy[i], cz[i], r[i], l);
  memcpy(words, m.words.data(), m.words.size() * sizeof(uint64_t));
}

extern "C" int mask_quick(uint32_t n, const uint64_t* words, uint32_t ix, uint32_t iy,
  uint32_t iz, uint32_t r) {
  Mask m;
  m.n = n;
  m.words.assign(words, words + ((uint64_t)n * n * n + 63) / 64);
  return overlap_quick(m, ix, iy, iz, r) ? 1 : 0;
}
