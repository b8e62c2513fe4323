// The device text of the batch's pair effects (batch_pairs.hip launches it;
// tests/emu runs the same text on the CPU under a wave emulator). One
// wavefront per candidate b, its lanes over the earlier candidates a < b.
#pragma once
#include <cstdint>

#include "batch_pairs.h"

namespace sif_amd {

constexpr int BP_WAVES = 4;

// The lanes of a wave see each other's LDS stores made before it.
__device__ __forceinline__ void pairs_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// COUNT: each candidate's number of effects (at most BATCH_EFF_FULL, which
// stands for "that many or more"). WRITE: its effects, in the order of the
// earlier candidates, at its offset (the exclusive sum of the counts) in one
// list of `capacity` effects; a candidate at BATCH_EFF_FULL, or whose effects
// run past the capacity, writes none.
template <bool WRITE>
__global__ void __launch_bounds__(64 * BP_WAVES) batch_pairs_kernel(
  const uint32_t* __restrict__ flat, const float* __restrict__ rad,
  const uint32_t* __restrict__ grown_state, uint32_t count, BatchPairParams P,
  const float* __restrict__ set_max_r, const uint32_t* __restrict__ cells,
  uint32_t* __restrict__ cnt32,
  uint8_t* __restrict__ cnt8, const uint64_t* __restrict__ offset,
  PairEffect* __restrict__ eff, uint64_t capacity) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t b = blockIdx.x * BP_WAVES + (threadIdx.x >> 6);
  if (b >= count)
    return;
  uint64_t at0 = 0;
  if constexpr (WRITE) {
    at0 = offset[b];
    const uint32_t c = cnt8[b];
    if (c == 0 || c >= BATCH_EFF_FULL || at0 + c > capacity)
      return;
  }
  const uint32_t n = P.n, nn = n * n;
  const uint32_t fb = flat[b];
  const uint32_t ix = fb / nn, iy = (fb / n) % n, iz = fb % n;
  const float x = (float)ix * P.l, y = (float)iy * P.l, z = (float)iz * P.l;
  const float r = rad[b];
  const uint32_t gs = grown_state[b];
  const float bmax = *set_max_r;
  const int32_t pole_r = r >= 0.0f ? (int32_t)(P.scale * r * (1.0f - P.f) / P.l) - 1 : 0;
  // every test b may still need (batch_pairs.h NEED_*), as far as the
  // device can tell
  const bool want_node = true;
  const bool want_pole = P.proxy > 0;
  const bool need_state = true;
  const bool want_pole_r = r >= 0.0f && pole_r > 0 && !(gs & GROWN_POLE_BIT);
  const bool need_state_r = r >= 0.0f && (gs & 3u) == 0u;
  OverlapQuery q, qr;
  CellRegion R;
  if (bmax >= 0.0f) {
    if (need_state) {
      overlap_query(q, x, y, z, P.radius, P.cat_max_r, P.box, P.f, P.scale, P.inv_l);
      R.add_overlap(q, bmax, P.inv_l);
    }
    if (need_state_r) {
      overlap_query(qr, x, y, z, r, P.cat_max_r, P.box, P.f, P.scale, P.inv_l);
      R.add_overlap(qr, bmax, P.inv_l);
    }
    const float paint = P.scale * bmax + 2.0f * P.l;
    if (want_node || want_pole)
      R.add_paint(x, y, z, paint + (want_pole ? (float)P.proxy * P.l : 0.0f), P.box, P.inv_l);
    if (want_pole_r)
      R.add_paint(x, y, z, paint + (float)pole_r * P.l, P.box, P.inv_l);
  }
  PoleNodes poles, poles_r;
  if (want_pole)
    poles.set(ix, iy, iz, P.proxy, n);
  if (want_pole_r)
    poles_r.set(ix, iy, iz, pole_r, n);
  // The region as one bit word per axis: most earlier candidates lie outside
  // it, and a lane tells so from the candidate's cell code alone, four
  // groups of 64 loaded at once. Those inside queue up in the wave's LDS, in
  // batch order, and are tested 64 at a time, so the pair tests run on full
  // waves.
  __shared__ uint32_t queue_s[BP_WAVES][128];
  uint32_t* queue = queue_s[threadIdx.x >> 6];
  const uint32_t bx = R.axis_bits(0), by = R.axis_bits(1), bz = R.axis_bits(2);
  constexpr uint32_t U = 4;
  uint32_t found = 0;  // wave-uniform
  uint32_t queued = 0;  // wave-uniform
  const uint64_t below = (1ull << lane) - 1ull;
  // the pair tests of queue[0 .. take), in queue order
  auto test_queued = [&](uint32_t take) {
    const bool act = lane < take;
    PairEffect e{};
    bool any = false;
    if (act) {
      const uint32_t a = queue[lane];
      const uint32_t c = cells[a];
      const float ra = rad[a];
      const uint32_t fa = flat[a];
      const uint32_t jx = fa / nn, jy = (fa / n) % n, jz = fa % n;
      const float ax = (float)jx * P.l, ay = (float)jy * P.l, az = (float)jz * P.l;
      const uint32_t wa = c >> 10, wb = (c >> 5) & 31u, wc = c & 31u;
      e.a = a;
      if (need_state)
        e.state = (uint8_t)pair_state(q, wa, wb, wc, ax, ay, az, ra);
      if (need_state_r)
        e.state_r = (uint8_t)pair_state(qr, wa, wb, wc, ax, ay, az, ra);
      PaintSphere S;
      paint_sphere(S, ax, ay, az, P.scale * ra, P.l, n);
      e.node = want_node && paint_hits(S, ix, iy, iz);
      e.pole = want_pole && poles.painted(S, ix, iy, iz);
      e.pole_r = want_pole_r && poles_r.painted(S, ix, iy, iz);
      any = e.node | e.pole | e.pole_r | e.state | e.state_r;
    }
    const uint64_t m = __ballot(any);
    if constexpr (WRITE) {
      const uint32_t at = found + (uint32_t)__popcll(m & below);
      if (any)
        eff[at0 + at] = e;
    }
    found += (uint32_t)__popcll(m);
  };
  for (uint32_t a0 = 0; a0 < b && found < BATCH_EFF_FULL && !R.empty(); a0 += 64 * U) {
    uint32_t code[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t a = a0 + u * 64 + lane;
      code[u] = a < b ? cells[a] : BATCH_NO_CELL;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint32_t c = code[u];
      const bool near = c != BATCH_NO_CELL &&
                        ((bx >> (c >> 10)) & (by >> ((c >> 5) & 31u)) & (bz >> (c & 31u)) & 1u);
      const uint64_t m = __ballot(near);
      if (m == 0ull)
        continue;
      if (near)
        queue[queued + (uint32_t)__popcll(m & below)] = a0 + u * 64 + lane;
      queued += (uint32_t)__popcll(m);
      if (queued >= 64) {
        pairs_wave_sync();
        test_queued(64);
        // the entries past the first 64 move to the front
        const uint32_t rest = queued - 64;
        const uint32_t moved = lane < rest ? queue[64 + lane] : 0u;
        pairs_wave_sync();
        if (lane < rest)
          queue[lane] = moved;
        queued = rest;
      }
    }
  }
  pairs_wave_sync();
  if (queued > 0 && found < BATCH_EFF_FULL)
    test_queued(queued);
  if constexpr (!WRITE) {
    if (lane == 0) {
      const uint32_t c = found < BATCH_EFF_FULL ? found : BATCH_EFF_FULL;
      cnt32[b] = c;
      cnt8[b] = (uint8_t)c;
    }
  }
}

}  // namespace sif_amd
