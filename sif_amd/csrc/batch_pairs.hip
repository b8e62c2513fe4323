// The effects of a growth batch's earlier candidates on each later one
// (batch_pairs.h), found on the device right after the batch is grown.
//
// The commit of exodus_finder.c:2174-2329 takes a batch's candidates in order,
// and one accepted becomes a void of the radius it was grown to. What an
// earlier candidate a would do to a later candidate b, were a accepted --
// paint b's node or a pole node, or reach it in the overlap test at the rung
// radius or at b's grown radius -- depends on the two candidates alone, so
// every pair is settled here at once, one wavefront per candidate b with its
// lanes over the earlier candidates, and the commit only ORs the effects of
// those it accepts.
//
// Each effect is the host's pair_effects() answer for the pair, by the same
// functions (overlap.h, paint.h, rung_voids.h). The device does not know what
// the rung's newer voids already decide for b (the commit's probes of them),
// so it lists the effects of every test b could still need; an effect of a
// test already decided changes nothing in an OR. The overlap states are
// taken with the catalogue's largest radius as the batch was launched, no
// larger than at its commit: a state 1 stays 1 as that radius grows, a 2 sends
// the candidate to the reference's own scan, and a 0 holds for any radius.
#include "common.h"
#include "internal.h"
#include "batch_pairs.h"
#include "primitives.h"

namespace sif_amd {

constexpr int BP_WAVES = 4;

// The largest grown radius of the batch (the set's max_r of its region
// scans); -1 when none grew.
__global__ void __launch_bounds__(1024) batch_max_radius_kernel(const float* __restrict__ rad,
  uint32_t count, float* __restrict__ out) {
  __shared__ float wmax[16];
  float m = -1.0f;
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x)
    m = fmaxf(m, rad[i]);
  for (int off = 32; off > 0; off >>= 1)
    m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0)
    wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < blockDim.x / 64; w++)
      m = fmaxf(m, wmax[w]);
    *out = m;
  }
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
  const float* __restrict__ set_max_r, uint32_t* __restrict__ cnt32,
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
  uint32_t found = 0;  // wave-uniform
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t a0 = 0; a0 < b && found < BATCH_EFF_FULL && !R.empty(); a0 += 64) {
    const uint32_t a = a0 + lane;
    PairEffect e{};
    bool any = false;
    if (a < b) {
      const float ra = rad[a];
      if (ra >= 0.0f) {
        const uint32_t fa = flat[a];
        const uint32_t jx = fa / nn, jy = (fa / n) % n, jz = fa % n;
        const float ax = (float)jx * P.l, ay = (float)jy * P.l, az = (float)jz * P.l;
        const uint32_t wa = void_axis_cell(ax, P.inv_l), wb = void_axis_cell(ay, P.inv_l),
                       wc = void_axis_cell(az, P.inv_l);
        if (R.holds(wa, wb, wc)) {
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
      }
    }
    const uint64_t m = __ballot(any);
    if constexpr (WRITE) {
      const uint32_t at = found + (uint32_t)__popcll(m & below);
      if (any)
        eff[at0 + at] = e;
    }
    found += (uint32_t)__popcll(m);
  }
  if constexpr (!WRITE) {
    if (lane == 0) {
      const uint32_t c = found < BATCH_EFF_FULL ? found : BATCH_EFF_FULL;
      cnt32[b] = c;
      cnt8[b] = (uint8_t)c;
    }
  }
}

int batch_pairs(const uint32_t* flat, const float* rad, const uint32_t* grown_state,
  uint32_t count, const BatchPairParams& P, BatchPairScratch& W) {
  if (count == 0)
    return SIF_OK;
  batch_max_radius_kernel<<<1, 1024, 0, stream()>>>(rad, count, W.set_max_r);
  SIF_HIP(hipGetLastError());
  const unsigned blocks = (count + BP_WAVES - 1) / BP_WAVES;
  batch_pairs_kernel<false><<<blocks, 64 * BP_WAVES, 0, stream()>>>(flat, rad, grown_state,
    count, P, W.set_max_r, W.cnt32, W.cnt8, nullptr, nullptr, 0);
  SIF_HIP(hipGetLastError());
  int st = exclusive_scan_u32(W.cnt32, count, W.offset, nullptr);
  if (st != SIF_OK)
    return st;
  batch_pairs_kernel<true><<<blocks, 64 * BP_WAVES, 0, stream()>>>(flat, rad, grown_state,
    count, P, W.set_max_r, nullptr, W.cnt8, W.offset, W.eff, W.capacity);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

}  // namespace sif_amd
