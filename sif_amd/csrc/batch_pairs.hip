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
#include "batch_pairs_device.h"
#include "primitives.h"

namespace sif_amd {

// The largest grown radius of the batch (the set's max_r of its region
// scans; -1 when none grew), and each candidate's cell code.
__global__ void __launch_bounds__(1024) batch_prep_kernel(const uint32_t* __restrict__ flat,
  const float* __restrict__ rad, uint32_t count, BatchPairParams P, float* __restrict__ out,
  uint32_t* __restrict__ cells) {
  __shared__ float wmax[16];
  float m = -1.0f;
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) {
    const float r = rad[i];
    m = fmaxf(m, r);
    cells[i] = batch_cell_code(flat[i], r, P);
  }
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

int batch_pairs(const uint32_t* flat, const float* rad, const uint32_t* grown_state,
  uint32_t count, const BatchPairParams& P, BatchPairScratch& W) {
  if (count == 0)
    return SIF_OK;
  batch_prep_kernel<<<1, 1024, 0, stream()>>>(flat, rad, count, P, W.set_max_r, W.cells);
  SIF_HIP(hipGetLastError());
  const unsigned blocks = (count + BP_WAVES - 1) / BP_WAVES;
  batch_pairs_kernel<false><<<blocks, 64 * BP_WAVES, 0, stream()>>>(flat, rad, grown_state,
    count, P, W.set_max_r, W.cells, W.cnt32, W.cnt8, nullptr, nullptr, 0);
  SIF_HIP(hipGetLastError());
  int st = exclusive_scan_u32(W.cnt32, count, W.offset, nullptr);
  if (st != SIF_OK)
    return st;
  batch_pairs_kernel<true><<<blocks, 64 * BP_WAVES, 0, stream()>>>(flat, rad, grown_state,
    count, P, W.set_max_r, W.cells, nullptr, W.cnt8, W.offset, W.eff, W.capacity);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

}  // namespace sif_amd
