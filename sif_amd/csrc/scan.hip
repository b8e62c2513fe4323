// Candidate scan: every cell with delta <= threshold whose mask bit is clear,
// compacted in cell order and sorted deepest first.
//
// Replaces sif__finder_scan_candidates and sort_candidates
// (src/finder/utils.c:123-213, include/sif/utils/sort.h:56-145) of
// sif v0.5.0. The reference orders equal depths by its quicksort's
// partitioning; here the order is (delta ascending, flat index ascending) --
// a stable radix sort of a list already in flat order. Where no two depths
// are equal, the two orders are the same.
#include "common.h"
#include "internal.h"
#include "primitives.h"
#include "sif_amd.h"

namespace sif_amd {

__global__ void __launch_bounds__(256) cand_flag_kernel(const float* __restrict__ delta,
  const uint64_t* __restrict__ mask, uint64_t total, float t,
  uint32_t* __restrict__ flags) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += stride) {
    const bool masked = mask && ((mask[i >> 6] >> (i & 63)) & 1ull);
    flags[i] = (delta[i] <= t && !masked) ? 1u : 0u;
  }
}

__device__ __forceinline__ uint32_t float_key(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float key_float(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void cand_scatter_kernel(const float* __restrict__ delta,
  const uint32_t* __restrict__ flags, const uint64_t* __restrict__ pos,
  uint64_t total, uint32_t* __restrict__ keys, uint32_t* __restrict__ flat) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += stride) {
    if (flags[i]) {
      const uint64_t p = pos[i];
      keys[p] = float_key(delta[i]);
      flat[p] = (uint32_t)i;
    }
  }
}

__global__ void cand_unkey_kernel(const uint32_t* __restrict__ keys, uint64_t n,
  float* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    out[i] = key_float(keys[i]);
}

namespace {
DeviceBuffer g_flags, g_pos, g_count;
}

// Writes the candidates, in flat order, to keys/flat (device, capacity `cap`)
// and their number to *count; keys are the order-preserving integer images of
// the depths. When the count exceeds `cap` nothing is written and the caller
// grows the buffers and calls again.
int scan_candidates_flat(const float* delta, const uint64_t* mask, uint64_t total,
  float t, uint32_t* keys, uint32_t* flat, uint64_t cap, uint64_t* count) {
  if (total >= 0x80000000ull) {
    error_set(SIF_ERR_RANGE, "scan: %llu cells (at most 2^31 - 1)",
      (unsigned long long)total);
    return SIF_ERR_RANGE;
  }
  int st = g_flags.reserve(total * 4);
  if (st == SIF_OK)
    st = g_pos.reserve(total * 8);
  if (st == SIF_OK)
    st = g_count.reserve(8);
  if (st != SIF_OK)
    return st;
  timing_begin(SIF_AMD_STAGE_SCAN);
  unsigned g = grid_for(total, 256);
  if (g > 8192)
    g = 8192;
  cand_flag_kernel<<<g, 256, 0, stream()>>>(delta, mask, total, t, g_flags.as<uint32_t>());
  SIF_HIP(hipGetLastError());
  st = exclusive_scan_u32(g_flags.as<uint32_t>(), total, g_pos.as<uint64_t>(),
    g_count.as<uint64_t>());
  if (st != SIF_OK)
    return st;
  uint64_t n = 0;
  SIF_HIP(hipMemcpyAsync(&n, g_count.ptr, 8, hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  *count = n;
  if (n <= cap) {
    cand_scatter_kernel<<<g, 256, 0, stream()>>>(delta, g_flags.as<uint32_t>(),
      g_pos.as<uint64_t>(), total, keys, flat);
    SIF_HIP(hipGetLastError());
  }
  timing_end(SIF_AMD_STAGE_SCAN);
  return SIF_OK;
}

// Deepest first; equal depths keep their order (a stable radix sort), so a
// list in flat order comes out ordered by (depth, flat index). Bit 31 of a
// flat value rides along untouched.
int sort_candidates(uint32_t* keys, uint32_t* flat, uint64_t n) {
  if (n == 0)
    return SIF_OK;
  timing_begin(SIF_AMD_STAGE_SORT);
  const int st = radix_sort_u32(keys, flat, (uint32_t)n, 32);
  timing_end(SIF_AMD_STAGE_SORT);
  return st;
}

// The sorted candidates (sif__finder_scan_candidates + sort_candidates).
int scan_candidates(const float* delta, const uint64_t* mask, uint64_t total,
  float t, uint32_t* keys, uint32_t* flat, uint64_t cap, uint64_t* count) {
  int st = scan_candidates_flat(delta, mask, total, t, keys, flat, cap, count);
  if (st != SIF_OK || *count > cap)
    return st;
  return sort_candidates(keys, flat, *count);
}

int keys_to_delta(const uint32_t* keys, uint64_t n, float* out) {
  if (n == 0)
    return SIF_OK;
  cand_unkey_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(keys, n, out);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

}  // namespace sif_amd

using namespace sif_amd;

extern "C" {

int sif_amd_scan_candidates(const float* delta, const uint64_t* mask_words,
  uint64_t total_cells, float threshold, float* out_delta, uint32_t* out_flat,
  uint64_t capacity, uint64_t* count) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!delta || !out_delta || !out_flat || !count) {
    error_set(SIF_ERR_INVALID, "sif_amd_scan_candidates: invalid arguments");
    return SIF_ERR_INVALID;
  }
  // out_delta doubles as the key buffer (same size); keys become depths at
  // the end.
  uint32_t* keys = reinterpret_cast<uint32_t*>(out_delta);
  st = scan_candidates(delta, mask_words, total_cells, threshold, keys, out_flat,
    capacity, count);
  if (st != SIF_OK)
    return st;
  if (*count <= capacity)
    st = keys_to_delta(keys, *count, out_delta);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

}  // extern "C"
