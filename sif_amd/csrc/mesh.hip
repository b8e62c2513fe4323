// Chain mesh: a counting sort of the particles by mesh cell.
//
// Replaces chain_mesh_create (src/structures/chain_mesh.c:680-832), its
// canonical per-cell order (chain_mesh.c:364-542) and the weight sums
// (chain_mesh.c:624-672) of sif v0.5.0.
//
// The reference claims slots with atomics and then sorts each cell by source
// index. Here the ordering is made deterministic from the start: a particle's
// slot is offsets[cell] + (the number of particles with a smaller index in the
// same cell), obtained with a stable LSD radix sort on the cell key, whose
// input is in source order. That is the reference's canonical order exactly.
// Cells per side are at most 1024 (SIF_FINDER_MESH_MAX_CELLS), so keys fit in
// 30 bits; the sort runs only over the bits the mesh needs.
#include "common.h"
#include "internal.h"
#include "primitives.h"
#include "sif_amd.h"

namespace sif_amd {

__device__ __forceinline__ float box_fold_m(float v, float box) {
  if (v >= box)
    v -= box;
  else if (v < 0.0f)
    v += box;
  return v >= box ? 0.0f : v;
}

// chain_mesh.c:316-331: ix = (int32)(fold(x) * (1 / l)), clamped.
__global__ void __launch_bounds__(256) mesh_key_kernel(const float* __restrict__ xs,
  const float* __restrict__ ys, const float* __restrict__ zs, uint32_t n,
  uint32_t nc, float inv_l, float box, uint32_t* __restrict__ keys,
  uint32_t* __restrict__ vals, uint32_t* __restrict__ counts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint32_t key = 0xffffffffu;
  if (i < n) {
    int32_t ix = (int32_t)(box_fold_m(xs[i], box) * inv_l);
    int32_t iy = (int32_t)(box_fold_m(ys[i], box) * inv_l);
    int32_t iz = (int32_t)(box_fold_m(zs[i], box) * inv_l);
    const int32_t m = (int32_t)nc;
    ix = ix < 0 ? 0 : (ix >= m ? m - 1 : ix);
    iy = iy < 0 ? 0 : (iy >= m ? m - 1 : iy);
    iz = iz < 0 ? 0 : (iz >= m ? m - 1 : iz);
    key = ((uint32_t)ix * nc + (uint32_t)iy) * nc + (uint32_t)iz;
    keys[i] = key;
    vals[i] = i;
  }
  // Runs of equal keys in neighbouring lanes add their count once.
  const uint32_t prev = __shfl_up(key, 1, 64);
  const bool head = (i < n) && (lane == 0 || prev != key);
  const uint64_t heads = __ballot(head);
  const uint64_t valid = __ballot(i < n);
  if (head) {
    const uint64_t later = lane == 63 ? 0ull : (heads >> (lane + 1)) << (lane + 1);
    const int end = later ? __ffsll((long long)later) - 1 : 64 - __clzll((long long)valid);
    atomicAdd(&counts[key], (uint32_t)(end - lane));
  }
}

__global__ void __launch_bounds__(256) mesh_gather_kernel(const uint32_t* __restrict__ perm,
  uint32_t n, const float* __restrict__ xs, const float* __restrict__ ys,
  const float* __restrict__ zs, const float* __restrict__ ws, float box,
  float* __restrict__ mx, float* __restrict__ my, float* __restrict__ mz,
  float* __restrict__ mw, uint64_t* __restrict__ orig) {
  const uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= n)
    return;
  const uint32_t i = perm[pos];
  mx[pos] = box_fold_m(xs[i], box);
  my[pos] = box_fold_m(ys[i], box);
  mz[pos] = box_fold_m(zs[i], box);
  if (mw)
    mw[pos] = ws[i];
  if (orig)
    orig[pos] = i;
}

// chain_mesh.c:646-664: each cell summed in double in mesh order.
__global__ void cell_weight_kernel(const uint64_t* __restrict__ offsets,
  uint64_t cells, const float* __restrict__ mw, float* __restrict__ cw,
  double* __restrict__ cd) {
  const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cells)
    return;
  double s = 0.0;
  for (uint64_t p = offsets[c]; p < offsets[c + 1]; p++)
    s += (double)mw[p];
  cw[c] = (float)s;
  cd[c] = s;
}

// chain_mesh.c:639-669: 64 fixed blocks of cells, each summed serially.
__global__ void weight_blocks_kernel(const double* __restrict__ cd, uint64_t cells,
  double* __restrict__ out) {
  const int b = threadIdx.x;
  const uint64_t block = (cells + 63) / 64;
  const uint64_t s = (uint64_t)b * block;
  const uint64_t e = s + block < cells ? s + block : cells;
  double sum = 0.0;
  for (uint64_t c = s; c < e; c++)
    sum += cd[c];
  out[b] = sum;
}

namespace {
DeviceBuffer g_keys, g_vals, g_counts, g_cd, g_blocks;
}

int mesh_build(const float* x, const float* y, const float* z, const float* w,
  uint64_t n64, uint32_t nc, float box, uint64_t* offsets, float* mx, float* my,
  float* mz, float* mw, uint64_t* orig, float* cw, double* total_weight) {
  if (n64 >= 0xffffffffull) {
    error_set(SIF_ERR_RANGE, "chain mesh: %llu particles (at most 2^32 - 2)",
      (unsigned long long)n64);
    return SIF_ERR_RANGE;
  }
  const uint32_t n = (uint32_t)n64;
  const uint64_t cells = (uint64_t)nc * nc * nc;
  const float cell = box / (float)nc;
  const float inv_l = 1.0f / cell;
  int st = g_keys.reserve((uint64_t)(n ? n : 1) * 4);
  if (st == SIF_OK)
    st = g_vals.reserve((uint64_t)(n ? n : 1) * 4);
  if (st == SIF_OK)
    st = g_counts.reserve(cells * 4);
  if (st != SIF_OK)
    return st;
  timing_begin(SIF_AMD_STAGE_MESH);
  SIF_HIP(hipMemsetAsync(g_counts.ptr, 0, cells * 4, stream()));
  if (n) {
    mesh_key_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(x, y, z, n, nc,
      inv_l, box, g_keys.as<uint32_t>(), g_vals.as<uint32_t>(),
      g_counts.as<uint32_t>());
    SIF_HIP(hipGetLastError());
  }
  st = exclusive_scan_u32(g_counts.as<uint32_t>(), cells, offsets, offsets + cells);
  if (st != SIF_OK)
    return st;
  int bits = 1;
  while (bits < 32 && (1ull << bits) < cells)
    bits++;
  st = radix_sort_u32(g_keys.as<uint32_t>(), g_vals.as<uint32_t>(), n, bits);
  if (st != SIF_OK)
    return st;
  if (n) {
    mesh_gather_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(
      g_vals.as<uint32_t>(), n, x, y, z, w, box, mx, my, mz, w ? mw : nullptr, orig);
    SIF_HIP(hipGetLastError());
  }
  if (w && cw) {
    st = g_cd.reserve(cells * sizeof(double));
    if (st == SIF_OK)
      st = g_blocks.reserve(64 * sizeof(double));
    if (st != SIF_OK)
      return st;
    cell_weight_kernel<<<grid_for(cells, 256), 256, 0, stream()>>>(
      offsets, cells, mw, cw, g_cd.as<double>());
    weight_blocks_kernel<<<1, 64, 0, stream()>>>(g_cd.as<double>(), cells,
      g_blocks.as<double>());
    SIF_HIP(hipGetLastError());
    double blocks[64];
    SIF_HIP(hipMemcpyAsync(blocks, g_blocks.ptr, sizeof(blocks),
      hipMemcpyDeviceToHost, stream()));
    SIF_HIP(hipStreamSynchronize(stream()));
    double total = 0.0;
    for (int b = 0; b < 64; b++)
      total += blocks[b];
    if (total_weight)
      *total_weight = total;
  } else if (total_weight) {
    *total_weight = (double)n;
  }
  timing_end(SIF_AMD_STAGE_MESH);
  if (timing_active()) {  // positions in and out once, offsets, indices
    const double per = w ? 16.0 : 12.0;
    timing_bytes(SIF_AMD_STAGE_MESH, 2.0 * per * (double)n + 8.0 * (double)(cells + 1) +
      (orig ? 8.0 * (double)n : 0.0) + (cw ? 4.0 * (double)cells : 0.0));
  }
  return SIF_OK;
}

}  // namespace sif_amd

using namespace sif_amd;

extern "C" {

int sif_amd_chain_mesh_build(const float* x, const float* y, const float* z,
  const float* w, uint64_t n, uint32_t n_cells, float box, uint64_t* offsets,
  float* mx, float* my, float* mz, float* mw, uint64_t* original_indices,
  float* cell_weights, double* total_weight) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!offsets || !mx || !my || !mz || n_cells == 0 || n_cells > 1024 ||
      !(box > 0.0f) || (n && (!x || !y || !z)) || (w && !mw)) {
    error_set(SIF_ERR_INVALID, "sif_amd_chain_mesh_build: invalid arguments");
    return SIF_ERR_INVALID;
  }
  st = mesh_build(x, y, z, w, n, n_cells, box, offsets, mx, my, mz, mw,
    original_indices, cell_weights, total_weight);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

}  // extern "C"
