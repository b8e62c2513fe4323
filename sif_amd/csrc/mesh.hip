// Chain mesh: a counting sort of the particles by mesh cell.
//
// Replaces chain_mesh_create (src/structures/chain_mesh.c:680-832), its
// canonical per-cell order (chain_mesh.c:364-542) and the weight sums
// (chain_mesh.c:624-672) of sif v0.5.0.
//
// The order is the reference's canonical one: a particle's slot is
// offsets[cell] + (the number of particles with a smaller source index in the
// same cell). It is reached in three passes over the particles:
//
//   count    each particle's cell, the counts per cell (a wavefront's run of
//            equal cells adds its length once);
//   scatter  the folded positions (and weights, and the source index) to a
//            slot of their cell claimed with an atomic cursor, in whatever
//            order the wavefronts arrive;
//   order    a workgroup per cell puts its slots in source order: each
//            element's rank among the cell's source indices (sorted in LDS,
//            then searched) is its place, so the result does not depend on
//            the arrival order. One wavefront takes a cell of up to
//            MESH_ORDER_CAP particles; a larger one gets MESH_BIG_THREADS
//            threads and up to MESH_BIG_BYTES of indices (the count pass
//            lists those cells).
//
// A cell past what a workgroup's LDS holds sends the whole mesh through the
// stable LSD radix sort of (cell, source index) instead. Source indices are
// 32-bit below 2^32 particles and 64-bit above.
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

// --- count, scatter, order ---

constexpr uint32_t MESH_ORDER_CAP = 1024;  // particles a cell, for the ordering pass
// The large cells' pass: a bigger workgroup a cell, its source indices in LDS.
constexpr uint32_t MESH_BIG_THREADS = 1024;
constexpr uint32_t MESH_BIG_BYTES = 64 * 1024;
template <class IdxT>
constexpr uint32_t mesh_big_cap() {
  return MESH_BIG_BYTES / (uint32_t)sizeof(IdxT);
}

__device__ __forceinline__ uint32_t mesh_cell_of(float x, float y, float z, uint32_t nc,
  float inv_l, float box) {
  int32_t ix = (int32_t)(box_fold_m(x, box) * inv_l);
  int32_t iy = (int32_t)(box_fold_m(y, box) * inv_l);
  int32_t iz = (int32_t)(box_fold_m(z, box) * inv_l);
  const int32_t m = (int32_t)nc;
  ix = ix < 0 ? 0 : (ix >= m ? m - 1 : ix);
  iy = iy < 0 ? 0 : (iy >= m ? m - 1 : iy);
  iz = iz < 0 ? 0 : (iz >= m ? m - 1 : iz);
  return ((uint32_t)ix * nc + (uint32_t)iy) * nc + (uint32_t)iz;
}

// A wavefront's lanes holding the same key in a row form a run: its first
// lane and (on that lane) its length.
__device__ __forceinline__ int mesh_run(uint32_t key, bool valid, int lane, uint32_t* len) {
  const uint32_t prev = __shfl_up(key, 1, 64);
  const bool head = valid && (lane == 0 || prev != key);
  const uint64_t heads = __ballot(head);
  const uint64_t live = __ballot(valid);
  const uint64_t upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
  const int first = 63 - __clzll((long long)(heads & upto));
  if (head) {
    const uint64_t later = lane == 63 ? 0ull : (heads >> (lane + 1)) << (lane + 1);
    const int end = later ? __ffsll((long long)later) - 1 : 64 - __clzll((long long)live);
    *len = (uint32_t)(end - lane);
  }
  return first;
}

__global__ void __launch_bounds__(256) mesh_count_kernel(const float* __restrict__ xs,
  const float* __restrict__ ys, const float* __restrict__ zs, uint64_t n, uint32_t nc,
  float inv_l, float box, uint32_t* __restrict__ counts) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = i < n;
  const uint32_t key = valid ? mesh_cell_of(xs[i], ys[i], zs[i], nc, inv_l, box) : 0xffffffffu;
  uint32_t len = 0;
  const int first = mesh_run(key, valid, lane, &len);
  if (valid && first == lane)
    atomicAdd(&counts[key], len);
}

template <class IdxT>
__global__ void __launch_bounds__(256) mesh_scatter_kernel(const float* __restrict__ xs,
  const float* __restrict__ ys, const float* __restrict__ zs, const float* __restrict__ ws,
  uint64_t n, uint32_t nc, float inv_l, float box, const uint64_t* __restrict__ offsets,
  uint32_t* __restrict__ cursor, float* __restrict__ mx, float* __restrict__ my,
  float* __restrict__ mz, float* __restrict__ mw, IdxT* __restrict__ src) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = i < n;
  float x = 0.0f, y = 0.0f, z = 0.0f;
  uint32_t key = 0xffffffffu;
  if (valid) {
    x = xs[i];
    y = ys[i];
    z = zs[i];
    key = mesh_cell_of(x, y, z, nc, inv_l, box);
  }
  uint32_t len = 0;
  const int first = mesh_run(key, valid, lane, &len);
  uint32_t base = 0;
  if (valid && first == lane)
    base = atomicAdd(&cursor[key], len);
  base = __shfl(base, first, 64);
  if (valid) {
    const uint64_t slot = offsets[key] + base + (uint32_t)(lane - first);
    mx[slot] = box_fold_m(x, box);
    my[slot] = box_fold_m(y, box);
    mz[slot] = box_fold_m(z, box);
    if (mw)
      mw[slot] = ws[i];
    src[slot] = (IdxT)i;
  }
}

// The ordering pass: a workgroup of THREADS threads a cell of at most CAP
// particles -- cells q = blockIdx.x, blockIdx.x + gridDim.x, ... below
// `count`: cell q, or big[q] when a list is given.
// Every element's place is its rank among the cell's source indices, which
// are distinct: the indices are sorted in LDS (bitonic, padded to a power of
// two with the largest value) and each element finds its own there by binary
// search, so the result does not depend on the order the scatter left. Every
// thread holds its own elements (thread t the elements t, t + THREADS, ...)
// in registers before any is written back (the cell's slots are rewritten in
// place). A cell of more than CAP particles is left to the other launch.
template <class IdxT, uint32_t THREADS, uint32_t CAP>
__global__ void __launch_bounds__(THREADS) mesh_order_kernel(
  const uint64_t* __restrict__ offsets, const uint64_t* __restrict__ big, uint64_t count,
  const IdxT* __restrict__ src, float* mx, float* my, float* mz, float* mw,
  uint64_t* __restrict__ orig) {
  constexpr int PER = (int)(CAP / THREADS);
  __shared__ IdxT ids[CAP];
  for (uint64_t q = blockIdx.x; q < count; q += gridDim.x) {
    const uint64_t c = big ? big[q] : q;
    const uint64_t b = offsets[c];
    uint32_t cnt = (uint32_t)(offsets[c + 1] - b);
    if (cnt > CAP)
      cnt = 0;
    uint32_t p2 = 1;
    while (p2 < cnt)
      p2 <<= 1;
    const uint32_t tid = threadIdx.x;
    IdxT own[PER];
    float px[PER], py[PER], pz[PER], pw[PER];
  #pragma unroll
    for (int t = 0; t < PER; t++) {
      const uint32_t j = tid + THREADS * (uint32_t)t;
      own[t] = 0;
      px[t] = py[t] = pz[t] = pw[t] = 0.0f;
      if (j < cnt) {
        own[t] = src[b + j];
        px[t] = mx[b + j];
        py[t] = my[b + j];
        pz[t] = mz[b + j];
        pw[t] = mw ? mw[b + j] : 0.0f;
      }
      if (j < p2)
        ids[j] = j < cnt ? own[t] : (IdxT)~(IdxT)0;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= p2; k <<= 1) {
      for (uint32_t h = k >> 1; h > 0; h >>= 1) {
        for (uint32_t i = tid; i < p2; i += THREADS) {
          const uint32_t l = i ^ h;
          if (l > i) {
            const IdxT u = ids[i], v = ids[l];
            const bool up = (i & k) == 0;
            if (up ? (u > v) : (u < v)) {
              ids[i] = v;
              ids[l] = u;
            }
          }
        }
        __syncthreads();
      }
    }
  #pragma unroll
    for (int t = 0; t < PER; t++) {
      const uint32_t j = tid + THREADS * (uint32_t)t;
      if (j < cnt) {
        uint32_t lo = 0, hi = cnt;  // the first place holding own[t] or more
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (ids[mid] < own[t])
            lo = mid + 1;
          else
            hi = mid;
        }
        const uint64_t at = b + lo;
        mx[at] = px[t];
        my[at] = py[t];
        mz[at] = pz[t];
        if (mw)
          mw[at] = pw[t];
        if (orig)
          orig[at] = (uint64_t)own[t];
      }
    }
    __syncthreads();  // the next cell rewrites ids
  }
}

// The largest cell count, and the cells of more than MESH_ORDER_CAP particles
// (out[1] of them; the first `big_cap` are listed in `big`, in no order).
__global__ void count_max_kernel(const uint32_t* __restrict__ counts, uint64_t cells,
  uint32_t* __restrict__ out, uint64_t* __restrict__ big, uint32_t big_cap) {
  uint32_t m = 0;
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < cells;
       c += (uint64_t)gridDim.x * blockDim.x) {
    m = counts[c] > m ? counts[c] : m;
    if (counts[c] > MESH_ORDER_CAP) {
      const uint32_t k = atomicAdd(&out[1], 1u);
      if (k < big_cap)
        big[k] = c;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m)
    atomicMax(out, m);
}

namespace {
thread_local DeviceBuffer g_keys, g_vals, g_counts, g_cd, g_blocks, g_cursor, g_src, g_max, g_big;
}

// The stable radix path (a cell past what a workgroup's LDS holds):
// (cell, source index) pairs sorted by cell, then the positions gathered.
static int mesh_by_radix(const float* x, const float* y, const float* z, const float* w,
  uint32_t n, uint32_t nc, float inv_l, float box, uint64_t cells, float* mx, float* my,
  float* mz, float* mw, uint64_t* orig) {
  int st = g_keys.reserve((uint64_t)(n ? n : 1) * 4);
  if (st == SIF_OK)
    st = g_vals.reserve((uint64_t)(n ? n : 1) * 4);
  if (st != SIF_OK)
    return st;
  // the counts are known already: this pass only writes the keys
  SIF_HIP(hipMemsetAsync(g_cursor.ptr, 0, cells * 4, stream()));
  mesh_key_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(x, y, z, n, nc, inv_l, box,
    g_keys.as<uint32_t>(), g_vals.as<uint32_t>(), g_cursor.as<uint32_t>());
  SIF_HIP(hipGetLastError());
  int bits = 1;
  while (bits < 32 && (1ull << bits) < cells)
    bits++;
  st = radix_sort_u32(g_keys.as<uint32_t>(), g_vals.as<uint32_t>(), n, bits);
  if (st != SIF_OK)
    return st;
  mesh_gather_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(
    g_vals.as<uint32_t>(), n, x, y, z, w, box, mx, my, mz, w ? mw : nullptr, orig);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

template <class IdxT>
static int mesh_scatter_order(const float* x, const float* y, const float* z, const float* w,
  uint64_t n, uint32_t nc, float inv_l, float box, uint64_t cells, const uint64_t* offsets,
  float* mx, float* my, float* mz, float* mw, uint64_t* orig, uint32_t n_big) {
  int st = g_src.reserve((n ? n : 1) * sizeof(IdxT));
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemsetAsync(g_cursor.ptr, 0, cells * 4, stream()));
  mesh_scatter_kernel<IdxT><<<grid_for(n, 256), 256, 0, stream()>>>(x, y, z, w, n, nc, inv_l,
    box, offsets, g_cursor.as<uint32_t>(), mx, my, mz, w ? mw : nullptr, g_src.as<IdxT>());
  SIF_HIP(hipGetLastError());
  const uint64_t blocks = cells < (1ull << 22) ? cells : (1ull << 22);
  mesh_order_kernel<IdxT, 64, MESH_ORDER_CAP><<<(unsigned)blocks, 64, 0, stream()>>>(offsets,
    nullptr, cells, g_src.as<IdxT>(), mx, my, mz, w ? mw : nullptr, orig);
  SIF_HIP(hipGetLastError());
  if (n_big) {
    mesh_order_kernel<IdxT, MESH_BIG_THREADS, mesh_big_cap<IdxT>()>
      <<<n_big, MESH_BIG_THREADS, 0, stream()>>>(offsets, g_big.as<uint64_t>(), n_big,
        g_src.as<IdxT>(), mx, my, mz, w ? mw : nullptr, orig);
    SIF_HIP(hipGetLastError());
  }
  return SIF_OK;
}

int mesh_build(const float* x, const float* y, const float* z, const float* w,
  uint64_t n, uint32_t nc, float box, uint64_t* offsets, float* mx, float* my,
  float* mz, float* mw, uint64_t* orig, float* cw, double* total_weight) {
  const uint64_t cells = (uint64_t)nc * nc * nc;
  const float cell = box / (float)nc;
  const float inv_l = 1.0f / cell;
  int st = g_counts.reserve(cells * 4);
  if (st == SIF_OK)
    st = g_cursor.reserve(cells * 4);
  if (st == SIF_OK)
    st = g_max.reserve(8);
  // cells of more than MESH_ORDER_CAP particles: at most n / (cap + 1)
  const uint64_t big_room = n / (MESH_ORDER_CAP + 1) + 1;
  const uint32_t big_cap = (uint32_t)(big_room < cells ? big_room : cells);
  if (st == SIF_OK)
    st = g_big.reserve((uint64_t)big_cap * 8);
  if (st != SIF_OK)
    return st;
  timing_begin(SIF_AMD_STAGE_MESH);
  SIF_HIP(hipMemsetAsync(g_counts.ptr, 0, cells * 4, stream()));
  SIF_HIP(hipMemsetAsync(g_max.ptr, 0, 8, stream()));
  if (n) {
    mesh_count_kernel<<<grid_for(n, 256), 256, 0, stream()>>>(x, y, z, n, nc, inv_l, box,
      g_counts.as<uint32_t>());
    SIF_HIP(hipGetLastError());
  }
  st = exclusive_scan_u32(g_counts.as<uint32_t>(), cells, offsets, offsets + cells);
  if (st != SIF_OK)
    return st;
  if (n) {
    const unsigned mb = grid_for(cells, 256) > 1024 ? 1024 : grid_for(cells, 256);
    count_max_kernel<<<mb, 256, 0, stream()>>>(g_counts.as<uint32_t>(), cells,
      g_max.as<uint32_t>(), g_big.as<uint64_t>(), big_cap);
    SIF_HIP(hipGetLastError());
    uint32_t got[2] = {0, 0};  // the largest cell, the cells past MESH_ORDER_CAP
    SIF_HIP(hipMemcpyAsync(got, g_max.ptr, 8, hipMemcpyDeviceToHost, stream()));
    SIF_HIP(hipStreamSynchronize(stream()));
    const uint32_t largest = got[0], n_big = got[1];
    const bool narrow = n < 0xffffffffull;
    const uint32_t cap = narrow ? mesh_big_cap<uint32_t>() : mesh_big_cap<uint64_t>();
    if (largest <= cap && n_big <= big_cap) {
      st = narrow ? mesh_scatter_order<uint32_t>(x, y, z, w, n, nc, inv_l, box, cells, offsets,
                      mx, my, mz, mw, orig, n_big)
                  : mesh_scatter_order<uint64_t>(x, y, z, w, n, nc, inv_l, box, cells, offsets,
                      mx, my, mz, mw, orig, n_big);
    } else if (narrow) {
      st = mesh_by_radix(x, y, z, w, (uint32_t)n, nc, inv_l, box, cells, mx, my, mz, mw, orig);
    } else {
      error_set(SIF_ERR_RANGE, "chain mesh: a cell of %u particles among %llu (cells of more "
        "than %u need fewer than 2^32 particles)", largest, (unsigned long long)n, cap);
      st = SIF_ERR_RANGE;
    }
    if (st != SIF_OK)
      return st;
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
