// Candidate scan: every cell with delta <= threshold whose mask bit is clear,
// compacted in cell order and sorted deepest first. Under
// SIF_FINDER_SEEDS_MINIMA only the local minima among them (on the 27 nodes
// around each) are kept.
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

// The scan is a two-pass compaction in tiles of 4096 consecutive cells: each
// thread owns 16 consecutive cells (four float4 loads and one 16-bit slice of
// the mask), a tile's count and then its in-tile offsets come from a block
// scan, and candidates are written in flat order. delta is read twice and
// nothing else of the grid's size is written.
constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 16;
constexpr uint64_t SC_TILE = (uint64_t)SC_THREADS * SC_ITEMS;

// Node i is a local minimum of the periodic n^3 field: strictly below each of
// its 26 neighbours, an equal neighbour winning when its flat index is lower,
// so of a run of equal values only the lowest-indexed node is a minimum
// (utils.c:228-262). A NaN neighbour makes no minimum. Only below-threshold
// cells come here, a few per cent of the grid; their stencils are L2 hits.
__device__ __forceinline__ bool node_is_minimum(const float* __restrict__ f, uint32_t n,
  uint64_t i) {
  const uint64_t nn = (uint64_t)n * n;
  const uint32_t ix = (uint32_t)(i / nn), iy = (uint32_t)((i / n) % n), iz = (uint32_t)(i % n);
  const float v = f[i];
  const uint32_t xs[3] = {ix ? ix - 1 : n - 1, ix, ix + 1 == n ? 0 : ix + 1};
  const uint32_t ys[3] = {iy ? iy - 1 : n - 1, iy, iy + 1 == n ? 0 : iy + 1};
  const uint32_t zs[3] = {iz ? iz - 1 : n - 1, iz, iz + 1 == n ? 0 : iz + 1};
  bool ok = true;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      const uint64_t row = (uint64_t)xs[a] * nn + (uint64_t)ys[b] * n;
      for (int c = 0; c < 3; c++) {
        const uint64_t j = row + zs[c];
        if (a == 1 && b == 1 && c == 1)
          continue;
        const float u = f[j];
        ok = ok && (j < i ? v < u : v <= u);
      }
    }
  return ok;
}

// Bit j of the result: cell first + j is a candidate (delta <= t, mask clear,
// and a local minimum when minima_n, the grid's side, is not 0).
__device__ __forceinline__ uint32_t cand_bits(const float* __restrict__ delta,
  const uint64_t* __restrict__ mask, uint64_t total, float t, uint64_t first,
  uint32_t minima_n) {
  uint32_t bits = 0;
  if (first + SC_ITEMS <= total) {
    const float4* d4 = reinterpret_cast<const float4*>(delta + first);
    for (int q = 0; q < SC_ITEMS / 4; q++) {
      const float4 v = d4[q];
      bits |= (uint32_t)(v.x <= t) << (4 * q);
      bits |= (uint32_t)(v.y <= t) << (4 * q + 1);
      bits |= (uint32_t)(v.z <= t) << (4 * q + 2);
      bits |= (uint32_t)(v.w <= t) << (4 * q + 3);
    }
    if (mask)
      bits &= ~(uint32_t)(mask[first >> 6] >> (first & 63)) & 0xffffu;
  } else {
    for (int j = 0; j < SC_ITEMS; j++) {
      const uint64_t i = first + j;
      if (i < total && delta[i] <= t && !(mask && ((mask[i >> 6] >> (i & 63)) & 1ull)))
        bits |= 1u << j;
    }
  }
  if (minima_n)
    for (uint32_t b = bits; b; b &= b - 1) {
      const int j = __ffs(b) - 1;
      if (!node_is_minimum(delta, minima_n, first + j))
        bits &= ~(1u << j);
    }
  return bits;
}

// Block-wide exclusive scan of one count per thread.
__device__ __forceinline__ uint32_t sc_block_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wave_sums[SC_THREADS / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off)
      x += y;
  }
  if (lane == 63)
    wave_sums[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (int w = 0; w < SC_THREADS / 64; w++) {
    if (w < wave)
      before += wave_sums[w];
    all += wave_sums[w];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

// Also the AND and the OR of each tile's candidate keys (tile_span[2 tile],
// [2 tile + 1]): the sort skips the digits on which every key agrees.
__global__ void __launch_bounds__(SC_THREADS) cand_count_kernel(
  const float* __restrict__ delta, const uint64_t* __restrict__ mask, uint64_t total,
  float t, uint32_t minima_n, uint64_t* __restrict__ tile_counts,
  uint32_t* __restrict__ tile_span) {
  __shared__ uint32_t w_and[SC_THREADS / 64], w_or[SC_THREADS / 64];
  const uint64_t first = (uint64_t)blockIdx.x * SC_TILE + (uint64_t)threadIdx.x * SC_ITEMS;
  const uint32_t bits = cand_bits(delta, mask, total, t, first, minima_n);
  uint32_t k_and = 0xffffffffu, k_or = 0u;
  for (uint32_t b = bits; b; b &= b - 1) {
    const uint32_t k = float_key(delta[first + (__ffs(b) - 1)]);
    k_and &= k;
    k_or |= k;
  }
  for (int off = 32; off > 0; off >>= 1) {
    k_and &= __shfl_xor(k_and, off, 64);
    k_or |= __shfl_xor(k_or, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    w_and[threadIdx.x >> 6] = k_and;
    w_or[threadIdx.x >> 6] = k_or;
  }
  uint32_t all;
  sc_block_scan((uint32_t)__popc(bits), &all);  // its barriers publish w_and, w_or
  if (threadIdx.x == 0) {
    tile_counts[blockIdx.x] = all;
    uint32_t a = 0xffffffffu, o = 0u;
    for (int w = 0; w < SC_THREADS / 64; w++) {
      a &= w_and[w];
      o |= w_or[w];
    }
    tile_span[2 * (uint64_t)blockIdx.x] = a;
    tile_span[2 * (uint64_t)blockIdx.x + 1] = o;
  }
}

// The tiles' spans into one (span[0] AND, span[1] OR), one block.
// The keys' AND and OR over all tiles and, for a scan of cell lists, the
// cells listed (the per-tile counts summed here, not by an atomic per tile).
__global__ void __launch_bounds__(1024) span_reduce_kernel(const uint32_t* __restrict__ tile_span,
  uint64_t tiles, uint32_t* __restrict__ span, const uint32_t* __restrict__ tile_listed,
  uint64_t* __restrict__ listed) {
  __shared__ uint32_t w_and[16], w_or[16];
  __shared__ uint64_t w_sum[16];
  uint32_t a = 0xffffffffu, o = 0u;
  uint64_t c = 0;
  for (uint64_t i = threadIdx.x; i < tiles; i += blockDim.x) {
    a &= tile_span[2 * i];
    o |= tile_span[2 * i + 1];
    if (tile_listed)
      c += tile_listed[i];
  }
  for (int off = 32; off > 0; off >>= 1) {
    a &= __shfl_xor(a, off, 64);
    o |= __shfl_xor(o, off, 64);
    c += __shfl_xor(c, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    w_and[threadIdx.x >> 6] = a;
    w_or[threadIdx.x >> 6] = o;
    w_sum[threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      a &= w_and[w];
      o |= w_or[w];
      c += w_sum[w];
    }
    span[0] = a;
    span[1] = o;
    if (tile_listed)
      *listed = c;
  }
}

__global__ void __launch_bounds__(SC_THREADS) cand_write_kernel(
  const float* __restrict__ delta, const uint64_t* __restrict__ mask, uint64_t total,
  float t, uint32_t minima_n, const uint64_t* __restrict__ tile_offsets,
  uint32_t* __restrict__ keys, uint32_t* __restrict__ flat, const uint64_t* __restrict__ found,
  uint64_t cap) {
  if (*found > cap)  // the buffers are short: the caller grows them and writes again
    return;
  const uint64_t first = (uint64_t)blockIdx.x * SC_TILE + (uint64_t)threadIdx.x * SC_ITEMS;
  uint32_t bits = cand_bits(delta, mask, total, t, first, minima_n);
  uint32_t all;
  uint64_t pos = tile_offsets[blockIdx.x] + sc_block_scan((uint32_t)__popc(bits), &all);
  while (bits) {
    const int j = __ffs(bits) - 1;
    bits &= bits - 1;
    const uint64_t i = first + j;
    keys[pos] = float_key(delta[i]);
    flat[pos] = (uint32_t)i;
    pos++;
  }
}

__global__ void cand_unkey_kernel(const uint32_t* __restrict__ keys, uint64_t n,
  float* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    out[i] = key_float(keys[i]);
}

namespace {
thread_local DeviceBuffer g_tiles, g_tile_pos, g_count, g_tile_span, g_tile_listed;
}

// --- the same scan from cell lists (CellLists, a listing smoothing's) ---
//
// One workgroup per LISTS_PER_TILE list blocks (a tile), which it takes in
// order: the listed cells whose mask bit is clear, counted (with their keys'
// AND and OR), then written in list order -- cell order -- at the tile's
// offset among all tiles. Every list is read twice, nothing of the grid's
// size.
constexpr uint32_t LISTS_PER_TILE = 4;

uint64_t list_tiles(uint64_t blocks) {
  return (blocks + LISTS_PER_TILE - 1) / LISTS_PER_TILE;
}
__global__ void __launch_bounds__(SC_THREADS) list_count_kernel(
  const uint32_t* __restrict__ lkeys, const uint16_t* __restrict__ llocal,
  const uint32_t* __restrict__ lcount, uint64_t blocks, const uint64_t* __restrict__ mask,
  uint64_t* __restrict__ tile_counts, uint32_t* __restrict__ tile_span,
  uint32_t* __restrict__ tile_listed) {
  __shared__ uint32_t w_and[SC_THREADS / 64], w_or[SC_THREADS / 64], w_cnt[SC_THREADS / 64];
  uint32_t cnt = 0, k_and = 0xffffffffu, k_or = 0u, c_all = 0;
  for (uint32_t j = 0; j < LISTS_PER_TILE; j++) {
    const uint64_t b = (uint64_t)blockIdx.x * LISTS_PER_TILE + j;
    if (b >= blocks)
      break;
    const uint64_t base = b * LIST_BLOCK;
    const uint32_t c = lcount[b];
    c_all += c;
    for (uint32_t i = threadIdx.x; i < c; i += SC_THREADS) {
      const uint64_t cell = base + llocal[base + i];
      if (!((mask[cell >> 6] >> (cell & 63)) & 1ull)) {
        const uint32_t k = lkeys[base + i];
        cnt++;
        k_and &= k;
        k_or |= k;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    cnt += __shfl_xor(cnt, off, 64);
    k_and &= __shfl_xor(k_and, off, 64);
    k_or |= __shfl_xor(k_or, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    w_and[threadIdx.x >> 6] = k_and;
    w_or[threadIdx.x >> 6] = k_or;
    w_cnt[threadIdx.x >> 6] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t a = 0xffffffffu, o = 0u, all = 0;
    for (int w = 0; w < SC_THREADS / 64; w++) {
      a &= w_and[w];
      o |= w_or[w];
      all += w_cnt[w];
    }
    tile_counts[blockIdx.x] = all;
    tile_span[2 * (uint64_t)blockIdx.x] = a;
    tile_span[2 * (uint64_t)blockIdx.x + 1] = o;
    tile_listed[blockIdx.x] = c_all;
  }
}

__global__ void __launch_bounds__(SC_THREADS) list_write_kernel(
  const uint32_t* __restrict__ lkeys, const uint16_t* __restrict__ llocal,
  const uint32_t* __restrict__ lcount, uint64_t blocks, const uint64_t* __restrict__ mask,
  const uint64_t* __restrict__ tile_offsets, uint32_t* __restrict__ keys,
  uint32_t* __restrict__ flat, const uint64_t* __restrict__ found, uint64_t cap) {
  if (*found > cap)  // the buffers are short: the caller grows them and writes again
    return;
  uint64_t pos = tile_offsets[blockIdx.x];
  for (uint32_t j = 0; j < LISTS_PER_TILE; j++) {
    const uint64_t b = (uint64_t)blockIdx.x * LISTS_PER_TILE + j;
    if (b >= blocks)
      break;
    const uint64_t base = b * LIST_BLOCK;
    const uint32_t c = lcount[b];
    // chunks of SC_THREADS entries in list order; a block scan places each
    // kept one (every thread runs every chunk: the scan's barriers)
    for (uint32_t i0 = 0; i0 < c; i0 += SC_THREADS) {
      const uint32_t i = i0 + threadIdx.x;
      uint64_t cell = 0;
      bool keep = false;
      if (i < c) {
        cell = base + llocal[base + i];
        keep = !((mask[cell >> 6] >> (cell & 63)) & 1ull);
      }
      uint32_t all;
      const uint32_t at = sc_block_scan(keep ? 1u : 0u, &all);
      if (keep) {
        keys[pos + at] = lkeys[base + i];
        flat[pos + at] = (uint32_t)cell;
      }
      pos += all;
    }
  }
}

namespace {
thread_local uint64_t* g_back = nullptr;  // pinned: the count, AND | OR << 32, cells listed
thread_local hipEvent_t g_back_ready = nullptr;
}

// The counting pass of a scan job, and the copy of its results to host
// memory (scan_finish reads them).
// The writing pass into keys/flat (capacity cap): writes nothing when the
// count (on the device) exceeds cap.
static int scan_write(const ScanJob& J, uint32_t* keys, uint32_t* flat, uint64_t cap,
  uint64_t tiles) {
  if (J.lists)
    list_write_kernel<<<(unsigned)tiles, SC_THREADS, 0, stream()>>>(J.lists->keys,
      J.lists->local, J.lists->count, J.lists->blocks, J.mask, g_tile_pos.as<uint64_t>(),
      keys, flat, g_count.as<uint64_t>(), cap);
  else
    cand_write_kernel<<<(unsigned)tiles, SC_THREADS, 0, stream()>>>(J.delta, J.mask, J.total,
      J.t, J.minima_n, g_tile_pos.as<uint64_t>(), keys, flat, g_count.as<uint64_t>(), cap);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

int scan_begin(ScanJob& J) {
  J.count = 0;
  J.key_differ = 0xffffffffu;
  J.listed = 0;
  J.begun = false;
  if (J.total >= 0x80000000ull) {
    error_set(SIF_ERR_RANGE, "scan: %llu cells (at most 2^31 - 1)",
      (unsigned long long)J.total);
    return SIF_ERR_RANGE;
  }
  if (!J.lists) {
    if (reinterpret_cast<uintptr_t>(J.delta) % 16 != 0) {
      error_set(SIF_ERR_INVALID, "scan: the grid must be 16-byte aligned");
      return SIF_ERR_INVALID;
    }
    if (J.minima_n && (uint64_t)J.minima_n * J.minima_n * J.minima_n != J.total) {
      error_set(SIF_ERR_INVALID, "scan: minima on a %u^3 grid of %llu cells", J.minima_n,
        (unsigned long long)J.total);
      return SIF_ERR_INVALID;
    }
  }
  const uint64_t tiles = J.lists ? list_tiles(J.lists->blocks) : (J.total + SC_TILE - 1) / SC_TILE;
  int st = g_tiles.reserve((tiles ? tiles : 1) * 8);
  if (st == SIF_OK)
    st = g_tile_pos.reserve((tiles ? tiles : 1) * 8);
  if (st == SIF_OK)
    st = g_count.reserve(32);  // the count, the keys' AND and OR, the cells listed
  if (st == SIF_OK)
    st = g_tile_span.reserve((tiles ? tiles : 1) * 8);
  if (st == SIF_OK && J.lists)
    st = g_tile_listed.reserve((tiles ? tiles : 1) * 4);
  if (st != SIF_OK)
    return st;
  if (!g_back)
    SIF_HIP(hipHostMalloc((void**)&g_back, 32, hipHostMallocDefault));
  if (!g_back_ready)
    SIF_HIP(hipEventCreateWithFlags(&g_back_ready, hipEventDisableTiming));
  if (J.total == 0 || tiles == 0)
    return SIF_OK;
  timing_begin(SIF_AMD_STAGE_SCAN);
  if (J.lists) {
    list_count_kernel<<<(unsigned)tiles, SC_THREADS, 0, stream()>>>(J.lists->keys,
      J.lists->local, J.lists->count, J.lists->blocks, J.mask, g_tiles.as<uint64_t>(),
      g_tile_span.as<uint32_t>(), g_tile_listed.as<uint32_t>());
  } else {
    cand_count_kernel<<<(unsigned)tiles, SC_THREADS, 0, stream()>>>(J.delta, J.mask, J.total,
      J.t, J.minima_n, g_tiles.as<uint64_t>(), g_tile_span.as<uint32_t>());
  }
  SIF_HIP(hipGetLastError());
  span_reduce_kernel<<<1, 1024, 0, stream()>>>(g_tile_span.as<uint32_t>(), tiles,
    g_count.as<uint32_t>() + 2, J.lists ? g_tile_listed.as<uint32_t>() : nullptr,
    g_count.as<uint64_t>() + 2);
  SIF_HIP(hipGetLastError());
  st = exclusive_scan_u64(g_tiles.as<uint64_t>(), tiles, g_tile_pos.as<uint64_t>(),
    g_count.as<uint64_t>());
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemcpyAsync(g_back, g_count.ptr, J.lists ? 24 : 16, hipMemcpyDeviceToHost,
    stream()));
  SIF_HIP(hipEventRecord(g_back_ready, stream()));
  if (J.keys && J.flat) {
    st = scan_write(J, J.keys, J.flat, J.cap, tiles);
    if (st != SIF_OK)
      return st;
  }
  timing_end(SIF_AMD_STAGE_SCAN);
  J.begun = true;
  return SIF_OK;
}

// Waits for the counting pass's results and, when the count fits cap,
// enqueues the writing pass: the candidates in flat order with their keys.
// When it does not, the job stays begun: the caller grows its buffers and
// finishes again, with the same counts (so the mask must not change between).
int scan_finish(ScanJob& J, uint32_t* keys, uint32_t* flat, uint64_t cap) {
  if (!J.begun)
    return SIF_OK;
  if (int w = wait_event(g_back_ready))
    return w;
  const uint64_t n = g_back[0];
  J.count = n;
  J.key_differ = (uint32_t)g_back[1] ^ (uint32_t)(g_back[1] >> 32);
  J.listed = J.lists ? g_back[2] : 0;
  if (n > cap)
    return SIF_OK;
  J.begun = false;
  const uint64_t tiles = J.lists ? list_tiles(J.lists->blocks) : (J.total + SC_TILE - 1) / SC_TILE;
  const bool written = J.keys == keys && J.flat == flat && J.keys && J.flat && n <= J.cap;
  if (n > 0 && !written) {
    timing_begin(SIF_AMD_STAGE_SCAN);
    const int st = scan_write(J, keys, flat, cap, tiles);
    if (st != SIF_OK)
      return st;
    timing_end(SIF_AMD_STAGE_SCAN);
  }
  if (timing_active()) {
    const double written = 8.0 * (double)n;
    if (J.lists) {
      // the lists twice and the candidates once; the smoothing that made
      // the lists wrote them once
      const double listed = (double)J.listed;
      timing_bytes(SIF_AMD_STAGE_SCAN, 12.0 * listed + written);
      timing_bytes(SIF_AMD_STAGE_SMOOTH, 6.0 * listed);
    } else {  // the field and the mask once, the candidates once
      timing_bytes(SIF_AMD_STAGE_SCAN, 4.0 * (double)J.total + (double)J.total / 8.0 + written);
    }
  }
  return SIF_OK;
}

// Writes the candidates, in flat order, to keys/flat (device, capacity `cap`)
// and their number to *count; keys are the order-preserving integer images of
// the depths. When the count exceeds `cap` nothing is written and the caller
// grows the buffers and calls again. minima_n: 0 for every cell below the
// threshold; the grid's side (total = minima_n^3) for the local minima alone
// (sif__finder_scan_refined with minima_only and no refinement,
// utils.c:406-552).
int scan_candidates_flat(const float* delta, const uint64_t* mask, uint64_t total,
  float t, uint32_t minima_n, uint32_t* keys, uint32_t* flat, uint64_t cap, uint64_t* count,
  uint32_t* key_differ) {
  ScanJob J;
  J.delta = delta;
  J.mask = mask;
  J.total = total;
  J.t = t;
  J.minima_n = minima_n;
  int st = scan_begin(J);
  if (st == SIF_OK)
    st = scan_finish(J, keys, flat, cap);
  *count = J.count;
  if (key_differ)
    *key_differ = J.key_differ;
  return st;
}

int scan_lists_flat(const CellLists& lists, const uint64_t* mask, uint64_t total,
  uint32_t* keys, uint32_t* flat, uint64_t cap, uint64_t* count, uint32_t* key_differ) {
  ScanJob J;
  J.lists = &lists;
  J.mask = mask;
  J.total = total;
  int st = scan_begin(J);
  if (st == SIF_OK)
    st = scan_finish(J, keys, flat, cap);
  *count = J.count;
  if (key_differ)
    *key_differ = J.key_differ;
  return st;
}

// A list of at most SMALL_SORT candidates sorts in one workgroup: each entry
// becomes (key << 32 | its place in the list), every such value distinct, so
// the ascending order is (key, place) -- the stable order -- whatever the
// sorting network; a bitonic network in LDS, then each place's value
// gathered. One launch instead of the radix sort's few per digit; the
// network's LDS traffic grows as n log^2 n on one CU, so past 8192 entries
// the radix sort over the whole device is the faster.
constexpr uint32_t SMALL_SORT = 8192;
constexpr uint32_t SMALL_SORT_THREADS = 1024;

__global__ void __launch_bounds__(SMALL_SORT_THREADS) small_sort_kernel(
  uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, uint32_t n) {
  __shared__ uint64_t s[SMALL_SORT];
  uint32_t m = 1;
  while (m < n)
    m <<= 1;
  for (uint32_t i = threadIdx.x; i < m; i += SMALL_SORT_THREADS)
    s[i] = i < n ? ((uint64_t)keys[i] << 32) | i : ~0ull;
  __syncthreads();
  for (uint32_t k = 2; k <= m; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < m; i += SMALL_SORT_THREADS) {
        const uint32_t p = i ^ j;
        if (p > i) {
          const uint64_t a = s[i], b = s[p];
          if ((a > b) == ((i & k) == 0)) {
            s[i] = b;
            s[p] = a;
          }
        }
      }
      __syncthreads();
    }
  constexpr int PER = SMALL_SORT / SMALL_SORT_THREADS;
  uint32_t kv[PER], vv[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const uint32_t i = threadIdx.x + q * SMALL_SORT_THREADS;
    if (i < n) {
      const uint64_t e = s[i];
      kv[q] = (uint32_t)(e >> 32);
      vv[q] = vals[(uint32_t)e];
    }
  }
  __syncthreads();  // every read of vals before any write
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const uint32_t i = threadIdx.x + q * SMALL_SORT_THREADS;
    if (i < n) {
      keys[i] = kv[q];
      vals[i] = vv[q];
    }
  }
}

// Deepest first; equal depths keep their order (a stable radix sort, or the
// one-workgroup sort of (key, place)), so a list in flat order comes out
// ordered by (depth, flat index). Bit 31 of a flat value rides along
// untouched. key_differ: the bits on which the keys may differ (from the scan
// that listed them, or a superset of them).
int sort_candidates(uint32_t* keys, uint32_t* flat, uint64_t n, uint32_t key_differ) {
  if (n == 0)
    return SIF_OK;
  timing_begin(SIF_AMD_STAGE_SORT);
  int st = SIF_OK;
  if (n <= SMALL_SORT) {
    small_sort_kernel<<<1, SMALL_SORT_THREADS, 0, stream()>>>(keys, flat, (uint32_t)n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      st = hip_check(e, "small_sort_kernel");
  } else {
    st = radix_sort_u32(keys, flat, (uint32_t)n, 32, key_differ);
  }
  timing_end(SIF_AMD_STAGE_SORT);
  return st;
}

// The sorted candidates (sif__finder_scan_candidates + sort_candidates).
int scan_candidates(const float* delta, const uint64_t* mask, uint64_t total,
  float t, uint32_t minima_n, uint32_t* keys, uint32_t* flat, uint64_t cap, uint64_t* count) {
  uint32_t differ = 0xffffffffu;
  int st = scan_candidates_flat(delta, mask, total, t, minima_n, keys, flat, cap, count,
    &differ);
  if (st != SIF_OK || *count > cap)
    return st;
  return sort_candidates(keys, flat, *count, differ);
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
  st = scan_candidates(delta, mask_words, total_cells, threshold, 0, keys, out_flat,
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

int sif_amd_scan_minima(const float* delta, const uint64_t* mask_words, uint32_t n_cells,
  float threshold, float* out_delta, uint32_t* out_flat, uint64_t capacity, uint64_t* count) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!delta || !out_delta || !out_flat || !count || n_cells == 0) {
    error_set(SIF_ERR_INVALID, "sif_amd_scan_minima: invalid arguments");
    return SIF_ERR_INVALID;
  }
  uint32_t* keys = reinterpret_cast<uint32_t*>(out_delta);
  const uint64_t total = (uint64_t)n_cells * n_cells * n_cells;
  st = scan_candidates(delta, mask_words, total, threshold, n_cells, keys, out_flat, capacity,
    count);
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
