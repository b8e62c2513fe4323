// Device-wide prefix sums, stream compaction and a stable LSD radix sort,
// shared by the chain mesh and the candidate scan. Everything here is integer
// and order-preserving, so results do not depend on launch shapes.
#include "common.h"
#include "internal.h"
#include "primitives.h"

namespace sif_amd {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr uint64_t SCAN_TILE = (uint64_t)SCAN_THREADS * SCAN_ITEMS;

// Block-wide exclusive scan of one value per thread (64-wide wavefronts).
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t* total) {
  __shared__ uint64_t wave_sums[SCAN_THREADS / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t y = __shfl_up(x, off, 64);
    if (lane >= off)
      x += y;
  }
  if (lane == 63)
    wave_sums[wave] = x;
  __syncthreads();
  uint64_t before = 0, all = 0;
  for (int w = 0; w < SCAN_THREADS / 64; w++) {
    if (w < wave)
      before += wave_sums[w];
    all += wave_sums[w];
  }
  __syncthreads();
  if (total)
    *total = all;
  return before + x - v;
}

__global__ void __launch_bounds__(SCAN_THREADS) tile_sums_u32(const uint32_t* __restrict__ in,
  uint64_t n, uint64_t* __restrict__ sums) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  uint64_t s = 0;
  for (int j = 0; j < SCAN_ITEMS; j++)
    if (base + j < n)
      s += in[base + j];
  uint64_t total;
  block_exclusive_scan(s, &total);
  if (threadIdx.x == 0)
    sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_THREADS) tile_sums_u64(const uint64_t* __restrict__ in,
  uint64_t n, uint64_t* __restrict__ sums) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  uint64_t s = 0;
  for (int j = 0; j < SCAN_ITEMS; j++)
    if (base + j < n)
      s += in[base + j];
  uint64_t total;
  block_exclusive_scan(s, &total);
  if (threadIdx.x == 0)
    sums[blockIdx.x] = total;
}

template <class T>
__global__ void __launch_bounds__(SCAN_THREADS) tile_scan(const T* __restrict__ in, uint64_t n,
  const uint64_t* __restrict__ tile_offsets, uint64_t* __restrict__ out,
  uint64_t* __restrict__ total_out) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  uint64_t v[SCAN_ITEMS];
  uint64_t s = 0;
  for (int j = 0; j < SCAN_ITEMS; j++) {
    v[j] = base + j < n ? (uint64_t)in[base + j] : 0;
    s += v[j];
  }
  uint64_t run = block_exclusive_scan(s, nullptr) + (tile_offsets ? tile_offsets[blockIdx.x] : 0);
  for (int j = 0; j < SCAN_ITEMS; j++) {
    if (base + j < n)
      out[base + j] = run;
    run += v[j];
  }
  // The last thread of the last tile ends on the grand total.
  if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1)
    *total_out = run;
}

// Exclusive scan of 64-bit values in place-compatible form (out may not
// alias in). Recursive over tiles.
static int scan_u64_rec(const uint64_t* in, uint64_t n, uint64_t* out,
  uint64_t* total_dev, DeviceBuffer* scratch, int depth);

static int scan_tiles(uint64_t n_tiles, uint64_t** tile_offsets,
  DeviceBuffer* scratch, int depth, const void* in, uint64_t n, bool is32) {
  uint64_t* sums = scratch[depth].as<uint64_t>();
  if (is32)
    tile_sums_u32<<<(unsigned)n_tiles, SCAN_THREADS, 0, stream()>>>(
      (const uint32_t*)in, n, sums);
  else
    tile_sums_u64<<<(unsigned)n_tiles, SCAN_THREADS, 0, stream()>>>(
      (const uint64_t*)in, n, sums);
  SIF_HIP(hipGetLastError());
  uint64_t* offs = sums + n_tiles;
  int st = scan_u64_rec(sums, n_tiles, offs, nullptr, scratch, depth + 1);
  if (st != SIF_OK)
    return st;
  *tile_offsets = offs;
  return SIF_OK;
}

static int scan_u64_rec(const uint64_t* in, uint64_t n, uint64_t* out,
  uint64_t* total_dev, DeviceBuffer* scratch, int depth) {
  const uint64_t n_tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (depth >= 4) {
    error_set(SIF_ERR_RANGE, "scan: %llu elements is beyond the supported size",
      (unsigned long long)n);
    return SIF_ERR_RANGE;
  }
  int st = scratch[depth].reserve(2 * (n_tiles + 1) * sizeof(uint64_t));
  if (st != SIF_OK)
    return st;
  uint64_t* offs = nullptr;
  if (n_tiles > 1) {
    st = scan_tiles(n_tiles, &offs, scratch, depth, in, n, false);
    if (st != SIF_OK)
      return st;
  }
  tile_scan<uint64_t><<<(unsigned)(n_tiles ? n_tiles : 1), SCAN_THREADS, 0, stream()>>>(
    in, n, offs, out, total_dev);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

namespace {
thread_local DeviceBuffer g_scan_scratch[4];
}

// The whole scan of a short array in one workgroup (one launch instead of
// three): the array is staged in LDS with coalesced loads, each thread sums a
// contiguous run of it, the runs' sums are scanned across the workgroup, and
// each thread writes its run's prefixes.
constexpr uint32_t ONE_SCAN_THREADS = 1024;
constexpr uint64_t ONE_SCAN_MAX = 16384;

__global__ void __launch_bounds__(ONE_SCAN_THREADS) one_block_scan_u32(
  const uint32_t* __restrict__ in, uint32_t n, uint64_t* __restrict__ out,
  uint64_t* __restrict__ total) {
  __shared__ uint64_t wave_sums[ONE_SCAN_THREADS / 64];
  __shared__ uint32_t vals[ONE_SCAN_MAX];
  for (uint32_t i = threadIdx.x; i < n; i += ONE_SCAN_THREADS)
    vals[i] = in[i];
  __syncthreads();
  const uint32_t per = (n + ONE_SCAN_THREADS - 1) / ONE_SCAN_THREADS;
  const uint32_t b = min(threadIdx.x * per, n), e = min(b + per, n);
  uint64_t v = 0;
  for (uint32_t i = b; i < e; i++)
    v += vals[i];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t y = __shfl_up(x, off, 64);
    if (lane >= off)
      x += y;
  }
  if (lane == 63)
    wave_sums[wave] = x;
  __syncthreads();
  uint64_t run = x - v, all = 0;
  for (int w = 0; w < (int)(ONE_SCAN_THREADS / 64); w++) {
    if (w < wave)
      run += wave_sums[w];
    all += wave_sums[w];
  }
  for (uint32_t i = b; i < e; i++) {
    out[i] = run;
    run += vals[i];
  }
  if (total && threadIdx.x == 0)
    *total = all;
}

int exclusive_scan_u32(const uint32_t* in, uint64_t n, uint64_t* out,
  uint64_t* total_dev) {
  if (n > SCAN_TILE && n <= ONE_SCAN_MAX) {  // one tile is one launch already
    one_block_scan_u32<<<1, ONE_SCAN_THREADS, 0, stream()>>>(in, (uint32_t)n, out, total_dev);
    SIF_HIP(hipGetLastError());
    return SIF_OK;
  }
  const uint64_t n_tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  int st = g_scan_scratch[0].reserve(2 * (n_tiles + 1) * sizeof(uint64_t));
  if (st != SIF_OK)
    return st;
  uint64_t* offs = nullptr;
  if (n_tiles > 1) {
    st = scan_tiles(n_tiles, &offs, g_scan_scratch, 0, in, n, true);
    if (st != SIF_OK)
      return st;
  }
  tile_scan<uint32_t><<<(unsigned)(n_tiles ? n_tiles : 1), SCAN_THREADS, 0, stream()>>>(
    in, n, offs, out, total_dev);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

int exclusive_scan_u64(const uint64_t* in, uint64_t n, uint64_t* out,
  uint64_t* total_dev) {
  return scan_u64_rec(in, n, out, total_dev, g_scan_scratch, 0);
}

// --- stable LSD radix sort of (u32 key, u32 value), 4-bit digits ---

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr uint32_t RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_RADIX = 16;

__global__ void __launch_bounds__(RS_THREADS) radix_hist(const uint32_t* __restrict__ keys,
  uint32_t n, int shift, uint32_t n_tiles, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[RS_RADIX];
  if (threadIdx.x < RS_RADIX)
    h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * RS_TILE + threadIdx.x * RS_ITEMS;
  for (int j = 0; j < RS_ITEMS; j++)
    if (base + j < n)
      atomicAdd(&h[(keys[base + j] >> shift) & (RS_RADIX - 1)], 1u);
  __syncthreads();
  if (threadIdx.x < RS_RADIX)
    hist[(uint64_t)threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

// Each thread owns RS_ITEMS consecutive items, so block order = thread-major
// then item order, and the rank of an item among equal digits follows it.
__global__ void __launch_bounds__(RS_THREADS) radix_scatter(const uint32_t* __restrict__ keys,
  const uint32_t* __restrict__ vals, uint32_t n, int shift, uint32_t n_tiles,
  const uint64_t* __restrict__ offsets, uint32_t* __restrict__ keys_out,
  uint32_t* __restrict__ vals_out) {
  __shared__ uint32_t cnt[RS_RADIX][RS_THREADS + 1];  // padded: the segment sums read across digits
  const uint32_t base = blockIdx.x * RS_TILE + threadIdx.x * RS_ITEMS;
  uint32_t k[RS_ITEMS], v[RS_ITEMS];
  uint32_t local[RS_RADIX];
  for (int d = 0; d < RS_RADIX; d++)
    local[d] = 0;
  for (int j = 0; j < RS_ITEMS; j++) {
    if (base + j < n) {
      k[j] = keys[base + j];
      v[j] = vals[base + j];
      local[(k[j] >> shift) & (RS_RADIX - 1)]++;
    }
  }
  for (int d = 0; d < RS_RADIX; d++)
    cnt[d][threadIdx.x] = local[d];
  __syncthreads();
  // exclusive scan along threads for each digit, in three short steps: thread
  // (digit d, segment g) sums its 16 threads' counts, one thread a digit
  // scans the 16 segment sums, and each segment writes its prefixes.
  __shared__ uint32_t seg_sum[RS_RADIX][RS_THREADS / 16];
  {
    const int d = threadIdx.x & (RS_RADIX - 1), g = threadIdx.x >> 4;
    uint32_t sum = 0;
    for (int t = 0; t < 16; t++)
      sum += cnt[d][g * 16 + t];
    seg_sum[d][g] = sum;
    __syncthreads();
    if (threadIdx.x < RS_RADIX) {
      uint32_t run = 0;
      for (int q = 0; q < RS_THREADS / 16; q++) {
        const uint32_t c = seg_sum[threadIdx.x][q];
        seg_sum[threadIdx.x][q] = run;
        run += c;
      }
    }
    __syncthreads();
    uint32_t run = seg_sum[d][g];
    for (int t = 0; t < 16; t++) {
      const uint32_t c = cnt[d][g * 16 + t];
      cnt[d][g * 16 + t] = run;
      run += c;
    }
  }
  __syncthreads();
  uint32_t pos[RS_RADIX];
  for (int d = 0; d < RS_RADIX; d++)
    pos[d] = (uint32_t)offsets[(uint64_t)d * n_tiles + blockIdx.x] + cnt[d][threadIdx.x];
  for (int j = 0; j < RS_ITEMS; j++) {
    if (base + j < n) {
      const uint32_t d = (k[j] >> shift) & (RS_RADIX - 1);
      keys_out[pos[d]] = k[j];
      vals_out[pos[d]] = v[j];
      pos[d]++;
    }
  }
}

namespace {
thread_local DeviceBuffer g_rs_hist, g_rs_offs, g_rs_k, g_rs_v;
}

// --- the same sort with 8-bit digits for short lists (half the passes) ---
// A tile is R8_ROUNDS rounds of one item a thread, taken in index order; the
// rank of an item among its tile's items of the same digit comes from the
// items of earlier rounds (run), of earlier waves in its round (wave counts)
// and of lower lanes in its wave (the lanes whose digit matches, found by
// eight ballots).
constexpr int R8_RADIX = 256;
constexpr int R8_THREADS = 256;
constexpr int R8_ROUNDS = 8;
constexpr uint32_t R8_TILE = R8_THREADS * R8_ROUNDS;
constexpr uint32_t R8_MAX_TILES = 64;  // 256 x 64 counts: a one-workgroup scan

__global__ void __launch_bounds__(R8_THREADS) radix8_hist(const uint32_t* __restrict__ keys,
  uint32_t n, int shift, uint32_t n_tiles, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[R8_RADIX];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * R8_TILE;
  for (int j = 0; j < R8_ROUNDS; j++) {
    const uint32_t i = base + j * R8_THREADS + threadIdx.x;
    if (i < n)
      atomicAdd(&h[(keys[i] >> shift) & (R8_RADIX - 1)], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(R8_THREADS) radix8_scatter(const uint32_t* __restrict__ keys,
  const uint32_t* __restrict__ vals, uint32_t n, int shift, uint32_t n_tiles,
  const uint64_t* __restrict__ offsets, uint32_t* __restrict__ keys_out,
  uint32_t* __restrict__ vals_out) {
  __shared__ uint32_t run[R8_RADIX];
  __shared__ uint32_t wcount[R8_THREADS / 64][R8_RADIX];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t below = (1ull << lane) - 1ull;
  run[threadIdx.x] = (uint32_t)offsets[(uint64_t)threadIdx.x * n_tiles + blockIdx.x];
  const uint32_t base = blockIdx.x * R8_TILE;
  for (int j = 0; j < R8_ROUNDS; j++) {
    const uint32_t i = base + j * R8_THREADS + threadIdx.x;
    const bool valid = i < n;
    uint32_t k = 0, v = 0;
    if (valid) {
      k = keys[i];
      v = vals[i];
    }
    const uint32_t d = (k >> shift) & (R8_RADIX - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const int bit = (d >> b) & 1;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    // thread t clears the counts it summed into run[t] last round
#pragma unroll
    for (int w = 0; w < R8_THREADS / 64; w++)
      wcount[w][threadIdx.x] = 0;
    __syncthreads();
    const uint32_t rank = (uint32_t)__popcll(peers & below);
    if (valid && rank == 0)
      wcount[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + rank;
      for (int w = 0; w < wave; w++)
        pos += wcount[w][d];
      keys_out[pos] = k;
      vals_out[pos] = v;
    }
    __syncthreads();  // every read of run before it moves on
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < R8_THREADS / 64; w++)
      add += wcount[w][threadIdx.x];
    run[threadIdx.x] += add;
  }
}

int radix_sort_u32(uint32_t* keys, uint32_t* vals, uint32_t n, int bits, uint32_t differ) {
  if (n <= 1)
    return SIF_OK;
  const uint32_t n_tiles8 = (n + R8_TILE - 1) / R8_TILE;
  if (n_tiles8 <= R8_MAX_TILES) {
    const uint64_t hsize = (uint64_t)R8_RADIX * n_tiles8;
    int st = g_rs_hist.reserve(hsize * sizeof(uint32_t));
    if (st == SIF_OK)
      st = g_rs_offs.reserve((hsize + 1) * sizeof(uint64_t));
    if (st == SIF_OK)
      st = g_rs_k.reserve((uint64_t)n * sizeof(uint32_t));
    if (st == SIF_OK)
      st = g_rs_v.reserve((uint64_t)n * sizeof(uint32_t));
    if (st != SIF_OK)
      return st;
    uint32_t *ka = keys, *va = vals, *kb = g_rs_k.as<uint32_t>(), *vb = g_rs_v.as<uint32_t>();
    for (int p = 0; p < (bits + 7) / 8; p++) {
      const int shift = 8 * p;
      if (((differ >> shift) & (R8_RADIX - 1)) == 0)
        continue;  // one digit value in every key: the identity permutation
      radix8_hist<<<n_tiles8, R8_THREADS, 0, stream()>>>(ka, n, shift, n_tiles8,
        g_rs_hist.as<uint32_t>());
      SIF_HIP(hipGetLastError());
      st = exclusive_scan_u32(g_rs_hist.as<uint32_t>(), hsize, g_rs_offs.as<uint64_t>(), nullptr);
      if (st != SIF_OK)
        return st;
      radix8_scatter<<<n_tiles8, R8_THREADS, 0, stream()>>>(ka, va, n, shift, n_tiles8,
        g_rs_offs.as<uint64_t>(), kb, vb);
      SIF_HIP(hipGetLastError());
      std::swap(ka, kb);
      std::swap(va, vb);
    }
    if (ka != keys) {
      SIF_HIP(hipMemcpyAsync(keys, ka, (uint64_t)n * 4, hipMemcpyDeviceToDevice, stream()));
      SIF_HIP(hipMemcpyAsync(vals, va, (uint64_t)n * 4, hipMemcpyDeviceToDevice, stream()));
    }
    return SIF_OK;
  }
  const uint32_t n_tiles = (n + RS_TILE - 1) / RS_TILE;
  const uint64_t hsize = (uint64_t)RS_RADIX * n_tiles;
  int st = g_rs_hist.reserve(hsize * sizeof(uint32_t));
  if (st == SIF_OK)
    st = g_rs_offs.reserve((hsize + 1) * sizeof(uint64_t));
  if (st == SIF_OK)
    st = g_rs_k.reserve((uint64_t)n * sizeof(uint32_t));
  if (st == SIF_OK)
    st = g_rs_v.reserve((uint64_t)n * sizeof(uint32_t));
  if (st != SIF_OK)
    return st;
  uint32_t *ka = keys, *va = vals, *kb = g_rs_k.as<uint32_t>(), *vb = g_rs_v.as<uint32_t>();
  int passes = (bits + 3) / 4;
  for (int p = 0; p < passes; p++) {
    const int shift = 4 * p;
    if (((differ >> shift) & (RS_RADIX - 1)) == 0)
      continue;  // one digit value in every key: the identity permutation
    radix_hist<<<n_tiles, RS_THREADS, 0, stream()>>>(ka, n, shift, n_tiles,
      g_rs_hist.as<uint32_t>());
    SIF_HIP(hipGetLastError());
    st = exclusive_scan_u32(g_rs_hist.as<uint32_t>(), hsize, g_rs_offs.as<uint64_t>(), nullptr);
    if (st != SIF_OK)
      return st;
    radix_scatter<<<n_tiles, RS_THREADS, 0, stream()>>>(ka, va, n, shift, n_tiles,
      g_rs_offs.as<uint64_t>(), kb, vb);
    SIF_HIP(hipGetLastError());
    uint32_t* t = ka;
    ka = kb;
    kb = t;
    t = va;
    va = vb;
    vb = t;
  }
  if (ka != keys) {
    SIF_HIP(hipMemcpyAsync(keys, ka, (uint64_t)n * 4, hipMemcpyDeviceToDevice, stream()));
    SIF_HIP(hipMemcpyAsync(vals, va, (uint64_t)n * 4, hipMemcpyDeviceToDevice, stream()));
  }
  return SIF_OK;
}

// --- order-preserving compaction of u32 values by keep flags ---

__global__ void compact_scatter(const uint32_t* __restrict__ in,
  const uint32_t* __restrict__ keep, const uint64_t* __restrict__ pos, uint64_t n,
  uint32_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && keep[i])
    out[pos[i]] = in[i];
}

namespace {
thread_local DeviceBuffer g_cpos, g_ccount;
}

__global__ void compact2_scatter(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
  const uint32_t* __restrict__ keep, const uint64_t* __restrict__ pos, uint64_t n,
  uint32_t* __restrict__ out_a, uint32_t* __restrict__ out_b) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && keep[i]) {
    out_a[pos[i]] = a[i];
    out_b[pos[i]] = b[i];
  }
}

int compact2_u32(const uint32_t* a, const uint32_t* b, const uint32_t* keep, uint64_t n,
  uint32_t* out_a, uint32_t* out_b, uint64_t* count) {
  int st = g_cpos.reserve((n ? n : 1) * 8);
  if (st == SIF_OK)
    st = g_ccount.reserve(8);
  if (st != SIF_OK)
    return st;
  st = exclusive_scan_u32(keep, n, g_cpos.as<uint64_t>(), g_ccount.as<uint64_t>());
  if (st != SIF_OK)
    return st;
  if (n)
    compact2_scatter<<<grid_for(n, 256), 256, 0, stream()>>>(a, b, keep,
      g_cpos.as<uint64_t>(), n, out_a, out_b);
  SIF_HIP(hipGetLastError());
  SIF_HIP(hipMemcpyAsync(count, g_ccount.ptr, 8, hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

namespace {
thread_local uint64_t* g_ccount_host = nullptr;  // pinned
thread_local hipEvent_t g_ccount_ready = nullptr;
}

int compact2_u32_begin(const uint32_t* a, const uint32_t* b, const uint32_t* keep, uint64_t n,
  uint32_t* out_a, uint32_t* out_b) {
  int st = g_cpos.reserve((n ? n : 1) * 8);
  if (st == SIF_OK)
    st = g_ccount.reserve(8);
  if (st != SIF_OK)
    return st;
  if (!g_ccount_host)
    SIF_HIP(hipHostMalloc((void**)&g_ccount_host, 8, hipHostMallocDefault));
  if (!g_ccount_ready)
    SIF_HIP(hipEventCreateWithFlags(&g_ccount_ready, hipEventDisableTiming));
  st = exclusive_scan_u32(keep, n, g_cpos.as<uint64_t>(), g_ccount.as<uint64_t>());
  if (st != SIF_OK)
    return st;
  if (n)
    compact2_scatter<<<grid_for(n, 256), 256, 0, stream()>>>(a, b, keep,
      g_cpos.as<uint64_t>(), n, out_a, out_b);
  SIF_HIP(hipGetLastError());
  SIF_HIP(hipMemcpyAsync(g_ccount_host, g_ccount.ptr, 8, hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipEventRecord(g_ccount_ready, stream()));
  return SIF_OK;
}

int compact2_u32_finish(uint64_t* count) {
  if (int w = wait_event(g_ccount_ready))
    return w;
  *count = *g_ccount_host;
  return SIF_OK;
}

int compact_u32(const uint32_t* in, const uint32_t* keep, uint64_t n,
  uint32_t* out, uint64_t* count) {
  int st = g_cpos.reserve((n ? n : 1) * 8);
  if (st == SIF_OK)
    st = g_ccount.reserve(8);
  if (st != SIF_OK)
    return st;
  st = exclusive_scan_u32(keep, n, g_cpos.as<uint64_t>(), g_ccount.as<uint64_t>());
  if (st != SIF_OK)
    return st;
  if (n)
    compact_scatter<<<grid_for(n, 256), 256, 0, stream()>>>(in, keep,
      g_cpos.as<uint64_t>(), n, out);
  SIF_HIP(hipGetLastError());
  SIF_HIP(hipMemcpyAsync(count, g_ccount.ptr, 8, hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

}  // namespace sif_amd
