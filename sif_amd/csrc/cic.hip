// CIC deposit with deterministic fixed-point accumulation, and the density
// contrast.
//
// Replaces the deposit of src/structures/grid.c:455-746 (tile/slab counting
// sort and per-slab float accumulation) and sif_grid_to_density_contrast
// (grid.c:1177-1244) of sif v0.5.0.
//
// The eight corner weights of a particle are computed with exactly the float
// operations of grid.c:585-650. Each is rounded to an integer number of 2^-s
// units and added with a 64-bit integer atomic, so the sum in every cell is
// independent of launch shape and arrival order: the grid is bit-identical
// from run to run and from one launch configuration to the next. (The
// reference gets the same property from a fixed chunk/slab order; the bits of
// the two differ by float rounding, within 1e-5 of the peak.)
//
// The unit 2^-s is chosen per deposit so that no sum can overflow: a cell
// holds at most n |w|max (n particles, each corner at most its weight), and
// with n |w|max <= 2^e, s = 62 - e keeps every corner and every cell sum
// below 2^62. The bound depends only on the count and the largest weight,
// both order-free, so the grid stays bit-identical; small weights get finer
// units rather than losing their digits. A weight that is not finite is
// refused (SIF_ERR_RANGE).
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>

#include "common.h"
#include "internal.h"
#include "primitives.h"
#include "sif_amd.h"

namespace sif_amd {

__device__ __forceinline__ float box_fold(float v, float box) {
  // src/structures/field_internal.h:58-64
  if (v >= box)
    v -= box;
  else if (v < 0.0f)
    v += box;
  return v >= box ? 0.0f : v;
}

__device__ __forceinline__ void atomic_add_i64(int64_t* p, int64_t v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}

__device__ __forceinline__ int64_t to_fixed(float v, double unit_inv) {
  return (int64_t)__double2ll_rn((double)v * unit_inv);
}

// The CIC footprint of one particle in the reference's float operations
// (grid.c:585-650): base node (ix, iy, iz) of the scaled, folded position
// c = fold(x) * n / L, and the eight corner values in fixed point (unit_inv =
// 2^s units per 1), in corner order (dx, dy, dz) = 000, 001, 010, 011, 100,
// 101, 110, 111.
__device__ __forceinline__ void cic_corners(float cx, float cy, float cz, float pw, uint32_t N,
  double unit_inv, uint32_t* ix, uint32_t* iy, uint32_t* iz, int64_t v[8]) {
  const uint32_t rx = (uint32_t)cx, ry = (uint32_t)cy, rz = (uint32_t)cz;
  *ix = rx >= N ? 0u : rx;
  *iy = ry >= N ? 0u : ry;
  *iz = rz >= N ? 0u : rz;
  const float tx = cx - (float)rx, ty = cy - (float)ry, tz = cz - (float)rz;
  const float wx0 = 1.0f - tx, wx1 = tx;
  const float wy0 = 1.0f - ty, wy1 = ty;
  const float wz0 = 1.0f - tz, wz1 = tz;
  const float w00 = wx0 * wy0, w01 = wx0 * wy1;
  const float w10 = wx1 * wy0, w11 = wx1 * wy1;
  const float mz0 = pw * wz0, mz1 = pw * wz1;
  v[0] = to_fixed(w00 * mz0, unit_inv);
  v[1] = to_fixed(w00 * mz1, unit_inv);
  v[2] = to_fixed(w01 * mz0, unit_inv);
  v[3] = to_fixed(w01 * mz1, unit_inv);
  v[4] = to_fixed(w10 * mz0, unit_inv);
  v[5] = to_fixed(w10 * mz1, unit_inv);
  v[6] = to_fixed(w11 * mz0, unit_inv);
  v[7] = to_fixed(w11 * mz1, unit_inv);
}

// The deposit is binned by tiles of CIC_T^3 nodes. A particle belongs to the
// tile of its base node; its eight corners land in that tile or in the next
// node layer up, so a tile's particles are accumulated in LDS over
// (CIC_T + 1)^3 nodes, and the tile is flushed with one 64-bit integer atomic
// per node it touched. Integer sums do not depend on order: the grid is the
// same bits for any launch shape or particle order.
// The tile edge T is 8 or 16 nodes (cic_tile_edge(), a tuning: any edge
// gives the same bits); the LDS accumulators hold (T + 1)^3 int64.
template <uint32_t T>
__device__ __forceinline__ uint32_t cic_tile(float cx, float cy, float cz, uint32_t N,
  uint32_t nt) {
  const uint32_t rx = (uint32_t)cx, ry = (uint32_t)cy, rz = (uint32_t)cz;
  const uint32_t ix = rx >= N ? 0u : rx, iy = ry >= N ? 0u : ry, iz = rz >= N ? 0u : rz;
  return ((ix / T) * nt + iy / T) * nt + iz / T;
}

// Lanes holding the same key in a row (common: neighbouring particles share
// a tile) form a run; returns the run's first lane and sets *len on it.
__device__ __forceinline__ int key_run(uint32_t key, bool valid, int lane, uint32_t* len) {
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

template <uint32_t T>
__global__ void __launch_bounds__(256) cic_count_kernel(const float* __restrict__ xs,
  const float* __restrict__ ys, const float* __restrict__ zs, uint64_t n, uint32_t N,
  float idx, float box, uint32_t nt, uint32_t* __restrict__ counts) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = p < n;
  uint32_t key = 0xffffffffu;
  if (valid)
    key = cic_tile<T>(box_fold(xs[p], box) * idx, box_fold(ys[p], box) * idx,
      box_fold(zs[p], box) * idx, N, nt);
  uint32_t len = 0;
  const int first = key_run(key, valid, lane, &len);
  if (valid && first == lane)
    atomicAdd(&counts[key], len);
}

// Particles into tile order: the scaled, folded coordinates (and weights).
template <uint32_t T>
__global__ void __launch_bounds__(256) cic_bin_kernel(const float* __restrict__ xs,
  const float* __restrict__ ys, const float* __restrict__ zs, const float* __restrict__ ws,
  uint64_t n, uint32_t N, float idx, float box, uint32_t nt,
  const uint64_t* __restrict__ offsets, uint32_t* __restrict__ cursor,
  float* __restrict__ bx, float* __restrict__ by, float* __restrict__ bz,
  float* __restrict__ bw) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = p < n;
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
  uint32_t key = 0xffffffffu;
  if (valid) {
    cx = box_fold(xs[p], box) * idx;
    cy = box_fold(ys[p], box) * idx;
    cz = box_fold(zs[p], box) * idx;
    key = cic_tile<T>(cx, cy, cz, N, nt);
  }
  uint32_t len = 0;
  const int first = key_run(key, valid, lane, &len);
  uint32_t base = 0;
  if (valid && first == lane)
    base = atomicAdd(&cursor[key], len);
  base = __shfl(base, first, 64);
  if (valid) {
    const uint64_t slot = offsets[key] + base + (uint32_t)(lane - first);
    bx[slot] = cx;
    by[slot] = cy;
    bz[slot] = cz;
    if (bw)
      bw[slot] = ws[p];
  }
}

// One workgroup per tile: LDS accumulation over (T + 1)^3 nodes, then the
// flush.
template <uint32_t T>
__global__ void __launch_bounds__(256) cic_tile_kernel(const float* __restrict__ bx,
  const float* __restrict__ by, const float* __restrict__ bz, const float* __restrict__ bw,
  const uint64_t* __restrict__ offsets, uint32_t N, uint32_t nt, double unit_inv,
  int64_t* __restrict__ acc) {
  constexpr uint32_t CIC_T = T, CIC_L = T + 1;
  __shared__ unsigned long long L[CIC_L * CIC_L * CIC_L];
  const uint32_t t = blockIdx.x;
  const uint64_t p0 = offsets[t], p1 = offsets[t + 1];
  if (p0 == p1)
    return;
  for (uint32_t i = threadIdx.x; i < CIC_L * CIC_L * CIC_L; i += blockDim.x)
    L[i] = 0ull;
  __syncthreads();
  const uint32_t tx0 = (t / (nt * nt)) * CIC_T, ty0 = ((t / nt) % nt) * CIC_T,
                 tz0 = (t % nt) * CIC_T;
  for (uint64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    uint32_t ix, iy, iz;
    int64_t v[8];
    cic_corners(bx[p], by[p], bz[p], bw ? bw[p] : 1.0f, N, unit_inv, &ix, &iy, &iz, v);
    const uint32_t lx = ix - tx0, ly = iy - ty0, lz = iz - tz0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const uint32_t a = lx + (c >> 2), b = ly + ((c >> 1) & 1), d = lz + (c & 1);
      atomicAdd(&L[(a * CIC_L + b) * CIC_L + d], (unsigned long long)v[c]);
    }
  }
  __syncthreads();
  const uint64_t NN = (uint64_t)N * N;
  for (uint32_t i = threadIdx.x; i < CIC_L * CIC_L * CIC_L; i += blockDim.x) {
    const unsigned long long s = L[i];
    if (!s)
      continue;
    uint32_t gx = tx0 + i / (CIC_L * CIC_L), gy = ty0 + (i / CIC_L) % CIC_L,
             gz = tz0 + i % CIC_L;
    gx = gx >= N ? gx - N : gx;  // the layer past the last node wraps to 0
    gy = gy >= N ? gy - N : gy;
    gz = gz >= N ? gz - N : gz;
    atomicAdd(reinterpret_cast<unsigned long long*>(acc + gx * NN + (uint64_t)gy * N + gz), s);
  }
}

// Fixed point -> density per unit volume (grid.c:735-738 divides by l^3).
__global__ void cic_finish_kernel(const int64_t* __restrict__ acc, uint64_t total,
  double unit, float inv_vol, float* __restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += stride)
    out[i] = (float)((double)acc[i] * unit) * inv_vol;
}

// The largest |w| by its bits (sign cleared): integer max, order-free; a
// non-finite weight has the largest bits of all.
__global__ void abs_max_bits_kernel(const uint32_t* __restrict__ w, uint64_t n,
  uint32_t* __restrict__ out) {
  uint32_t m = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t b = w[i] & 0x7fffffffu;
    m = b > m ? b : m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m)
    atomicMax(out, m);
}

// Deterministic mean: block b covers cells [b*MEAN_BLOCK, (b+1)*MEAN_BLOCK);
// thread t sums cells b*MEAN_BLOCK + t + 256 j in j order, in double; the 256
// partials fold in a fixed binary tree; the host adds block sums in order.
constexpr uint32_t MEAN_BLOCK = 65536;

__global__ void __launch_bounds__(256) mean_blocks_kernel(const float* __restrict__ v,
  uint64_t total, double* __restrict__ block_sums) {
  __shared__ double s[256];
  const uint64_t base = (uint64_t)blockIdx.x * MEAN_BLOCK;
  double acc = 0.0;
  for (uint32_t j = 0; j < MEAN_BLOCK / 256; j++) {
    const uint64_t i = base + threadIdx.x + 256ull * j;
    if (i < total)
      acc += (double)v[i];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t stride = 128; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride)
      s[threadIdx.x] += s[threadIdx.x + stride];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    block_sums[blockIdx.x] = s[0];
}

__global__ void contrast_kernel(float* __restrict__ v, uint64_t total, float inv) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += stride)
    v[i] = v[i] * inv - 1.0f;
}

// Field validation of src/structures/field.c:889-935: counts outside the box
// beyond 4 float ulps of its length, and those accepted only by that slack.
__global__ void check_box_kernel(const float* __restrict__ xs,
  const float* __restrict__ ys, const float* __restrict__ zs, uint64_t n,
  float box, double tol, unsigned long long* __restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long bad = 0, edge = 0, first = ~0ull;
  const double L = (double)box;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    const float p[3] = {xs[i], ys[i], zs[i]};
    bool inside = true, ok = true;
    for (int a = 0; a < 3; a++) {
      if (!(p[a] >= 0.0f && p[a] < box))
        inside = false;
      const double d = (double)p[a];
      if (!(d >= -tol && d <= L + tol))
        ok = false;
    }
    if (inside)
      continue;
    if (ok) {
      edge++;
    } else {
      bad++;
      if (i < first)
        first = i;
    }
  }
  if (bad)
    atomicAdd(&out[0], bad);
  if (edge)
    atomicAdd(&out[1], edge);
  if (first != ~0ull)
    atomicMin(&out[2], first);
}

namespace {
thread_local DeviceBuffer g_mean_buf;
thread_local DeviceBuffer g_box_buf;
}  // namespace

// --- the density contrast with no allocation ---
//
// The block sums of the mean live in `scratch`, a fixed array in device
// memory (one per device the library runs on, held until the process ends),
// for callers that cannot allocate. Its last slot receives their ordered sum.
// Calls on one device share it: calls that overlap on a device race.
constexpr uint32_t NOALLOC_MEAN_BLOCKS = 32768;  // grids of up to 2^31 cells
static __device__ __align__(256) double scratch[NOALLOC_MEAN_BLOCKS + 1];

// The block sums added in block order by one thread: the same additions, in
// the same order, as grid_mean's loop on the host.
__global__ void ordered_sum_kernel(const double* __restrict__ blocks, uint64_t n,
  double* __restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double sum = 0.0;
    for (uint64_t b = 0; b < n; b++)
      sum += blocks[b];
    *out = sum;
  }
}

namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[SCRATCH_DEVICES] = {};

// The device address of `scratch` on the current device, looked up on the
// first call there (the name is a host-side stand-in, not an address).
int noalloc_scratch(double** out) {
  int dev = 0;
  SIF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= SCRATCH_DEVICES) {
    error_set(SIF_ERR_UNSUPPORTED, "device %d: beyond the %d devices this build serves", dev,
      SCRATCH_DEVICES);
    return SIF_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SIF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return SIF_OK;
}
}  // namespace

int grid_mean(const float* values, uint64_t total, double* mean) {
  const uint64_t blocks = (total + MEAN_BLOCK - 1) / MEAN_BLOCK;
  SIF_HIP(hipGetLastError());
  int st = g_mean_buf.reserve(blocks * sizeof(double));
  if (st != SIF_OK)
    return st;
  mean_blocks_kernel<<<(unsigned)blocks, 256, 0, stream()>>>(
    values, total, g_mean_buf.as<double>());
  SIF_HIP(hipGetLastError());
  double* host = new double[blocks];
  hipError_t e = hipMemcpyAsync(host, g_mean_buf.ptr, blocks * sizeof(double),
    hipMemcpyDeviceToHost, stream());
  if (e == hipSuccess)
    e = hipStreamSynchronize(stream());
  double sum = 0.0;
  for (uint64_t b = 0; b < blocks; b++)
    sum += host[b];
  delete[] host;
  SIF_HIP(e);
  *mean = sum / (double)total;
  return SIF_OK;
}

int density_contrast(float* values, uint64_t total, double mean) {
  const float inv = (float)(1.0 / mean);
  contrast_kernel<<<grid_for(total, 256), 256, 0, stream()>>>(values, total, inv);
  SIF_HIP(hipGetLastError());
  return SIF_OK;
}

int check_in_box(const float* x, const float* y, const float* z, uint64_t n,
  float box, uint64_t* n_bad, uint64_t* n_edge, uint64_t* first_bad) {
  int st = g_box_buf.reserve(3 * sizeof(unsigned long long));
  if (st != SIF_OK)
    return st;
  unsigned long long init[3] = {0, 0, ~0ull};
  SIF_HIP(hipMemcpyAsync(
    g_box_buf.ptr, init, sizeof(init), hipMemcpyHostToDevice, stream()));
  const double tol = (double)box * 4.0 * (double)FLT_EPSILON;
  if (n)
    check_box_kernel<<<grid_for(n, 256) > 4096 ? 4096 : grid_for(n, 256), 256,
      0, stream()>>>(x, y, z, n, box, tol, g_box_buf.as<unsigned long long>());
  SIF_HIP(hipGetLastError());
  unsigned long long out[3];
  SIF_HIP(hipMemcpyAsync(
    out, g_box_buf.ptr, sizeof(out), hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  if (n_bad)
    *n_bad = out[0];
  if (n_edge)
    *n_edge = out[1];
  if (first_bad)
    *first_bad = out[2];
  return SIF_OK;
}

namespace {
thread_local DeviceBuffer g_tile_counts, g_tile_offsets, g_tile_cursor, g_bx, g_by, g_bz, g_bw, g_wmax;
}

// The fixed-point exponent of a deposit of n particles whose largest |weight|
// is wmax (the rule above): s = 62 - e with n wmax <= 2^e.
int cic_unit_exponent(uint64_t n, float wmax) {
  const double bound = (double)n * (double)wmax;
  if (!(bound > 0.0))
    return 32;
  int e;
  (void)frexp(bound, &e);  // bound = m 2^e, m in [0.5, 1): bound < 2^e
  return 62 - e;
}

// The largest |w| of a weighted deposit on the device; SIF_ERR_RANGE for a
// weight that is not finite.
int cic_weight_max(const float* w, uint64_t n, float* wmax) {
  int st = g_wmax.reserve(sizeof(uint32_t));
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipMemsetAsync(g_wmax.ptr, 0, sizeof(uint32_t), stream()));
  if (n) {
    const unsigned blocks = grid_for(n, 256) > 2048 ? 2048 : grid_for(n, 256);
    abs_max_bits_kernel<<<blocks, 256, 0, stream()>>>(
      reinterpret_cast<const uint32_t*>(w), n, g_wmax.as<uint32_t>());
    SIF_HIP(hipGetLastError());
  }
  uint32_t bits = 0;
  SIF_HIP(hipMemcpyAsync(&bits, g_wmax.ptr, sizeof(bits), hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  if (bits >= 0x7f800000u) {
    error_set(SIF_ERR_RANGE, "cic: a particle weight is not finite");
    return SIF_ERR_RANGE;
  }
  memcpy(wmax, &bits, sizeof(float));
  return SIF_OK;
}

namespace {
std::atomic<uint32_t> g_cic_tile{8};
}

uint32_t cic_tile_edge() { return g_cic_tile.load(std::memory_order_relaxed); }
int set_cic_tile_edge(uint32_t t) {
  if (t != 8 && t != 16) {
    error_set(SIF_ERR_INVALID, "cic tile edge %u (8 or 16)", t);
    return SIF_ERR_INVALID;
  }
  g_cic_tile.store(t, std::memory_order_relaxed);
  return SIF_OK;
}

template <uint32_t CIC_T>
static int cic_deposit_t(const float* x, const float* y, const float* z, const float* w,
  uint64_t n, uint32_t N, float box, float* values, int64_t* acc);

int cic_deposit(const float* x, const float* y, const float* z, const float* w,
  uint64_t n, uint32_t N, float box, float* values, int64_t* acc) {
  return cic_tile_edge() == 16 ? cic_deposit_t<16>(x, y, z, w, n, N, box, values, acc)
                               : cic_deposit_t<8>(x, y, z, w, n, N, box, values, acc);
}

template <uint32_t CIC_T>
static int cic_deposit_t(const float* x, const float* y, const float* z, const float* w,
  uint64_t n, uint32_t N, float box, float* values, int64_t* acc) {
  const uint64_t total = (uint64_t)N * N * N;
  const float idx = (float)N / box;
  const float cell = box / (float)N;
  const float inv_vol = 1.0f / (float)(cell * cell * cell);
  const uint32_t nt = (N + CIC_T - 1) / CIC_T;
  const uint64_t tiles = (uint64_t)nt * nt * nt;
  // slots and offsets are 64-bit; a tile's count is 32-bit
  if (tiles >= 0xffffffffull) {
    error_set(SIF_ERR_RANGE, "cic: %llu tiles of %u^3 nodes (at most 2^32 - 2)",
      (unsigned long long)tiles, CIC_T);
    return SIF_ERR_RANGE;
  }
  int st = g_tile_counts.reserve((tiles + 1) * 4);
  if (st == SIF_OK) st = g_tile_offsets.reserve((tiles + 1) * 8);
  if (st == SIF_OK) st = g_tile_cursor.reserve(tiles * 4);
  if (st == SIF_OK) st = g_bx.reserve((n ? n : 1) * 4);
  if (st == SIF_OK) st = g_by.reserve((n ? n : 1) * 4);
  if (st == SIF_OK) st = g_bz.reserve((n ? n : 1) * 4);
  if (st == SIF_OK && w) st = g_bw.reserve((n ? n : 1) * 4);
  if (st != SIF_OK)
    return st;
  float wmax = 1.0f;
  if (w && (st = cic_weight_max(w, n, &wmax)) != SIF_OK)
    return st;
  const int s_exp = cic_unit_exponent(n, wmax);
  const double unit_inv = ldexp(1.0, s_exp), unit = ldexp(1.0, -s_exp);
  timing_begin(SIF_AMD_STAGE_CIC);
  SIF_HIP(hipMemsetAsync(acc, 0, total * sizeof(int64_t), stream()));
  if (n) {
    SIF_HIP(hipMemsetAsync(g_tile_counts.ptr, 0, (tiles + 1) * 4, stream()));
    SIF_HIP(hipMemsetAsync(g_tile_cursor.ptr, 0, tiles * 4, stream()));
    const unsigned g = grid_for(n, 256);
    cic_count_kernel<CIC_T><<<g, 256, 0, stream()>>>(x, y, z, n, N, idx, box, nt,
      g_tile_counts.as<uint32_t>());
    SIF_HIP(hipGetLastError());
    st = exclusive_scan_u32(g_tile_counts.as<uint32_t>(), tiles + 1,
      g_tile_offsets.as<uint64_t>(), nullptr);
    if (st != SIF_OK)
      return st;
    cic_bin_kernel<CIC_T><<<g, 256, 0, stream()>>>(x, y, z, w, n, N, idx, box, nt,
      g_tile_offsets.as<uint64_t>(), g_tile_cursor.as<uint32_t>(), g_bx.as<float>(),
      g_by.as<float>(), g_bz.as<float>(), w ? g_bw.as<float>() : nullptr);
    SIF_HIP(hipGetLastError());
    cic_tile_kernel<CIC_T><<<(unsigned)tiles, 256, 0, stream()>>>(g_bx.as<float>(), g_by.as<float>(),
      g_bz.as<float>(), w ? g_bw.as<float>() : nullptr, g_tile_offsets.as<uint64_t>(), N, nt,
      unit_inv, acc);
    SIF_HIP(hipGetLastError());
  }
  cic_finish_kernel<<<grid_for(total, 256), 256, 0, stream()>>>(
    acc, total, unit, inv_vol, values);
  SIF_HIP(hipGetLastError());
  timing_end(SIF_AMD_STAGE_CIC);
  if (timing_active())  // positions (and weights) once, the grid once
    timing_bytes(SIF_AMD_STAGE_CIC, (double)n * (w ? 16.0 : 12.0) + 4.0 * (double)total);
  return SIF_OK;
}

}  // namespace sif_amd

using namespace sif_amd;

extern "C" {

uint64_t sif_amd_cic_work_bytes(uint32_t n_cells) {
  return (uint64_t)n_cells * n_cells * n_cells * sizeof(int64_t);
}

int sif_amd_check_in_box(const float* x, const float* y, const float* z,
  uint64_t n, float box, uint64_t* n_bad, uint64_t* n_edge,
  uint64_t* first_bad) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  return check_in_box(x, y, z, n, box, n_bad, n_edge, first_bad);
}

int sif_amd_grid_assign_cic(const float* x, const float* y, const float* z,
  const float* w, uint64_t n, uint32_t n_cells, float box, float* values,
  void* work) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  if (!x || !y || !z || !values || !work || n_cells < 4 || !(box > 0.0f)) {
    error_set(SIF_ERR_INVALID, "sif_amd_grid_assign_cic: invalid arguments");
    return SIF_ERR_INVALID;
  }
  st = cic_deposit(x, y, z, w, n, n_cells, box, values, (int64_t*)work);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

int sif_amd_grid_mean(const float* values, uint64_t total_cells, double* mean) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  return grid_mean(values, total_cells, mean);
}

int sif_amd_grid_to_density_contrast(float* values, uint64_t total_cells,
  double* mean_out) {
  int st = ensure_init();
  if (st != SIF_OK)
    return st;
  double mean = 0.0;
  timing_begin(SIF_AMD_STAGE_CONTRAST);
  st = grid_mean(values, total_cells, &mean);
  if (st != SIF_OK)
    return st;
  if (mean_out)
    *mean_out = mean;
  if (!(mean > 0.0)) {
    error_set(SIF_ERR_RANGE, "grid: mean density %g (expected > 0)", mean);
    return SIF_ERR_RANGE;
  }
  st = density_contrast(values, total_cells, mean);
  timing_end(SIF_AMD_STAGE_CONTRAST);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

int sif_amd_grid_to_density_contrast_noalloc(float* values, uint64_t total_cells,
  double* mean_out) {
  if (!initialised()) {
    error_set(SIF_ERR_UNINITIALISED,
      "sif_amd_grid_to_density_contrast_noalloc: call sif_system_init first");
    return SIF_ERR_UNINITIALISED;
  }
  if (!values || total_cells == 0) {
    error_set(SIF_ERR_INVALID, "sif_amd_grid_to_density_contrast_noalloc: no grid");
    return SIF_ERR_INVALID;
  }
  const uint64_t blocks = (total_cells + MEAN_BLOCK - 1) / MEAN_BLOCK;
  if (blocks > NOALLOC_MEAN_BLOCKS) {
    error_set(SIF_ERR_RANGE,
      "sif_amd_grid_to_density_contrast_noalloc: %llu cells (at most %llu without "
      "allocation)", (unsigned long long)total_cells,
      (unsigned long long)NOALLOC_MEAN_BLOCKS * MEAN_BLOCK);
    return SIF_ERR_RANGE;
  }
  double* d_scratch = nullptr;
  int st = noalloc_scratch(&d_scratch);
  if (st != SIF_OK)
    return st;
  // every block sum the call reads is written by it first: nothing to clear
  mean_blocks_kernel<<<(unsigned)blocks, 256, 0, stream()>>>(values, total_cells, d_scratch);
  SIF_HIP(hipGetLastError());
  ordered_sum_kernel<<<1, 64, 0, stream()>>>(d_scratch, blocks, d_scratch + NOALLOC_MEAN_BLOCKS);
  SIF_HIP(hipGetLastError());
  double sum = 0.0;
  SIF_HIP(hipMemcpyAsync(&sum, d_scratch + NOALLOC_MEAN_BLOCKS, sizeof(double),
    hipMemcpyDeviceToHost, stream()));
  SIF_HIP(hipStreamSynchronize(stream()));
  const double mean = sum / (double)total_cells;
  if (mean_out)
    *mean_out = mean;
  if (!(mean > 0.0)) {
    error_set(SIF_ERR_RANGE, "grid: mean density %g (expected > 0)", mean);
    return SIF_ERR_RANGE;
  }
  st = density_contrast(values, total_cells, mean);
  if (st != SIF_OK)
    return st;
  SIF_HIP(hipStreamSynchronize(stream()));
  return SIF_OK;
}

}  // extern "C"
